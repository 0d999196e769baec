/*
 * platoanalyze_amd.hpp — C++ host layer over the C ABI (header-only, C++17).
 *
 * Re-creates, for this path, the reference's class names, method names,
 * argument meaning and error behaviour, so that its call sites and unit tests
 * switch by namespace (Plato:: -> Plato::amd::):
 *
 *   Elliptic::VectorFunction<Mechanics<3> | Thermal<3>>
 *       value, gradient_u, gradient_u_T, gradient_z,    src/elliptic/VectorFunction.hpp:115-710
 *       gradient_x
 *   Elliptic::PhysicsScalarFunction<...>
 *       value, gradient_u, gradient_z, gradient_x       src/elliptic/PhysicsScalarFunction.hpp:225-494
 *   Elliptic::Problem<...>
 *       solution, criterionValue, criterionGradient,    src/elliptic/Problem.hpp:237-542
 *       criterionGradientX
 *   LocalVectorFunctionInc<InfinitesimalStrainPlasticity<3>>
 *       updateLocalState, value, gradient_{u,c,cp,z}    src/LocalVectorFunctionInc.hpp:236-874
 *   GlobalVectorFunctionInc<InfinitesimalStrainPlasticity<3>>
 *       value, gradient_{u,up,c,cp,z,n}                 src/GlobalVectorFunctionInc.hpp:919-1155
 *   PlasticityProblem<InfinitesimalStrainPlasticity<3>>
 *       solution, criterionValue, criterionGradient     src/PlasticityProblem.hpp:784-1068
 *   CrsMatrix (rowMap, columnIndices, entries, ...)     src/alg/CrsMatrix.hpp:20-113
 *
 * Inputs are the reference's parameter lists (the Teuchos XML of its input
 * decks and unit tests), read by a small built-in parser
 * (getParametersFromXmlString) with the reference's key names and defaults.
 * Only what this path implements is accepted; any other material model, load
 * type or criterion throws, naming the key.
 *
 * Errors throw std::runtime_error carrying the library's message, as the
 * reference's ANALYZE_THROWERR (src/AnalyzeMacros.hpp:29-33).  Arrays are
 * device arrays shared by reference, like Kokkos views; `toHost` copies one to
 * a std::vector.  Every call runs on the default stream and returns after its
 * work has been enqueued; reading a result to the host synchronizes.
 *
 * Link with -lplatoanalyze_amd (the library that implements the C ABI).
 */
#ifndef PLATOANALYZE_AMD_HPP
#define PLATOANALYZE_AMD_HPP

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "platoanalyze_amd.h"

namespace Plato {
namespace amd {

using Scalar = double;
using OrdinalType = int;

[[noreturn]] inline void throw_error(const std::string& aMsg) { throw std::runtime_error(aMsg); }

inline void check(int aStatus) {
    if (aStatus != PA_OK) throw_error(pa_last_error());
}

// ---------------------------------------------------------------- arrays
// Device array shared by reference (Kokkos::View<T*> in the reference).
template <class T>
class View {
  public:
    View() = default;
    View(const std::string& aLabel, size_t aSize) : mLabel(aLabel), mSize(aSize) {
        void* tPtr = nullptr;
        if (aSize) {
            check(pa_malloc(&tPtr, aSize * sizeof(T)));
            mPtr = std::shared_ptr<T>(static_cast<T*>(tPtr), [](T* p) { pa_free(p); });
            check(pa_memset(tPtr, 0, aSize * sizeof(T), nullptr));
        }
    }
    // a view of memory another object owns, kept alive through aOwner
    static View alias(std::shared_ptr<void> aOwner, T* aPtr, size_t aSize, const std::string& aLabel) {
        View tView;
        tView.mPtr = std::shared_ptr<T>(aOwner, aPtr);
        tView.mSize = aSize;
        tView.mLabel = aLabel;
        return tView;
    }
    T* data() const { return mPtr.get(); }
    size_t extent(int aDim) const { return aDim == 0 ? mSize : 1; }
    size_t size() const { return mSize; }
    const std::string& label() const { return mLabel; }

  private:
    std::string mLabel;
    std::shared_ptr<T> mPtr;
    size_t mSize = 0;
};

using ScalarVector = View<Scalar>;
using OrdinalVector = View<OrdinalType>;

template <class T>
std::vector<T> toHost(const View<T>& aView) {
    std::vector<T> tOut(aView.size());
    if (aView.size()) check(pa_memcpy_d2h(tOut.data(), aView.data(), aView.size() * sizeof(T)));
    return tOut;
}

template <class T>
View<T> toDevice(const std::vector<T>& aHost, const std::string& aLabel = "") {
    View<T> tView(aLabel, aHost.size());
    if (aHost.size()) check(pa_memcpy_h2d(tView.data(), aHost.data(), aHost.size() * sizeof(T)));
    return tView;
}

// Plato::blas1::fill
template <class T>
void fill(T aValue, const View<T>& aView) {
    std::vector<T> tHost(aView.size(), aValue);
    if (aView.size()) check(pa_memcpy_h2d(aView.data(), tHost.data(), tHost.size() * sizeof(T)));
}

// Cell workset (ScalarArray3D in the reference): [cell][row][column], LayoutRight.
struct ScalarArray3D {
    View<Scalar> values;
    size_t dims[3] = {0, 0, 0};
    ScalarArray3D() = default;
    ScalarArray3D(const std::string& aLabel, size_t a, size_t b, size_t c) : values(aLabel, a * b * c), dims{a, b, c} {}
    Scalar* data() const { return values.data(); }
    size_t extent(int aDim) const { return dims[aDim]; }
    size_t size() const { return values.size(); }
};

// Named state arrays of a solve (Plato::Solutions, src/Solutions.hpp).
class Solutions {
  public:
    void set(const std::string& aName, const ScalarVector& aState) { mStates[aName] = {aState}; }
    void set(const std::string& aName, const std::vector<ScalarVector>& aStates) { mStates[aName] = aStates; }
    const ScalarVector& get(const std::string& aName, size_t aStep = 0) const {
        auto tIt = mStates.find(aName);
        if (tIt == mStates.end() || aStep >= tIt->second.size())
            throw_error("Solutions: no state '" + aName + "' at step " + std::to_string(aStep));
        return tIt->second[aStep];
    }
    size_t numSteps(const std::string& aName = "State") const {
        auto tIt = mStates.find(aName);
        return tIt == mStates.end() ? 0 : tIt->second.size();
    }
    bool empty() const { return mStates.empty(); }

  private:
    std::map<std::string, std::vector<ScalarVector>> mStates;
};

// ------------------------------------------------------- parameter lists
// Subset of Teuchos::ParameterList: typed scalar, string and array entries
// and nested sublists, in insertion order.
class ParameterList {
  public:
    struct Entry {
        std::string type;  // "double", "int", "bool", "string", "Array(double)", "Array(int)", "Array(string)", "list"
        std::string text;  // the value as written
        std::shared_ptr<ParameterList> list;
    };

    explicit ParameterList(std::string aName = "ANONYMOUS") : mName(std::move(aName)) {}
    const std::string& name() const { return mName; }

    bool isParameter(const std::string& aName) const {
        const Entry* e = find(aName);
        return e && e->type != "list";
    }
    bool isSublist(const std::string& aName) const {
        const Entry* e = find(aName);
        return e && e->type == "list";
    }
    template <class T>
    bool isType(const std::string& aName) const {
        const Entry* e = find(aName);
        return e && matches<T>(e->type);
    }
    template <class T>
    T get(const std::string& aName) const {
        const Entry* e = find(aName);
        if (!e || e->type == "list")
            throw_error("ParameterList '" + mName + "': parameter '" + aName + "' does not exist");
        if (!matches<T>(e->type))
            throw_error("ParameterList '" + mName + "': parameter '" + aName + "' is not of the type asked for (" +
                        e->type + ")");
        return convert<T>(e->text, aName);
    }
    template <class T>
    T get(const std::string& aName, const T& aDefault) const {
        return isParameter(aName) ? get<T>(aName) : aDefault;
    }
    std::string get(const std::string& aName, const char* aDefault) const {
        return get<std::string>(aName, std::string(aDefault));
    }
    const ParameterList& sublist(const std::string& aName) const {
        const Entry* e = find(aName);
        if (!e || e->type != "list") throw_error("ParameterList '" + mName + "': no sublist '" + aName + "'");
        return *e->list;
    }
    ParameterList& sublist(const std::string& aName) {
        Entry* e = find(aName);
        if (e && e->type != "list") throw_error("ParameterList '" + mName + "': '" + aName + "' is not a sublist");
        if (!e) {
            mEntries.push_back({aName, {"list", "", std::make_shared<ParameterList>(aName)}});
            e = &mEntries.back().second;
        }
        return *e->list;
    }
    void set(const std::string& aName, const std::string& aType, const std::string& aText) {
        Entry* e = find(aName);
        if (e) *e = {aType, aText, nullptr};
        else mEntries.push_back({aName, {aType, aText, nullptr}});
    }
    void set(const std::string& aName, double aValue) {
        std::ostringstream s;
        s.precision(17);
        s << aValue;
        set(aName, "double", s.str());
    }
    void set(const std::string& aName, int aValue) { set(aName, "int", std::to_string(aValue)); }
    void set(const std::string& aName, bool aValue) { set(aName, "bool", aValue ? "true" : "false"); }
    void set(const std::string& aName, const std::string& aValue) { set(aName, "string", aValue); }
    void set(const std::string& aName, const char* aValue) { set(aName, "string", std::string(aValue)); }
    // a deep copy (sublists copied, not shared)
    std::shared_ptr<ParameterList> clone() const {
        auto tOut = std::make_shared<ParameterList>(mName);
        for (const auto& kv : mEntries) {
            Entry e = kv.second;
            if (e.list) e.list = e.list->clone();
            tOut->mEntries.push_back({kv.first, e});
        }
        return tOut;
    }
    // names of the sublists, in order
    std::vector<std::string> sublistNames() const {
        std::vector<std::string> tOut;
        for (const auto& kv : mEntries)
            if (kv.second.type == "list") tOut.push_back(kv.first);
        return tOut;
    }

  private:
    const Entry* find(const std::string& aName) const {
        for (const auto& kv : mEntries)
            if (kv.first == aName) return &kv.second;
        return nullptr;
    }
    Entry* find(const std::string& aName) {
        for (auto& kv : mEntries)
            if (kv.first == aName) return &kv.second;
        return nullptr;
    }
    template <class T>
    static bool matches(const std::string& aType) {
        if constexpr (std::is_same_v<T, double>) return aType == "double" || aType == "int";
        else if constexpr (std::is_same_v<T, int>) return aType == "int";
        else if constexpr (std::is_same_v<T, bool>) return aType == "bool";
        else if constexpr (std::is_same_v<T, std::string>) return aType == "string";
        else if constexpr (std::is_same_v<T, std::vector<double>>)
            return aType == "Array(double)" || aType == "Array(int)";
        else if constexpr (std::is_same_v<T, std::vector<int>>) return aType == "Array(int)";
        else if constexpr (std::is_same_v<T, std::vector<std::string>>) return aType == "Array(string)";
        else return false;
    }
    static std::vector<std::string> array_items(const std::string& aText) {
        std::string t = aText;
        const size_t a = t.find('{'), b = t.rfind('}');
        if (a == std::string::npos || b == std::string::npos || b < a) throw_error("malformed array value '" + aText + "'");
        t = t.substr(a + 1, b - a - 1);
        std::vector<std::string> tOut;
        std::string tItem;
        std::istringstream s(t);
        while (std::getline(s, tItem, ',')) {
            const size_t i = tItem.find_first_not_of(" \t\n\r"), j = tItem.find_last_not_of(" \t\n\r");
            if (i != std::string::npos) tOut.push_back(tItem.substr(i, j - i + 1));
        }
        return tOut;
    }
    static double to_double(const std::string& aText, const std::string& aName) {
        char* tEnd = nullptr;
        const double v = std::strtod(aText.c_str(), &tEnd);
        if (tEnd == aText.c_str()) throw_error("parameter '" + aName + "': '" + aText + "' is not a number");
        return v;
    }
    template <class T>
    static T convert(const std::string& aText, const std::string& aName) {
        if constexpr (std::is_same_v<T, double>) return to_double(aText, aName);
        else if constexpr (std::is_same_v<T, int>) return (int)std::lround(to_double(aText, aName));
        else if constexpr (std::is_same_v<T, bool>) {
            std::string t = aText;
            for (auto& ch : t) ch = (char)std::tolower((unsigned char)ch);
            return t == "true" || t == "1" || t == "yes";
        } else if constexpr (std::is_same_v<T, std::string>) return aText;
        else if constexpr (std::is_same_v<T, std::vector<double>>) {
            std::vector<double> tOut;
            for (const auto& s : array_items(aText)) tOut.push_back(to_double(s, aName));
            return tOut;
        } else if constexpr (std::is_same_v<T, std::vector<int>>) {
            std::vector<int> tOut;
            for (const auto& s : array_items(aText)) tOut.push_back((int)std::lround(to_double(s, aName)));
            return tOut;
        } else {
            return array_items(aText);
        }
    }

    std::string mName;
    std::vector<std::pair<std::string, Entry>> mEntries;
};

namespace detail {

inline std::string xml_unescape(const std::string& s) {
    std::string o;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] != '&') {
            o += s[i];
            continue;
        }
        const size_t j = s.find(';', i);
        const std::string e = j == std::string::npos ? "" : s.substr(i + 1, j - i - 1);
        if (e == "lt") o += '<';
        else if (e == "gt") o += '>';
        else if (e == "amp") o += '&';
        else if (e == "quot") o += '"';
        else if (e == "apos") o += '\'';
        else {
            o += s[i];
            continue;
        }
        i = j;
    }
    return o;
}

// attributes of one tag body: name='value' or name="value"
inline std::map<std::string, std::string> xml_attributes(const std::string& aBody) {
    std::map<std::string, std::string> tOut;
    size_t i = 0;
    while (i < aBody.size()) {
        while (i < aBody.size() && std::isspace((unsigned char)aBody[i])) ++i;
        size_t k = i;
        while (k < aBody.size() && (std::isalnum((unsigned char)aBody[k]) || aBody[k] == '_' || aBody[k] == '-')) ++k;
        if (k == i) break;
        const std::string tKey = aBody.substr(i, k - i);
        while (k < aBody.size() && std::isspace((unsigned char)aBody[k])) ++k;
        if (k >= aBody.size() || aBody[k] != '=') throw_error("XML: attribute '" + tKey + "' has no value");
        ++k;
        while (k < aBody.size() && std::isspace((unsigned char)aBody[k])) ++k;
        if (k >= aBody.size() || (aBody[k] != '\'' && aBody[k] != '"')) throw_error("XML: unquoted attribute");
        const char q = aBody[k];
        const size_t e = aBody.find(q, k + 1);
        if (e == std::string::npos) throw_error("XML: unterminated attribute");
        tOut[tKey] = xml_unescape(aBody.substr(k + 1, e - k - 1));
        i = e + 1;
    }
    return tOut;
}

}  // namespace detail

// Teuchos::getParametersFromXmlString for <ParameterList>/<Parameter> documents.
inline std::shared_ptr<ParameterList> getParametersFromXmlString(const std::string& aXml) {
    std::shared_ptr<ParameterList> tRoot;
    std::vector<ParameterList*> tStack;
    size_t i = 0;
    while ((i = aXml.find('<', i)) != std::string::npos) {
        if (aXml.compare(i, 4, "<!--") == 0) {
            const size_t e = aXml.find("-->", i);
            if (e == std::string::npos) throw_error("XML: unterminated comment");
            i = e + 3;
            continue;
        }
        const size_t e = aXml.find('>', i);
        if (e == std::string::npos) throw_error("XML: unterminated tag");
        std::string tTag = aXml.substr(i + 1, e - i - 1);
        i = e + 1;
        if (!tTag.empty() && tTag[0] == '?') continue;  // <?xml ...?>
        const bool tClosing = !tTag.empty() && tTag[0] == '/';
        bool tSelfClosing = !tTag.empty() && tTag.back() == '/';
        if (tSelfClosing) tTag.pop_back();
        if (tClosing) {
            if (tStack.empty()) throw_error("XML: unbalanced </ParameterList>");
            tStack.pop_back();
            continue;
        }
        size_t k = 0;
        while (k < tTag.size() && !std::isspace((unsigned char)tTag[k])) ++k;
        const std::string tElement = tTag.substr(0, k);
        auto tAttr = detail::xml_attributes(tTag.substr(k));
        if (tElement == "ParameterList") {
            const std::string tName = tAttr.count("name") ? tAttr["name"] : "ANONYMOUS";
            ParameterList* tList;
            if (tStack.empty()) {
                if (tRoot) throw_error("XML: more than one top-level ParameterList");
                tRoot = std::make_shared<ParameterList>(tName);
                tList = tRoot.get();
            } else {
                tList = &tStack.back()->sublist(tName);
            }
            if (!tSelfClosing) tStack.push_back(tList);
        } else if (tElement == "Parameter") {
            if (tStack.empty()) throw_error("XML: Parameter outside a ParameterList");
            if (!tAttr.count("name") || !tAttr.count("type") || !tAttr.count("value"))
                throw_error("XML: Parameter needs name, type and value");
            tStack.back()->set(tAttr["name"], tAttr["type"], tAttr["value"]);
        } else {
            throw_error("XML: unexpected element <" + tElement + ">");
        }
    }
    if (!tRoot) throw_error("XML: no ParameterList");
    if (!tStack.empty()) throw_error("XML: unterminated ParameterList '" + tStack.back()->name() + "'");
    return tRoot;
}

// The input's Function strings, evaluated by the library's reader
// (pa_expression_eval: numbers, x y z t, pi, + - * / ^, parentheses, the
// common functions, leading 'name = expr;'), here as functions of t (the
// reference's Teuchos MathExpr, src/PlatoMathExpr.hpp).
class Expression {
  public:
    explicit Expression(std::string aText) : mText(std::move(aText)) {}
    double operator()(double t) const {
        double v = 0.0;
        if (pa_expression_eval(mText.c_str(), 1, nullptr, nullptr, nullptr, t, &v) != PA_OK)
            throw_error(pa_last_error());
        return v;
    }
    const std::string& text() const { return mText; }

  private:
    std::string mText;
};

// ------------------------------------------------------------------ mesh
// Tet4 mesh with its assembly graph, side sets and node sets (the reference's
// Omega_h mesh behind Plato::SpatialModel).
class Mesh {
  public:
    Mesh(const std::vector<double>& aCoords, const std::vector<int>& aConn) {
        if (aCoords.size() % 3 || aConn.size() % 4) throw_error("Mesh: coords must be N*3 and conn C*4");
        pa_mesh* tMesh = nullptr;
        check(pa_mesh_create(aCoords.data(), (int)(aCoords.size() / 3), aConn.data(), (int)(aConn.size() / 4), &tMesh));
        mMesh = std::shared_ptr<pa_mesh>(tMesh, [](pa_mesh* m) { pa_mesh_destroy(m); });
        mCoords = aCoords;
        mConn = aConn;
    }
    int NumNodes() const { return (int)(mCoords.size() / 3); }
    int NumElements() const { return (int)(mConn.size() / 4); }
    long long NumBlocks() const {
        int n, c;
        int64_t b;
        check(pa_mesh_sizes(mMesh.get(), &n, &c, &b));
        return (long long)b;
    }
    pa_mesh* handle() const { return mMesh.get(); }
    std::shared_ptr<pa_mesh> shared() const { return mMesh; }
    const std::vector<double>& coords() const { return mCoords; }
    const std::vector<int>& connectivity() const { return mConn; }
    void addSideSet(const std::string& aName, const std::vector<int>& aFaceCells, const std::vector<int>& aFaceLocalNodes) {
        if (aFaceLocalNodes.size() != 3 * aFaceCells.size()) throw_error("Mesh: side set needs 3 local nodes per face");
        check(pa_mesh_add_side_set(mMesh.get(), aName.c_str(), (int)aFaceCells.size(), aFaceCells.data(),
                                   aFaceLocalNodes.data()));
    }
    void addNodeSet(const std::string& aName, const std::vector<int>& aNodes) {
        check(pa_mesh_add_node_set(mMesh.get(), aName.c_str(), (int)aNodes.size(), aNodes.data()));
    }
    std::vector<int> getNodeSetNodes(const std::string& aName) const {
        int n = 0;
        check(pa_mesh_node_set(mMesh.get(), aName.c_str(), &n, nullptr));
        std::vector<int> tOut(n);
        check(pa_mesh_node_set(mMesh.get(), aName.c_str(), &n, tOut.data()));
        return tOut;
    }
    // hex counts of a box built by getBoxMesh (0 otherwise): the multigrid needs them
    int boxDims[3] = {0, 0, 0};

  private:
    std::shared_ptr<pa_mesh> mMesh;
    std::vector<double> mCoords;
    std::vector<int> mConn;
};

// Box [0,lx]x[0,ly]x[0,lz] of nx*ny*nz hexes, 6 Tet4 each, with the side sets
// and node sets x-, x+, y-, y+, z-, z+ of the reference's BamG test meshes
// (unit_tests/PlatoTestHelpers.hpp:31-84).
inline std::shared_ptr<Mesh> getBoxMesh(int nx, int ny, int nz, double lx = 1.0, double ly = 1.0, double lz = 1.0) {
    if (nx < 1 || ny < 1 || nz < 1) throw_error("getBoxMesh: interval counts must be positive");
    const long N = (long)(nx + 1) * (ny + 1) * (nz + 1), C = 6L * nx * ny * nz;
    std::vector<double> tCoords(N * 3);
    std::vector<int> tConn(C * 4);
    check(pa_box_mesh(nx, ny, nz, lx, ly, lz, tCoords.data(), tConn.data()));
    auto tMesh = std::make_shared<Mesh>(tCoords, tConn);
    tMesh->boxDims[0] = nx;
    tMesh->boxDims[1] = ny;
    tMesh->boxDims[2] = nz;
    const double L[3] = {lx, ly, lz};
    const char* tAxes = "xyz";
    static const int kFaces[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    for (int ax = 0; ax < 3; ++ax)
        for (int side = 0; side < 2; ++side) {
            const double v = side ? L[ax] : 0.0;
            auto on = [&](int node) { return std::fabs(tCoords[(size_t)node * 3 + ax] - v) < 1e-12; };
            std::vector<int> tCells, tLocal, tNodes;
            for (long c = 0; c < C; ++c)
                for (const auto& f : kFaces)
                    if (on(tConn[c * 4 + f[0]]) && on(tConn[c * 4 + f[1]]) && on(tConn[c * 4 + f[2]])) {
                        tCells.push_back((int)c);
                        tLocal.insert(tLocal.end(), f, f + 3);
                    }
            for (long n = 0; n < N; ++n)
                if (on((int)n)) tNodes.push_back((int)n);
            const std::string tName = std::string(1, tAxes[ax]) + (side ? "+" : "-");
            tMesh->addSideSet(tName, tCells, tLocal);
            tMesh->addNodeSet(tName, tNodes);
        }
    return tMesh;
}

namespace detail {

// Non-record variables of a netCDF classic (CDF-1) or 64-bit-offset (CDF-2)
// file, the container of Exodus II meshes (the netCDF classic format
// specification: big-endian header of dimensions, attributes and variables,
// each variable's data contiguous at its `begin` offset).
class NetcdfClassic {
  public:
    struct Var {
        std::vector<int> dims;
        int type = 0;  // 1 byte, 2 char, 3 short, 4 int, 5 float, 6 double
        unsigned long long begin = 0;
        std::map<std::string, std::string> text_attrs;
    };
    explicit NetcdfClassic(const std::string& aPath) : mPath(aPath) {
        std::FILE* f = std::fopen(aPath.c_str(), "rb");
        if (!f) throw_error("netCDF: cannot open '" + aPath + "'");
        std::fseek(f, 0, SEEK_END);
        mBytes.resize((size_t)std::ftell(f));
        std::fseek(f, 0, SEEK_SET);
        const size_t n = std::fread(mBytes.data(), 1, mBytes.size(), f);
        std::fclose(f);
        if (n != mBytes.size() || mBytes.size() < 8 || mBytes[0] != 'C' || mBytes[1] != 'D' || mBytes[2] != 'F')
            throw_error("netCDF: '" + aPath + "' is not a netCDF classic file (netCDF-4/HDF5 files are not read here)");
        const int version = mBytes[3];
        if (version != 1 && version != 2) throw_error("netCDF: '" + aPath + "': unsupported format version");
        mOffset64 = version == 2;
        size_t p = 4;
        u32(p);  // numrecs
        // dimensions
        if (u32(p) == 0x0A) {
            const unsigned nd = u32(p);
            for (unsigned i = 0; i < nd; ++i) {
                const std::string name = text(p);
                mDimNames.push_back(name);
                mDims.push_back(u32(p));
            }
        } else {
            u32(p);
        }
        skip_attrs(p, nullptr);
        // variables
        if (u32(p) == 0x0B) {
            const unsigned nv = u32(p);
            for (unsigned i = 0; i < nv; ++i) {
                const std::string name = text(p);
                Var v;
                const unsigned nd = u32(p);
                for (unsigned k = 0; k < nd; ++k) {
                    const unsigned d = u32(p);
                    if (d >= mDims.size()) throw_error("netCDF: '" + aPath + "': variable '" + name + "' has an unknown dimension");
                    v.dims.push_back((int)d);
                }
                skip_attrs(p, &v.text_attrs);
                v.type = (int)u32(p);
                u32(p);  // vsize
                v.begin = mOffset64 ? u64(p) : u32(p);
                mVars[name] = v;
            }
        } else {
            u32(p);
        }
    }
    long long dim(const std::string& aName, long long aDefault = -1) const {
        for (size_t i = 0; i < mDimNames.size(); ++i)
            if (mDimNames[i] == aName) return mDims[i];
        if (aDefault < 0) throw_error("netCDF: '" + mPath + "' has no dimension '" + aName + "'");
        return aDefault;
    }
    bool has(const std::string& aName) const { return mVars.count(aName) != 0; }
    const Var& var(const std::string& aName) const {
        auto it = mVars.find(aName);
        if (it == mVars.end()) throw_error("netCDF: '" + mPath + "' has no variable '" + aName + "'");
        return it->second;
    }
    std::vector<long long> shape(const std::string& aName) const {
        std::vector<long long> s;
        for (int d : var(aName).dims) {
            if (mDims[d] == 0) throw_error("netCDF: record variable '" + aName + "' is not supported");
            s.push_back(mDims[d]);
        }
        return s;
    }
    // Bytes of a variable of n values of width w at v.begin, checked against the file.
    void check_extent(const std::string& aName, const Var& v, unsigned long long n, unsigned long long w) const {
        if (n > mBytes.size() || v.begin > mBytes.size() || n * w > mBytes.size() - v.begin)
            throw_error("netCDF: variable '" + aName + "' of '" + mPath + "' runs past the end of the file");
    }
    // numeric variable as doubles (int, short, byte, float or double)
    std::vector<double> numbers(const std::string& aName) const {
        const Var& v = var(aName);
        unsigned long long n = 1;
        for (long long s : shape(aName)) {
            n *= (unsigned long long)s;
            if (n > mBytes.size()) throw_error("netCDF: variable '" + aName + "' runs past the end of the file");
        }
        const int w = v.type == 6 ? 8 : (v.type == 4 || v.type == 5) ? 4 : v.type == 3 ? 2 : 1;
        if (v.type == 2) throw_error("netCDF: variable '" + aName + "' is text");
        if (v.type < 1 || v.type > 6) throw_error("netCDF: variable '" + aName + "' has an unknown type");
        check_extent(aName, v, n, (unsigned long long)w);
        std::vector<double> out(n);
        size_t p = (size_t)v.begin;
        for (size_t i = 0; i < n; ++i) {
            if (v.type == 6) {
                const unsigned long long b = u64(p);
                double d;
                std::memcpy(&d, &b, 8);
                out[i] = d;
            } else if (v.type == 5) {
                const unsigned b = u32(p);
                float f;
                std::memcpy(&f, &b, 4);
                out[i] = f;
            } else if (v.type == 4) {
                out[i] = (double)(int)u32(p);
            } else if (v.type == 3) {
                const int x = (mBytes[p] << 8) | mBytes[p + 1];
                p += 2;
                out[i] = (double)(short)x;
            } else {
                out[i] = (double)(signed char)mBytes[p++];
            }
        }
        return out;
    }
    // rows of a char variable (rows x len), each cut at its first NUL
    std::vector<std::string> strings(const std::string& aName) const {
        const Var& v = var(aName);
        const auto s = shape(aName);
        if (v.type != 2 || s.size() != 2) throw_error("netCDF: '" + aName + "' is not a text array");
        if ((unsigned long long)s[0] > mBytes.size() || (unsigned long long)s[1] > mBytes.size())
            throw_error("netCDF: variable '" + aName + "' runs past the end of the file");
        check_extent(aName, v, (unsigned long long)s[0] * (unsigned long long)s[1], 1);
        std::vector<std::string> out;
        for (long long r = 0; r < s[0]; ++r) {
            const char* c = reinterpret_cast<const char*>(mBytes.data() + v.begin + (size_t)(r * s[1]));
            std::string t(c, (size_t)s[1]);
            t = t.substr(0, t.find('\0'));
            while (!t.empty() && std::isspace((unsigned char)t.back())) t.pop_back();
            out.push_back(t);
        }
        return out;
    }

  private:
    unsigned u32(size_t& p) const {
        if (p + 4 > mBytes.size()) throw_error("netCDF: '" + mPath + "' truncated header");
        const unsigned v = (unsigned)mBytes[p] << 24 | (unsigned)mBytes[p + 1] << 16 | (unsigned)mBytes[p + 2] << 8 |
                           (unsigned)mBytes[p + 3];
        p += 4;
        return v;
    }
    unsigned long long u64(size_t& p) const {
        const unsigned long long hi = u32(p);
        return hi << 32 | u32(p);
    }
    std::string text(size_t& p) const {
        const unsigned n = u32(p);
        if (p + n > mBytes.size()) throw_error("netCDF: '" + mPath + "' truncated name");
        std::string s(reinterpret_cast<const char*>(mBytes.data() + p), n);
        p += (n + 3) / 4 * 4;
        return s;
    }
    void skip_attrs(size_t& p, std::map<std::string, std::string>* aText) const {
        if (u32(p) != 0x0C) {
            u32(p);
            return;
        }
        const unsigned na = u32(p);
        for (unsigned i = 0; i < na; ++i) {
            const std::string name = text(p);
            const unsigned type = u32(p), n = u32(p);
            const unsigned w = type == 6 ? 8 : (type == 4 || type == 5) ? 4 : type == 3 ? 2 : 1;
            const unsigned long long bytes = ((unsigned long long)n * w + 3) / 4 * 4;
            if (bytes > mBytes.size() - p) throw_error("netCDF: '" + mPath + "' truncated attribute");
            if (type == 2 && aText) {
                std::string v(reinterpret_cast<const char*>(mBytes.data() + p), n);
                (*aText)[name] = v.substr(0, v.find('\0'));
            }
            p += ((size_t)n * w + 3) / 4 * 4;
        }
    }
    std::string mPath;
    std::vector<unsigned char> mBytes;
    bool mOffset64 = false;
    std::vector<std::string> mDimNames;
    std::vector<long long> mDims;
    std::map<std::string, Var> mVars;
};

}  // namespace detail

// Tet4 mesh of an Exodus II file (netCDF classic or 64-bit offset) with its
// side sets and node sets, as the reference reads through Omega_h
// (src/OmegaHMesh.cpp:13-27): coordx/y/z (or coord), connect<b> of TETRA4
// blocks (1-based, blocks in order), elem_ss<i>/side_ss<i> with the TETRA
// side numbering 1 = (1,2,4), 2 = (2,3,4), 3 = (1,4,3), 4 = (1,3,2), node_ns<i>,
// and the names in ss_names / ns_names (the id from ss_prop1 / ns_prop1 when a
// name is empty).
// Host arrays of an Exodus II Tet4 mesh (no device work): coordinates,
// 0-based connectivity, side sets (cells and local node ordinals per face),
// node sets.  Every index is checked against the file's own counts.
struct ExodusData {
    std::vector<double> coords;
    std::vector<int> conn;
    std::vector<std::pair<std::string, std::pair<std::vector<int>, std::vector<int>>>> side_sets;
    std::vector<std::pair<std::string, std::vector<int>>> node_sets;
};

// 1-based index x of a file as a 0-based one in [0, n), else -1.
inline long long exodus_index(double x, long long n) {
    if (!(x >= 1.0 && x <= (double)n)) return -1;
    return (long long)x - 1;
}

inline ExodusData readExodusData(const std::string& aPath) {
    detail::NetcdfClassic nc(aPath);
    ExodusData tOut;
    // set and block counts are bounded by the variables they need
    auto count = [&](const char* aDim) {
        const long long c = nc.dim(aDim, 0);
        if (c < 0 || c > 1000000) throw_error(std::string("readExodusMesh: bad ") + aDim);
        return c;
    };
    if (nc.dim("num_dim", 3) != 3) throw_error("readExodusMesh: only three-dimensional meshes are on this path");
    const long long N = nc.dim("num_nodes");
    if (N <= 0 || N > (1LL << 31) - 1) throw_error("readExodusMesh: bad node count");
    std::vector<double>& tCoords = tOut.coords;
    tCoords.assign((size_t)N * 3, 0.0);
    if (nc.has("coord")) {
        const auto c = nc.numbers("coord");
        if ((long long)c.size() < 3 * N) throw_error("readExodusMesh: coord holds fewer than 3 x num_nodes values");
        for (long long n = 0; n < N; ++n)
            for (int d = 0; d < 3; ++d) tCoords[(size_t)n * 3 + d] = c[(size_t)d * N + n];
    } else {
        const char* names[3] = {"coordx", "coordy", "coordz"};
        for (int d = 0; d < 3; ++d) {
            const auto c = nc.numbers(names[d]);
            if ((long long)c.size() < N) throw_error(std::string("readExodusMesh: ") + names[d] + " is short");
            for (long long n = 0; n < N; ++n) tCoords[(size_t)n * 3 + d] = c[(size_t)n];
        }
    }
    std::vector<int>& tConn = tOut.conn;
    const long long nblk = count("num_el_blk");
    for (long long b = 1; b <= nblk; ++b) {
        const std::string v = "connect" + std::to_string(b);
        if (!nc.has(v)) continue;
        std::string etype = nc.var(v).text_attrs.count("elem_type") ? nc.var(v).text_attrs.at("elem_type") : "TETRA4";
        for (auto& ch : etype) ch = (char)std::toupper((unsigned char)ch);
        const auto s = nc.shape(v);
        if ((etype != "TETRA" && etype != "TETRA4" && etype != "TET4" && etype != "TET") || s.size() != 2 || s[1] != 4)
            throw_error("readExodusMesh: element block " + std::to_string(b) + " is " + etype +
                        "; only TETRA4 is on this path");
        for (double x : nc.numbers(v)) {
            const long long n = exodus_index(x, N);
            if (n < 0) throw_error("readExodusMesh: connectivity out of range");
            tConn.push_back((int)n);
        }
    }
    const long long C = (long long)tConn.size() / 4;
    auto set_names = [&](const std::string& aNames, const std::string& aIds, long long aCount) {
        std::vector<std::string> out((size_t)aCount);
        const auto names = nc.has(aNames) ? nc.strings(aNames) : std::vector<std::string>();
        const auto ids = nc.has(aIds) ? nc.numbers(aIds) : std::vector<double>();
        for (long long i = 0; i < aCount; ++i) {
            out[i] = i < (long long)names.size() ? names[i] : "";
            if (out[i].empty())
                out[i] = std::to_string(i < (long long)ids.size() && std::isfinite(ids[i]) && std::fabs(ids[i]) < 1e15
                                            ? (long long)ids[i]
                                            : i + 1);
        }
        return out;
    };
    static const int kSides[4][3] = {{0, 1, 3}, {1, 2, 3}, {0, 3, 2}, {0, 2, 1}};
    const long long nss = count("num_side_sets");
    const auto ss = set_names("ss_names", "ss_prop1", nss);
    for (long long i = 1; i <= nss; ++i) {
        const std::string ev = "elem_ss" + std::to_string(i), sv = "side_ss" + std::to_string(i);
        if (!nc.has(ev)) continue;
        const auto el = nc.numbers(ev), sd = nc.numbers(sv);
        if (el.size() != sd.size())
            throw_error("readExodusMesh: side set '" + ss[i - 1] + "' has " + std::to_string(el.size()) +
                        " elements and " + std::to_string(sd.size()) + " sides");
        std::vector<int> cells, local;
        for (size_t k = 0; k < el.size(); ++k) {
            const long long c = exodus_index(el[k], C), s = exodus_index(sd[k], 4);
            if (c < 0 || s < 0) throw_error("readExodusMesh: side set '" + ss[i - 1] + "' out of range");
            cells.push_back((int)c);
            local.insert(local.end(), kSides[s], kSides[s] + 3);
        }
        tOut.side_sets.push_back({ss[i - 1], {cells, local}});
    }
    const long long nns = count("num_node_sets");
    const auto ns = set_names("ns_names", "ns_prop1", nns);
    for (long long i = 1; i <= nns; ++i) {
        const std::string nv = "node_ns" + std::to_string(i);
        if (!nc.has(nv)) continue;
        std::vector<int> nodes;
        for (double x : nc.numbers(nv)) {
            const long long n = exodus_index(x, N);
            if (n < 0) throw_error("readExodusMesh: node set '" + ns[i - 1] + "' out of range");
            nodes.push_back((int)n);
        }
        tOut.node_sets.push_back({ns[i - 1], nodes});
    }
    return tOut;
}

// The mesh of an Exodus II Tet4 file with its side sets and node sets.
inline std::shared_ptr<Mesh> readExodusMesh(const std::string& aPath) {
    const ExodusData tData = readExodusData(aPath);
    auto tMesh = std::make_shared<Mesh>(tData.coords, tData.conn);
    for (const auto& s : tData.side_sets) tMesh->addSideSet(s.first, s.second.first, s.second.second);
    for (const auto& s : tData.node_sets) tMesh->addNodeSet(s.first, s.second);
    return tMesh;
}

// PlatoUtestHelpers::getBoxMesh("TET4", meshWidth): the unit cube with
// meshWidth intervals per direction.
inline std::shared_ptr<Mesh> getBoxMesh(const std::string& aType, int aMeshWidth) {
    if (aType != "TET4") throw_error("getBoxMesh: element type '" + aType + "' is not supported (TET4 only)");
    return getBoxMesh(aMeshWidth, aMeshWidth, aMeshWidth, 1.0, 1.0, 1.0);
}

// ------------------------------------------------------------ CrsMatrix
// Block CRS matrix on the mesh graph (src/alg/CrsMatrix.hpp): row = node,
// sorted column node ids with the diagonal; block entry e*bi*bj + i*bj + j.
template <class T = Scalar>
class CrsMatrix {
  public:
    CrsMatrix(const Mesh& aMesh, int aNumRowsPerBlock, int aNumColsPerBlock, const std::string& aLabel)
        : mNumRowsPerBlock(aNumRowsPerBlock), mNumColsPerBlock(aNumColsPerBlock),
          mIsBlockMatrix(aNumRowsPerBlock * aNumColsPerBlock > 1) {
        const int* r = nullptr;
        const int* c = nullptr;
        check(pa_mesh_graph_dev(aMesh.handle(), &r, &c));
        const long long nnzb = aMesh.NumBlocks();
        auto tOwner = std::static_pointer_cast<void>(aMesh.shared());
        mRowMap = OrdinalVector::alias(tOwner, const_cast<int*>(r), (size_t)aMesh.NumNodes() + 1, "row map");
        mColumnIndices = OrdinalVector::alias(tOwner, const_cast<int*>(c), (size_t)nnzb, "column indices");
        mEntries = View<T>(aLabel, (size_t)nnzb * aNumRowsPerBlock * aNumColsPerBlock);
        mNumRows = aMesh.NumNodes() * aNumRowsPerBlock;
        mNumCols = aMesh.NumNodes() * aNumColsPerBlock;
    }
    const OrdinalVector& rowMap() const { return mRowMap; }
    const OrdinalVector& columnIndices() const { return mColumnIndices; }
    const View<T>& entries() const { return mEntries; }
    int numRows() const { return mNumRows; }
    int numCols() const { return mNumCols; }
    int numRowsPerBlock() const { return mNumRowsPerBlock; }
    int numColsPerBlock() const { return mNumColsPerBlock; }
    bool isBlockMatrix() const { return mIsBlockMatrix; }

  private:
    OrdinalVector mRowMap, mColumnIndices;
    View<T> mEntries;
    int mNumRows = 0, mNumCols = 0, mNumRowsPerBlock = 1, mNumColsPerBlock = 1;
    bool mIsBlockMatrix = false;
};

using CrsMatrixType = CrsMatrix<Scalar>;

// --------------------------------------------------------------- physics
template <int SpaceDim>
struct Mechanics {
    static_assert(SpaceDim == 3, "this path is three-dimensional");
    static constexpr int mNumSpatialDims = SpaceDim;
    static constexpr int mNumDofsPerNode = SpaceDim;
    static constexpr int mPhysics = PA_MECHANICS;
    static constexpr int mEnergy = PA_INTERNAL_ELASTIC_ENERGY;
    static constexpr const char* mEnergyName = "Internal Elastic Energy";
    // the residual's dof names (src/elliptic/ElastostaticResidual.hpp:104-106)
    static std::vector<std::string> dofNames() { return {"displacement X", "displacement Y", "displacement Z"}; }
};

template <int SpaceDim>
struct Thermal {
    static_assert(SpaceDim == 3, "this path is three-dimensional");
    static constexpr int mNumSpatialDims = SpaceDim;
    static constexpr int mNumDofsPerNode = 1;
    static constexpr int mPhysics = PA_THERMAL;
    static constexpr int mEnergy = PA_INTERNAL_THERMAL_ENERGY;
    static constexpr const char* mEnergyName = "Internal Thermal Energy";
    // src/elliptic/ThermostaticResidual.hpp:91
    static std::vector<std::string> dofNames() { return {"temperature"}; }
};

// Physics "Plasticity": the stabilized mixed u-p J2 form (4 dofs per node).
template <int SpaceDim>
struct InfinitesimalStrainPlasticity {
    static_assert(SpaceDim == 3, "this path is three-dimensional");
    static constexpr int mNumSpatialDims = SpaceDim;
    static constexpr int mNumDofsPerNode = SpaceDim + 1;
    static constexpr int mNumLocalDofsPerCell = 14;
    static constexpr int mFormulation = PA_J2_MIXED;
    // the stabilized residual's dof names (src/StabilizedElastostaticResidual.hpp:86-89)
    static std::vector<std::string> dofNames() {
        return {"displacement X", "displacement Y", "displacement Z", "pressure"};
    }
};

// Pure-displacement J2 (3 dofs per node); no counterpart in the reference.
template <int SpaceDim>
struct DisplacementPlasticity {
    static_assert(SpaceDim == 3, "this path is three-dimensional");
    static constexpr int mNumSpatialDims = SpaceDim;
    static constexpr int mNumDofsPerNode = SpaceDim;
    static constexpr int mNumLocalDofsPerCell = 14;
    static constexpr int mFormulation = PA_J2_DISPLACEMENT;
    static std::vector<std::string> dofNames() { return {"displacement X", "displacement Y", "displacement Z"}; }
};

namespace detail {

// Material model sublist of the (single) domain: Spatial Model/Domains/<d>/
// Material Model names an entry of Material Models (src/SpatialModel.hpp).
inline const ParameterList& domain_material(const ParameterList& aParams) {
    if (!aParams.isSublist("Material Models")) throw_error("input: no 'Material Models' sublist");
    const auto& tModels = aParams.sublist("Material Models");
    std::string tName;
    if (aParams.isSublist("Spatial Model") && aParams.sublist("Spatial Model").isSublist("Domains")) {
        const auto& tDomains = aParams.sublist("Spatial Model").sublist("Domains");
        const auto tNames = tDomains.sublistNames();
        if (tNames.size() != 1) throw_error("input: exactly one domain is supported on this path");
        tName = tDomains.sublist(tNames[0]).get<std::string>("Material Model");
    } else {
        const auto tNames = tModels.sublistNames();
        if (tNames.size() != 1) throw_error("input: name the material model through 'Spatial Model/Domains'");
        tName = tNames[0];
    }
    if (!tModels.isSublist(tName)) throw_error("input: material model '" + tName + "' is not defined");
    return tModels.sublist(tName);
}

// SIMP penalty of a 'Penalty Function' sublist (src/Simp.hpp:41-52).
inline void simp(const ParameterList& aList, double aExponentDefault, double aMinDefault, double aOut[2]) {
    const std::string tType = aList.get<std::string>("Type", std::string("SIMP"));
    if (tType != "SIMP") throw_error("Penalty Function type '" + tType + "' is not supported (SIMP only)");
    aOut[0] = aList.get<double>("Exponent", aExponentDefault);
    aOut[1] = aList.get<double>("Minimum Value", aMinDefault);
}

inline double constant_function(const std::string& aText, const std::string& aWhat) {
    Expression e(aText);
    const double v0 = e(0.0), v1 = e(1.0);
    if (v0 != v1) throw_error(aWhat + ": function '" + aText + "' is not constant");
    return v0;
}

// A Function string that is a plain number (no names or operators beyond a sign).
inline bool plain_number(const std::string& aText, double& aValue) {
    const char* b = aText.c_str();
    char* e = nullptr;
    aValue = std::strtod(b, &e);
    if (e == b) return false;
    while (*e && std::isspace((unsigned char)*e)) ++e;
    return *e == 0;
}

// Dirichlet dofs and end values of 'Essential Boundary Conditions'
// (src/EssentialBCs.hpp:81-99, src/EssentialBC.hpp:31-65), sorted, the later
// entry winning on a repeated dof.  allowTime: 'Time Dependent' functions
// linear in t with f(0) = 0 give their t = 1 value (the load ramp scales it).
inline void essential_bcs(const ParameterList& aParams, const Mesh& aMesh, int aDofsPerNode, bool aAllowTime,
                          std::vector<int>& aDofs, std::vector<double>& aValues) {
    std::map<int, double> tBcs;
    if (aParams.isSublist("Essential Boundary Conditions")) {
        const auto& tList = aParams.sublist("Essential Boundary Conditions");
        for (const auto& tName : tList.sublistNames()) {
            const auto& bc = tList.sublist(tName);
            const std::string tType = bc.get<std::string>("Type");
            const int tIndex = bc.get<int>("Index", 0);
            if (tIndex < 0 || tIndex >= aDofsPerNode) throw_error("essential BC '" + tName + "': Index out of range");
            double tValue = 0.0;
            if (tType == "Zero Value") {
                tValue = 0.0;
            } else if (tType == "Fixed Value") {
                // a Function is taken at time 0, as the reference's elliptic
                // solve does (src/EssentialBC.hpp:95-105)
                tValue = bc.isType<std::string>("Function") ? Expression(bc.get<std::string>("Function"))(0.0)
                                                            : bc.get<double>("Value");
                if (aAllowTime && tValue != 0.0)
                    throw_error("essential BC '" + tName + "': a constant non-zero value is not supported by the "
                                "pseudo-time driver (use 'Time Dependent' with a function of t)");
            } else if (tType == "Time Dependent" && aAllowTime) {
                Expression e(bc.get<std::string>("Function"));
                const double f0 = e(0.0), f1 = e(1.0), fh = e(0.5);
                if (f0 != 0.0 || std::fabs(fh - 0.5 * f1) > 1e-12 * std::fabs(f1))
                    throw_error("essential BC '" + tName + "': only functions a*t are supported");
                tValue = f1;
            } else {
                throw_error("essential BC '" + tName + "': type '" + tType + "' is not supported on this path");
            }
            for (int n : aMesh.getNodeSetNodes(bc.get<std::string>("Sides")))
                tBcs[n * aDofsPerNode + tIndex] = tValue;
        }
    }
    aDofs.clear();
    aValues.clear();
    for (const auto& kv : tBcs) {
        aDofs.push_back(kv.first);
        aValues.push_back(kv.second);
    }
}

}  // namespace detail

namespace Elliptic {

// ----------------------------------------------------- VectorFunction
// Elliptic::VectorFunction (src/elliptic/VectorFunction.hpp:115-710) with the
// residual of Mechanics (ElastostaticResidual) or Thermal
// (ThermostaticResidual), SIMP from '<type>/Penalty Function', uniform
// natural BCs and constant body loads.
template <class PhysicsT>
class VectorFunction {
  public:
    static constexpr int mNumDofsPerNode = PhysicsT::mNumDofsPerNode;

    VectorFunction(std::shared_ptr<Mesh> aMesh, const ParameterList& aProblemParams,
                   const std::string& aType = "Elliptic")
        : mMesh(std::move(aMesh)) {
        if (aType != "Elliptic") throw_error("VectorFunction: PDE type '" + aType + "' is not supported (Elliptic)");
        const auto& tMaterial = detail::domain_material(aProblemParams);
        double tMat[2] = {0.0, 0.0}, tPen[2];
        if (PhysicsT::mPhysics == PA_MECHANICS) {
            if (!tMaterial.isSublist("Isotropic Linear Elastic"))
                throw_error("VectorFunction: the material needs an 'Isotropic Linear Elastic' sublist");
            const auto& e = tMaterial.sublist("Isotropic Linear Elastic");
            tMat[0] = e.get<double>("Youngs Modulus");
            tMat[1] = e.get<double>("Poissons Ratio");
        } else {
            if (!tMaterial.isSublist("Thermal Conduction"))
                throw_error("VectorFunction: the material needs a 'Thermal Conduction' sublist");
            tMat[0] = tMaterial.sublist("Thermal Conduction").get<double>("Thermal Conductivity");
        }
        const ParameterList tEmpty;
        const auto& tRes = aProblemParams.isSublist(aType) ? aProblemParams.sublist(aType) : tEmpty;
        detail::simp(tRes.isSublist("Penalty Function") ? tRes.sublist("Penalty Function") : tEmpty, 3.0, 1e-9, tPen);
        pa_vector_function* tVf = nullptr;
        check(pa_vector_function_create(mMesh->handle(), PhysicsT::mPhysics, tMat, tPen, &tVf));
        mVf = std::shared_ptr<pa_vector_function>(tVf, [](pa_vector_function* v) { pa_vector_function_destroy(v); });
        if (aProblemParams.isSublist("Natural Boundary Conditions")) readNaturalBCs(aProblemParams.sublist("Natural Boundary Conditions"));
        if (aProblemParams.isSublist("Body Loads")) readBodyLoads(aProblemParams.sublist("Body Loads"));
    }

    ScalarVector value(const ScalarVector& aState, const ScalarVector& aControl, Scalar /*aTimeStep*/ = 0.0) const {
        checkSizes(aState, aControl);
        ScalarVector tResidual("Residual", size());
        check(pa_vector_function_value(mVf.get(), aState.data(), aControl.data(), tResidual.data(), nullptr));
        return tResidual;
    }
    std::shared_ptr<CrsMatrixType> gradient_u(const ScalarVector& aState, const ScalarVector& aControl,
                                              Scalar /*aTimeStep*/ = 0.0) const {
        checkSizes(aState, aControl);
        auto tJ = std::make_shared<CrsMatrixType>(*mMesh, mNumDofsPerNode, mNumDofsPerNode, "Jacobian");
        check(pa_vector_function_gradient_u(mVf.get(), aState.data(), aControl.data(), tJ->entries().data(), PA_FP64,
                                            nullptr));
        return tJ;
    }
    // gradient_u with FP32 values (the mixed-precision solve's matrix); no reference counterpart
    std::shared_ptr<CrsMatrix<float>> gradient_u_fp32(const ScalarVector& aState, const ScalarVector& aControl) const {
        checkSizes(aState, aControl);
        auto tJ = std::make_shared<CrsMatrix<float>>(*mMesh, mNumDofsPerNode, mNumDofsPerNode, "Jacobian");
        check(pa_vector_function_gradient_u(mVf.get(), aState.data(), aControl.data(), tJ->entries().data(), PA_FP32,
                                            nullptr));
        return tJ;
    }
    std::shared_ptr<CrsMatrixType> gradient_u_T(const ScalarVector& aState, const ScalarVector& aControl,
                                                Scalar /*aTimeStep*/ = 0.0) const {
        checkSizes(aState, aControl);
        auto tJ = std::make_shared<CrsMatrixType>(*mMesh, mNumDofsPerNode, mNumDofsPerNode, "Jacobian Transpose");
        check(pa_vector_function_gradient_u_T(mVf.get(), aState.data(), aControl.data(), tJ->entries().data(),
                                              PA_FP64, nullptr));
        return tJ;
    }
    // (dR/dz)^T on the node graph with 1 x dpn blocks
    std::shared_ptr<CrsMatrixType> gradient_z(const ScalarVector& aState, const ScalarVector& aControl,
                                              Scalar /*aTimeStep*/ = 0.0) const {
        checkSizes(aState, aControl);
        auto tJ = std::make_shared<CrsMatrixType>(*mMesh, 1, mNumDofsPerNode, "Gradient Z");
        check(pa_vector_function_gradient_z(mVf.get(), aState.data(), aControl.data(), tJ->entries().data(), nullptr));
        return tJ;
    }
    // (dR/dx)^T on the node graph with 3 x dpn blocks (src/elliptic/VectorFunction.hpp:337-420)
    std::shared_ptr<CrsMatrixType> gradient_x(const ScalarVector& aState, const ScalarVector& aControl,
                                              Scalar /*aTimeStep*/ = 0.0) const {
        checkSizes(aState, aControl);
        auto tJ = std::make_shared<CrsMatrixType>(*mMesh, 3, mNumDofsPerNode, "Gradient X");
        check(pa_vector_function_gradient_x(mVf.get(), aState.data(), aControl.data(), tJ->entries().data(), nullptr));
        return tJ;
    }
    OrdinalType size() const { return mMesh->NumNodes() * mNumDofsPerNode; }
    OrdinalType numNodes() const { return mMesh->NumNodes(); }
    OrdinalType numCells() const { return mMesh->NumElements(); }
    OrdinalType numDofsPerNode() const { return mNumDofsPerNode; }
    OrdinalType numNodesPerCell() const { return 4; }
    OrdinalType numDofsPerCell() const { return 4 * mNumDofsPerNode; }
    OrdinalType numControlsPerNode() const { return 1; }
    std::vector<std::string> getDofNames() const { return PhysicsT::dofNames(); }
    pa_vector_function* handle() const { return mVf.get(); }
    const std::shared_ptr<Mesh>& mesh() const { return mMesh; }

  private:
    void checkSizes(const ScalarVector& aState, const ScalarVector& aControl) const {
        if ((OrdinalType)aState.size() != size() || (OrdinalType)aControl.size() != numNodes())
            throw_error("VectorFunction: state must hold N*dofs and control N values");
    }
    // NaturalBCs::appendNaturalBC (src/NaturalBCs.hpp:110-224): 'Uniform' with
    // 'Values' (one per dof) or 'Value' at 'Index'; 'Uniform Component' with
    // 'Value' at 'Component' x|y|z.  Values may be constant expressions.
    void readNaturalBCs(const ParameterList& aList) {
        for (const auto& tName : aList.sublistNames()) {
            const auto& bc = aList.sublist(tName);
            const std::string tType = bc.get<std::string>("Type");
            std::vector<double> tVec(mNumDofsPerNode, 0.0);
            auto scalar = [&](const std::string& aKey) {
                return bc.isType<std::string>(aKey) ? detail::constant_function(bc.get<std::string>(aKey), tName)
                                                    : bc.get<double>(aKey);
            };
            if (tType == "Uniform") {
                const bool tHasValues = bc.isParameter("Values"), tHasValue = bc.isParameter("Value");
                if (tHasValues == tHasValue)
                    throw_error("natural BC '" + tName + "': give exactly one of 'Values' and 'Value'");
                if (tHasValues) {
                    std::vector<double> v;
                    if (bc.isType<std::vector<std::string>>("Values")) {
                        for (const auto& s : bc.get<std::vector<std::string>>("Values"))
                            v.push_back(detail::constant_function(s, tName));
                    } else {
                        v = bc.get<std::vector<double>>("Values");
                    }
                    if ((int)v.size() != mNumDofsPerNode) throw_error("natural BC '" + tName + "': wrong 'Values' length");
                    tVec = v;
                } else {
                    const int tDof = bc.get<int>("Index", 0);
                    if (tDof < 0 || tDof >= mNumDofsPerNode) throw_error("natural BC '" + tName + "': Index out of range");
                    tVec[tDof] = scalar("Value");
                }
            } else if (tType == "Uniform Component") {
                const std::string c = bc.get<std::string>("Component");
                const int tDof = c == "x" || c == "X" ? 0 : c == "y" || c == "Y" ? 1 : c == "z" || c == "Z" ? 2 : -1;
                if (tDof < 0 || tDof >= mNumDofsPerNode) throw_error("natural BC '" + tName + "': bad Component");
                tVec[tDof] = scalar("Value");
            } else {
                throw_error("Natural Boundary Condition Type '" + tType + "' is not supported on this path");
            }
            check(pa_vector_function_add_uniform_natural_bc(mVf.get(), bc.get<std::string>("Sides").c_str(),
                                                           tVec.data()));
        }
    }
    // BodyLoads (src/BodyLoads.hpp:40-122): constant 'Function' at 'Index'.
    // 'Body Loads' (src/BodyLoads.hpp:38-122): a constant thermal source takes
    // the fused path; any other Function of x, y, z (or any elastic load) is
    // evaluated at the cell centroids by the library.
    void readBodyLoads(const ParameterList& aList) {
        for (const auto& tName : aList.sublistNames()) {
            const auto& bl = aList.sublist(tName);
            const std::string tFunction = bl.get<std::string>("Function");
            const int tIndex = bl.get<int>("Index", 0);
            double v = 0.0;
            if (mNumDofsPerNode == 1 && tIndex == 0 && detail::plain_number(tFunction, v))
                check(pa_vector_function_add_body_load(mVf.get(), tIndex, v));
            else
                check(pa_vector_function_add_body_load_function(mVf.get(), tIndex, tFunction.c_str()));
        }
    }

    std::shared_ptr<Mesh> mMesh;
    std::shared_ptr<pa_vector_function> mVf;
};

// ------------------------------------------------ PhysicsScalarFunction
// Elliptic::PhysicsScalarFunction with 'Internal Elastic Energy' or
// 'Internal Thermal Energy' (src/elliptic/PhysicsScalarFunction.hpp:60-494),
// SIMP from 'Criteria/<name>/Penalty Function'.
template <class PhysicsT>
class PhysicsScalarFunction {
  public:
    PhysicsScalarFunction(std::shared_ptr<Mesh> aMesh, const ParameterList& aProblemParams,
                          const std::string& aFunctionName)
        : mMesh(std::move(aMesh)), mName(aFunctionName) {
        if (!aProblemParams.isSublist("Criteria") || !aProblemParams.sublist("Criteria").isSublist(aFunctionName))
            throw_error("PhysicsScalarFunction: no criterion '" + aFunctionName + "'");
        const auto& tCrit = aProblemParams.sublist("Criteria").sublist(aFunctionName);
        const std::string tType = tCrit.get<std::string>("Scalar Function Type", std::string(""));
        if (tType != PhysicsT::mEnergyName)
            throw_error("PhysicsScalarFunction: scalar function type '" + tType + "' is not supported for this physics ('" +
                        PhysicsT::mEnergyName + "')");
        const auto& tMaterial = detail::domain_material(aProblemParams);
        double tMat[2] = {0.0, 0.0}, tPen[2];
        if (PhysicsT::mPhysics == PA_MECHANICS) {
            const auto& e = tMaterial.sublist("Isotropic Linear Elastic");
            tMat[0] = e.get<double>("Youngs Modulus");
            tMat[1] = e.get<double>("Poissons Ratio");
        } else {
            tMat[0] = tMaterial.sublist("Thermal Conduction").get<double>("Thermal Conductivity");
        }
        const ParameterList tEmpty;
        detail::simp(tCrit.isSublist("Penalty Function") ? tCrit.sublist("Penalty Function") : tEmpty, 3.0, 1e-9, tPen);
        pa_scalar_function* tSf = nullptr;
        check(pa_scalar_function_create(mMesh->handle(), PhysicsT::mEnergy, tMat, tPen, &tSf));
        mSf = std::shared_ptr<pa_scalar_function>(tSf, [](pa_scalar_function* s) { pa_scalar_function_destroy(s); });
    }

    Scalar value(const ScalarVector& aState, const ScalarVector& aControl, Scalar /*aTimeStep*/ = 0.0) const {
        ScalarVector tValue("value", 1);
        check(pa_scalar_function_value(mSf.get(), aState.data(), aControl.data(), tValue.data(), nullptr));
        return toHost(tValue)[0];
    }
    Scalar value(const Solutions& aSolution, const ScalarVector& aControl, Scalar aTimeStep = 0.0) const {
        return value(aSolution.get("State"), aControl, aTimeStep);
    }
    ScalarVector gradient_u(const ScalarVector& aState, const ScalarVector& aControl, Scalar = 0.0) const {
        ScalarVector tOut("gradient_u", aState.size());
        check(pa_scalar_function_gradient_u(mSf.get(), aState.data(), aControl.data(), tOut.data(), nullptr));
        return tOut;
    }
    ScalarVector gradient_u(const Solutions& aSolution, const ScalarVector& aControl, OrdinalType aStepIndex = 0,
                            Scalar aTimeStep = 0.0) const {
        return gradient_u(aSolution.get("State", aStepIndex), aControl, aTimeStep);
    }
    ScalarVector gradient_z(const ScalarVector& aState, const ScalarVector& aControl, Scalar = 0.0) const {
        ScalarVector tOut("gradient_z", aControl.size());
        check(pa_scalar_function_gradient_z(mSf.get(), aState.data(), aControl.data(), tOut.data(), nullptr));
        return tOut;
    }
    ScalarVector gradient_z(const Solutions& aSolution, const ScalarVector& aControl, Scalar aTimeStep = 0.0) const {
        return gradient_z(aSolution.get("State"), aControl, aTimeStep);
    }
    // src/elliptic/PhysicsScalarFunction.hpp:293-355: N*3 entries, node*3 + dim
    ScalarVector gradient_x(const ScalarVector& aState, const ScalarVector& aControl, Scalar = 0.0) const {
        ScalarVector tOut("gradient_x", 3 * aControl.size());
        check(pa_scalar_function_gradient_x(mSf.get(), aState.data(), aControl.data(), tOut.data(), nullptr));
        return tOut;
    }
    ScalarVector gradient_x(const Solutions& aSolution, const ScalarVector& aControl, Scalar aTimeStep = 0.0) const {
        return gradient_x(aSolution.get("State"), aControl, aTimeStep);
    }
    std::string name() const { return mName; }
    pa_scalar_function* handle() const { return mSf.get(); }

  private:
    std::shared_ptr<Mesh> mMesh;
    std::string mName;
    std::shared_ptr<pa_scalar_function> mSf;
};

// -------------------------------------------------------------- Problem
// Elliptic::Problem (src/elliptic/Problem.hpp:96-450): one Newton step per
// solution (R(0, z), K, symmetric Dirichlet elimination, K du = -R), the
// criteria of 'Criteria' and their self-adjoint gradients ('Self-Adjoint'
// must be true: the general adjoint of this path is lambda = -u).
// 'Linear Solver' keys: Tolerance (default 1e-14, as AztecOO's), Iterations
// (1000); and, with no reference counterpart, Preconditioner ('Block Jacobi',
// 'Jacobi', 'None', 'Multigrid') and Precision ('Double' | 'Mixed').
template <class PhysicsT>
class Problem {
  public:
    Problem(std::shared_ptr<Mesh> aMesh, const ParameterList& aProblemParams)
        : mMesh(aMesh), mPDE(std::make_shared<VectorFunction<PhysicsT>>(
                            aMesh, aProblemParams, aProblemParams.get<std::string>("PDE Constraint"))) {
        if (!aProblemParams.get<bool>("Self-Adjoint", false))
            throw_error("Problem: only self-adjoint criteria ('Self-Adjoint' = true) are on this path");
        detail::essential_bcs(aProblemParams, *aMesh, PhysicsT::mNumDofsPerNode, false, mBcDofs, mBcValues);
        const ParameterList tEmpty;
        const auto& ls = aProblemParams.isSublist("Linear Solver") ? aProblemParams.sublist("Linear Solver") : tEmpty;
        mOptions.mode = ls.get<std::string>("Precision", std::string("Double")) == "Mixed" ? PA_FP32 : PA_FP64;
        mOptions.tol = ls.get<double>("Tolerance", 1e-14);
        mOptions.max_iter = ls.get<int>("Iterations", 1000);
        mOptions.replace_every = ls.get<int>("Replace Every", 200);
        mOptions.check_every = ls.get<int>("Check Every", 10);
        const std::string p = ls.get<std::string>("Preconditioner", std::string("Block Jacobi"));
        mMultigrid = p == "Multigrid";
        mOptions.precond = p == "None" ? PA_PRECOND_NONE : p == "Jacobi" ? PA_PRECOND_JACOBI : PA_PRECOND_BLOCK_JACOBI;
        if (aProblemParams.isSublist("Criteria")) {
            const auto& cl = aProblemParams.sublist("Criteria");
            for (const auto& tName : cl.sublistNames()) {
                if (cl.sublist(tName).get<std::string>("Type", std::string("Scalar Function")) != "Scalar Function")
                    throw_error("Problem: criterion '" + tName + "' is not a 'Scalar Function'");
                auto tSf = std::make_shared<PhysicsScalarFunction<PhysicsT>>(aMesh, aProblemParams, tName);
                pa_problem* tP = nullptr;
                check(pa_problem_create(mPDE->handle(), tSf->handle(), (int)mBcDofs.size(), mBcDofs.data(),
                                        mBcValues.data(), &mOptions, &tP));
                auto tProblem = std::shared_ptr<pa_problem>(tP, [](pa_problem* q) { pa_problem_destroy(q); });
                if (mMultigrid) {
                    if (!mMesh->boxDims[0]) throw_error("Problem: the multigrid preconditioner needs a box mesh");
                    pa_multigrid_options o{};
                    o.pre_sweeps = o.post_sweeps = 2;
                    o.omega = 0.6;
                    check(pa_problem_set_multigrid(tP, mMesh->boxDims[0], mMesh->boxDims[1], mMesh->boxDims[2], &o));
                }
                mCriteria[tName] = {tSf, tProblem};
                if (mFirst.empty()) mFirst = tName;
            }
        }
        if (mFirst.empty()) throw_error("Problem: no criteria");
    }

    Solutions solution(const ScalarVector& aControl) {
        ScalarVector tState("State", (size_t)mPDE->size());
        check(pa_problem_solution(mCriteria.at(mFirst).problem.get(), aControl.data(), tState.data(), &mStats, nullptr));
        Solutions tS;
        tS.set("State", tState);
        mLast = tS;
        return tS;
    }
    Scalar criterionValue(const ScalarVector& aControl, const Solutions& aSolution, const std::string& aName) {
        ScalarVector tValue("value", 1);
        check(pa_problem_criterion_value(criterion(aName).problem.get(), aControl.data(), aSolution.get("State").data(),
                                         tValue.data(), nullptr));
        return toHost(tValue)[0];
    }
    Scalar criterionValue(const ScalarVector& aControl, const std::string& aName) {
        return criterionValue(aControl, lastSolution(aControl), aName);
    }
    ScalarVector criterionGradient(const ScalarVector& aControl, const Solutions& aSolution, const std::string& aName) {
        ScalarVector tGrad("criterion gradient", aControl.size());
        check(pa_problem_criterion_gradient(criterion(aName).problem.get(), aControl.data(),
                                            aSolution.get("State").data(), tGrad.data(), nullptr));
        return tGrad;
    }
    ScalarVector criterionGradient(const ScalarVector& aControl, const std::string& aName) {
        return criterionGradient(aControl, lastSolution(aControl), aName);
    }
    // self-adjoint criterionGradientX: -df/dx (src/elliptic/Problem.hpp:492-515)
    ScalarVector criterionGradientX(const ScalarVector& aControl, const Solutions& aSolution,
                                    const std::string& aName) {
        ScalarVector tGrad("criterion gradient x", 3 * aControl.size());
        check(pa_problem_criterion_gradient_x(criterion(aName).problem.get(), aControl.data(),
                                              aSolution.get("State").data(), tGrad.data(), nullptr));
        return tGrad;
    }
    ScalarVector criterionGradientX(const ScalarVector& aControl, const std::string& aName) {
        return criterionGradientX(aControl, lastSolution(aControl), aName);
    }
    const pa_solve_stats& solverStats() const { return mStats; }
    const std::vector<int>& bcDofs() const { return mBcDofs; }
    const VectorFunction<PhysicsT>& getPDE() const { return *mPDE; }

  private:
    struct Criterion {
        std::shared_ptr<PhysicsScalarFunction<PhysicsT>> function;
        std::shared_ptr<pa_problem> problem;
    };
    const Criterion& criterion(const std::string& aName) const {
        auto it = mCriteria.find(aName);
        if (it == mCriteria.end()) throw_error("Problem: no criterion '" + aName + "'");
        return it->second;
    }
    const Solutions& lastSolution(const ScalarVector& aControl) {
        if (mLast.empty()) solution(aControl);
        return mLast;
    }

    std::shared_ptr<Mesh> mMesh;
    std::shared_ptr<VectorFunction<PhysicsT>> mPDE;
    std::map<std::string, Criterion> mCriteria;
    std::string mFirst;
    std::vector<int> mBcDofs;
    std::vector<double> mBcValues;
    pa_problem_options mOptions{};
    bool mMultigrid = false;
    pa_solve_stats mStats{};
    Solutions mLast;
};

}  // namespace Elliptic

// ----------------------------------------------------- J2 plasticity
namespace detail {

// pa_j2_params from the reference's input lists: Material Models/<m>/
// {Isotropic Linear Elastic, Plasticity Model/J2 Plasticity}
// (src/J2PlasticityLocalResidual.hpp:148-220), Material Models/Pressure
// Scaling, Elliptic/Penalty Function (src/InfinitesimalStrainPlasticityResidual.hpp:142-160),
// State Gradient Projection/Penalty Function (src/Projection.hpp:30-62) and the
// plastic work criterion's Penalty Function (src/PlasticWorkCriterion.hpp:295-303).
inline pa_j2_params j2_params(const ParameterList& aParams, int aFormulation, const std::string& aWorkCriterion) {
    pa_j2_params p;
    pa_j2_params_default(&p);
    p.formulation = aFormulation;
    const auto& tMaterial = domain_material(aParams);
    if (!tMaterial.isSublist("Isotropic Linear Elastic"))
        throw_error("plasticity: the material needs an 'Isotropic Linear Elastic' sublist");
    const auto& e = tMaterial.sublist("Isotropic Linear Elastic");
    p.youngs = e.get<double>("Youngs Modulus");
    p.poisson = e.get<double>("Poissons Ratio");
    if (!tMaterial.isSublist("Plasticity Model") || !tMaterial.sublist("Plasticity Model").isSublist("J2 Plasticity"))
        throw_error("plasticity: the material needs a 'Plasticity Model/J2 Plasticity' sublist");
    const auto& j = tMaterial.sublist("Plasticity Model").sublist("J2 Plasticity");
    p.hardening_isotropic = j.get<double>("Hardening Modulus Isotropic");
    p.hardening_kinematic = j.get<double>("Hardening Modulus Kinematic");
    p.initial_yield_stress = j.get<double>("Initial Yield Stress");
    p.elastic_penalty_exponent = j.get<double>("Elastic Properties Penalty Exponent", 1.0);
    p.elastic_min_ersatz = j.get<double>("Elastic Properties Minimum Ersatz", 1e-9);
    p.plastic_penalty_exponent = j.get<double>("Plastic Properties Penalty Exponent", 0.5);
    p.plastic_min_ersatz = j.get<double>("Plastic Properties Minimum Ersatz", 1e-4);
    p.pressure_scaling = aParams.sublist("Material Models").get<double>("Pressure Scaling", 1.0);
    const ParameterList tEmpty;
    double tPen[2];
    const auto& el = aParams.isSublist("Elliptic") ? aParams.sublist("Elliptic") : tEmpty;
    simp(el.isSublist("Penalty Function") ? el.sublist("Penalty Function") : tEmpty, 1.0, 1e-9, tPen);
    p.global_penalty_exponent = tPen[0];
    p.global_min_ersatz = tPen[1];
    if (aParams.isSublist("State Gradient Projection") &&
        aParams.sublist("State Gradient Projection").isSublist("Penalty Function")) {
        const auto& pp = aParams.sublist("State Gradient Projection").sublist("Penalty Function");
        const std::string t = pp.get<std::string>("Type", std::string("NoPenalty"));
        if (t == "SIMP") {
            p.projection_penalty = 1;
            p.projection_penalty_exponent = pp.get<double>("Exponent", 3.0);
            p.projection_min_ersatz = pp.get<double>("Minimum Value", 1e-9);
        } else if (t != "NoPenalty") {
            throw_error("State Gradient Projection penalty '" + t + "' is not supported (SIMP or NoPenalty)");
        }
    }
    if (!aWorkCriterion.empty()) {
        const auto& c = aParams.sublist("Criteria").sublist(aWorkCriterion);
        simp(c.isSublist("Penalty Function") ? c.sublist("Penalty Function") : tEmpty, 3.0, 1e-9, tPen);
        p.work_penalty_exponent = tPen[0];
        p.work_min_ersatz = tPen[1];
    }
    return p;
}

// the 'Plastic Work' criterion of the input ("" if none); more than one throws
inline std::string plastic_work_criterion(const ParameterList& aParams, double& aMultiplier) {
    std::string tFound;
    aMultiplier = 1.0;
    if (!aParams.isSublist("Criteria")) return tFound;
    const auto& cl = aParams.sublist("Criteria");
    for (const auto& n : cl.sublistNames()) {
        const std::string t = cl.sublist(n).get<std::string>("Scalar Function Type", std::string(""));
        if (t != "Plastic Work") throw_error("plasticity: criterion '" + n + "' of type '" + t + "' is not on this path");
        if (!tFound.empty()) throw_error("plasticity: more than one 'Plastic Work' criterion");
        tFound = n;
        aMultiplier = cl.sublist(n).get<double>("Multiplier", 1.0);  // src/BasicLocalScalarFunction.hpp:714
    }
    return tFound;
}

}  // namespace detail

// Shared J2 evaluator (the kernels behind the local and global functions).
class PlasticityEvaluator {
  public:
    PlasticityEvaluator(std::shared_ptr<Mesh> aMesh, const pa_j2_params& aParams) : mMesh(std::move(aMesh)) {
        pa_plasticity* p = nullptr;
        check(pa_plasticity_create(mMesh->handle(), &aParams, &p));
        mP = std::shared_ptr<pa_plasticity>(p, [](pa_plasticity* q) { pa_plasticity_destroy(q); });
        mDofsPerNode = aParams.formulation == PA_J2_MIXED ? 4 : 3;
    }
    pa_plasticity* handle() const { return mP.get(); }
    const std::shared_ptr<Mesh>& mesh() const { return mMesh; }
    int dofsPerNode() const { return mDofsPerNode; }

  private:
    std::shared_ptr<Mesh> mMesh;
    std::shared_ptr<pa_plasticity> mP;
    int mDofsPerNode = 4;
};

// LocalVectorFunctionInc (src/LocalVectorFunctionInc.hpp:236-874) for the J2
// local residual: 14 states per cell.  The previous global state and the
// time data of the reference's signatures are accepted and not needed here.
template <class PhysicsT>
class LocalVectorFunctionInc {
  public:
    static constexpr int mNumLocalDofsPerCell = 14;
    LocalVectorFunctionInc(std::shared_ptr<Mesh> aMesh, const ParameterList& aParams)
        : mEval(std::make_shared<PlasticityEvaluator>(aMesh, detail::j2_params(aParams, PhysicsT::mFormulation, ""))) {}
    explicit LocalVectorFunctionInc(std::shared_ptr<PlasticityEvaluator> aEval) : mEval(std::move(aEval)) {}

    void updateLocalState(const ScalarVector& aGlobalState, const ScalarVector& /*aPrevGlobalState*/,
                          const ScalarVector& aLocalState, const ScalarVector& aPrevLocalState,
                          const ScalarVector& aControl, Scalar /*aTimeData*/ = 0.0) const {
        check(pa_j2_update_local_state(mEval->handle(), aGlobalState.data(), aPrevLocalState.data(), aControl.data(),
                                       aLocalState.data(), nullptr));
    }
    ScalarVector value(const ScalarVector& aGlobalState, const ScalarVector& /*aPrevGlobalState*/,
                       const ScalarVector& aLocalState, const ScalarVector& aPrevLocalState,
                       const ScalarVector& aControl, Scalar /*aTimeData*/ = 0.0) const {
        ScalarVector tH("Local Residual", (size_t)numCells() * 14);
        check(pa_j2_local_value(mEval->handle(), aGlobalState.data(), aLocalState.data(), aPrevLocalState.data(),
                                aControl.data(), tH.data(), nullptr));
        return tH;
    }
    ScalarArray3D gradient_u(const ScalarVector& u, const ScalarVector& up, const ScalarVector& c,
                             const ScalarVector& cp, const ScalarVector& z, Scalar = 0.0) const {
        return workset(PA_WRT_U, 16, u, c, cp, z);
    }
    ScalarArray3D gradient_c(const ScalarVector& u, const ScalarVector&, const ScalarVector& c, const ScalarVector& cp,
                             const ScalarVector& z, Scalar = 0.0) const {
        return workset(PA_WRT_C, 14, u, c, cp, z);
    }
    ScalarArray3D gradient_cp(const ScalarVector& u, const ScalarVector&, const ScalarVector& c,
                              const ScalarVector& cp, const ScalarVector& z, Scalar = 0.0) const {
        return workset(PA_WRT_C_PREV, 14, u, c, cp, z);
    }
    ScalarArray3D gradient_z(const ScalarVector& u, const ScalarVector&, const ScalarVector& c, const ScalarVector& cp,
                             const ScalarVector& z, Scalar = 0.0) const {
        return workset(PA_WRT_Z, 4, u, c, cp, z);
    }
    OrdinalType numCells() const { return mEval->mesh()->NumElements(); }

  private:
    ScalarArray3D workset(int aWrt, int aWidth, const ScalarVector& u, const ScalarVector& c, const ScalarVector& cp,
                          const ScalarVector& z) const {
        ScalarArray3D tOut("Local Jacobian", (size_t)numCells(), 14, (size_t)aWidth);
        check(pa_j2_local_gradient(mEval->handle(), aWrt, u.data(), c.data(), cp.data(), z.data(), tOut.data(), nullptr));
        return tOut;
    }
    std::shared_ptr<PlasticityEvaluator> mEval;
};

// GlobalVectorFunctionInc (src/GlobalVectorFunctionInc.hpp:919-1155) for the
// J2 global residual.  Worksets are [cell][16][W] with the local dof node*4 + d.
// The residual does not depend on the previous global or local state, so
// gradient_up and gradient_cp are zero worksets.
template <class PhysicsT>
class GlobalVectorFunctionInc {
  public:
    GlobalVectorFunctionInc(std::shared_ptr<Mesh> aMesh, const ParameterList& aParams)
        : mEval(std::make_shared<PlasticityEvaluator>(aMesh, detail::j2_params(aParams, PhysicsT::mFormulation, ""))) {}
    explicit GlobalVectorFunctionInc(std::shared_ptr<PlasticityEvaluator> aEval) : mEval(std::move(aEval)) {}

    ScalarVector value(const ScalarVector& u, const ScalarVector& /*up*/, const ScalarVector& c,
                       const ScalarVector& /*cp*/, const ScalarVector& aProjPressGrad, const ScalarVector& z,
                       Scalar = 0.0) const {
        ScalarVector tR("Global Residual", (size_t)size());
        check(pa_j2_global_value(mEval->handle(), u.data(), c.data(), aProjPressGrad.data(), z.data(), tR.data(),
                                 nullptr));
        return tR;
    }
    ScalarArray3D gradient_u(const ScalarVector& u, const ScalarVector&, const ScalarVector& c, const ScalarVector&,
                             const ScalarVector& pi, const ScalarVector& z, Scalar = 0.0) const {
        return workset(PA_WRT_U, 16, u, c, pi, z);
    }
    ScalarArray3D gradient_up(const ScalarVector&, const ScalarVector&, const ScalarVector&, const ScalarVector&,
                              const ScalarVector&, const ScalarVector&, Scalar = 0.0) const {
        return ScalarArray3D("Global Jacobian", (size_t)numCells(), 16, 16);
    }
    ScalarArray3D gradient_c(const ScalarVector& u, const ScalarVector&, const ScalarVector& c, const ScalarVector&,
                             const ScalarVector& pi, const ScalarVector& z, Scalar = 0.0) const {
        return workset(PA_WRT_C, 14, u, c, pi, z);
    }
    ScalarArray3D gradient_cp(const ScalarVector&, const ScalarVector&, const ScalarVector&, const ScalarVector&,
                              const ScalarVector&, const ScalarVector&, Scalar = 0.0) const {
        return ScalarArray3D("Global Jacobian", (size_t)numCells(), 16, 14);
    }
    ScalarArray3D gradient_z(const ScalarVector& u, const ScalarVector&, const ScalarVector& c, const ScalarVector&,
                             const ScalarVector& pi, const ScalarVector& z, Scalar = 0.0) const {
        return workset(PA_WRT_Z, 4, u, c, pi, z);
    }
    // d R / d (projected pressure gradient), [cell][16][12]
    ScalarArray3D gradient_n(const ScalarVector& u, const ScalarVector&, const ScalarVector& c, const ScalarVector&,
                             const ScalarVector& pi, const ScalarVector& z, Scalar = 0.0) const {
        return workset(PA_WRT_PI, 12, u, c, pi, z);
    }
    // NewtonRaphsonSolver::assembleTangentMatrix + assembleResidual
    // (src/NewtonRaphsonSolver.hpp:204-330) in one call: K_t on the node graph
    // (dpn x dpn blocks) and R_t; transposed: K_t^T only (the adjoint's operator).
    std::shared_ptr<CrsMatrixType> condensedSystem(const ScalarVector& u, const ScalarVector& c, const ScalarVector& cp,
                                                   const ScalarVector& pi, const ScalarVector& z,
                                                   ScalarVector* aResidual, bool aTranspose = false) const {
        const int D = mEval->dofsPerNode();
        auto tK = std::make_shared<CrsMatrixType>(*mEval->mesh(), D, D, "Tangent");
        if (!aTranspose && !aResidual) throw_error("condensedSystem: a residual output is required");
        if (aResidual && !aTranspose) *aResidual = ScalarVector("Condensed Residual", (size_t)size());
        check(pa_j2_condensed_system(mEval->handle(), u.data(), c.data(), cp.data(), pi.data(), z.data(),
                                     tK->entries().data(), aTranspose ? nullptr : aResidual->data(),
                                     aTranspose ? 1 : 0, nullptr));
        return tK;
    }
    OrdinalType size() const { return mEval->mesh()->NumNodes() * mEval->dofsPerNode(); }
    OrdinalType numCells() const { return mEval->mesh()->NumElements(); }
    OrdinalType numDofsPerNode() const { return mEval->dofsPerNode(); }

  private:
    ScalarArray3D workset(int aWrt, int aWidth, const ScalarVector& u, const ScalarVector& c, const ScalarVector& pi,
                          const ScalarVector& z) const {
        ScalarArray3D tOut("Global Jacobian", (size_t)numCells(), 16, (size_t)aWidth);
        check(pa_j2_global_gradient(mEval->handle(), aWrt, u.data(), c.data(), pi.data(), z.data(), tOut.data(),
                                    nullptr));
        return tOut;
    }
    std::shared_ptr<PlasticityEvaluator> mEval;
};

// PlasticityProblem (src/PlasticityProblem.hpp:96-1068): 'Time Stepping/
// Initial Num. Pseudo Time Steps' (default 20) with load factor (k+1)/K,
// 'Newton-Raphson/{Maximum Number Iterations (10), Stopping Tolerance (1e-8),
// Current Residual Norm Stopping Tolerance (1e-8)}', the 'Plastic Work'
// criterion with its 'Multiplier', Dirichlet values from 'Essential Boundary
// Conditions' (Zero Value, Time Dependent a*t).  'Linear Solver' keys, with
// no reference counterpart: Solver ('GMRES' | 'BiCGStab'), Restart (50),
// Tolerance (1e-12), Iterations (20000), Preconditioner ('Block Jacobi' |
// 'Multigrid'; the multigrid needs a box mesh), Smoother Sweeps and Smoother
// Damping of the multigrid (5 and 0.45 for the mixed form, 3 and 0.7 for the
// displacement form).
template <class PhysicsT>
class PlasticityProblem {
  public:
    PlasticityProblem(std::shared_ptr<Mesh> aMesh, const ParameterList& aParams) : mMesh(aMesh) {
        const std::string tPhysics = aParams.get<std::string>("Physics", std::string("Plasticity"));
        if (tPhysics != "Plasticity") throw_error("PlasticityProblem: Physics '" + tPhysics + "' is not on this path");
        mCriterion = detail::plastic_work_criterion(aParams, mMultiplier);
        mEval = std::make_shared<PlasticityEvaluator>(aMesh,
                                                      detail::j2_params(aParams, PhysicsT::mFormulation, mCriterion));
        pa_plasticity_options_default(&mOptions);
        if (aParams.isSublist("Time Stepping"))
            mOptions.num_steps = aParams.sublist("Time Stepping").get<int>("Initial Num. Pseudo Time Steps", 20);
        else
            mOptions.num_steps = 20;
        if (aParams.isSublist("Newton-Raphson")) {
            const auto& nr = aParams.sublist("Newton-Raphson");
            mOptions.max_newton = nr.get<int>("Maximum Number Iterations", 10);
            mOptions.newton_stop_tol = nr.get<double>("Stopping Tolerance", 1e-8);
            mOptions.newton_current_norm_tol = nr.get<double>("Current Residual Norm Stopping Tolerance", 1e-8);
        }
        if (aParams.isSublist("Linear Solver")) {
            const auto& ls = aParams.sublist("Linear Solver");
            mOptions.krylov = ls.get<std::string>("Solver", std::string("GMRES")) == "BiCGStab" ? PA_KRYLOV_BICGSTAB
                                                                                              : PA_KRYLOV_GMRES;
            mOptions.restart = ls.get<int>("Restart", 50);
            mOptions.linear_tol = ls.get<double>("Tolerance", 1e-12);
            mOptions.linear_max_iter = ls.get<int>("Iterations", 20000);
            mMultigrid = ls.get<std::string>("Preconditioner", std::string("Block Jacobi")) == "Multigrid";
            mSmootherSweeps = ls.get<int>("Smoother Sweeps", 0);
            mSmootherDamping = ls.get<double>("Smoother Damping", 0.0);
        }
        readEssentialBoundaryConditions(aParams);
    }

    // PlasticityProblem::readEssentialBoundaryConditions
    void readEssentialBoundaryConditions(const ParameterList& aParams) {
        detail::essential_bcs(aParams, *mMesh, mEval->dofsPerNode(), true, mBcDofs, mBcValues);
        pa_plasticity_problem* q = nullptr;
        check(pa_plasticity_problem_create(mEval->handle(), (int)mBcDofs.size(), mBcDofs.data(), mBcValues.data(),
                                           &mOptions, &q));
        mProblem = std::shared_ptr<pa_plasticity_problem>(q, [](pa_plasticity_problem* r) {
            pa_plasticity_problem_destroy(r);
        });
        if (mMultigrid) {
            if (!mMesh->boxDims[0]) throw_error("PlasticityProblem: the multigrid preconditioner needs a box mesh");
            pa_multigrid_options o{};
            // the mixed form's saddle-point blocks take a smaller damping
            o.pre_sweeps = o.post_sweeps = mSmootherSweeps > 0 ? mSmootherSweeps : (PhysicsT::mNumDofsPerNode == 4 ? 5 : 3);
            o.omega = mSmootherDamping > 0.0 ? mSmootherDamping : (PhysicsT::mNumDofsPerNode == 4 ? 0.45 : 0.7);
            check(pa_plasticity_problem_set_multigrid(mProblem.get(), mMesh->boxDims[0], mMesh->boxDims[1],
                                                      mMesh->boxDims[2], &o));
        }
    }

    // Every step's global state ("State"), local state ("Local State") and
    // projected pressure gradient ("Projected Pressure Gradient"), as device
    // copies.
    Solutions solution(const ScalarVector& aControl) {
        int tConverged = 0;
        check(pa_plasticity_problem_solution(mProblem.get(), aControl.data(), &tConverged, nullptr));
        mConverged = tConverged != 0;
        const int K = mOptions.num_steps;
        std::vector<ScalarVector> us, cs, pis;
        const size_t nu = (size_t)mMesh->NumNodes() * mEval->dofsPerNode(), nc = (size_t)mMesh->NumElements() * 14,
                     np = (size_t)mMesh->NumNodes() * 3;
        for (int k = 0; k < K; ++k) {
            const double *u, *c, *pi;
            check(pa_plasticity_problem_state(mProblem.get(), k, &u, &c, &pi));
            us.emplace_back("State", nu);
            cs.emplace_back("Local State", nc);
            pis.emplace_back("Projected Pressure Gradient", np);
            check(pa_memcpy_d2d(us.back().data(), u, nu * sizeof(double), nullptr));
            check(pa_memcpy_d2d(cs.back().data(), c, nc * sizeof(double), nullptr));
            check(pa_memcpy_d2d(pis.back().data(), pi, np * sizeof(double), nullptr));
        }
        Solutions tS;
        tS.set("State", us);
        tS.set("Local State", cs);
        tS.set("Projected Pressure Gradient", pis);
        return tS;
    }
    Scalar criterionValue(const ScalarVector& aControl, const std::string& aName) {
        requireCriterion(aName);
        double v = 0.0;
        check(pa_plasticity_problem_criterion_value(mProblem.get(), aControl.data(), mMultiplier, &v, nullptr));
        return v;
    }
    ScalarVector criterionGradient(const ScalarVector& aControl, const std::string& aName) {
        requireCriterion(aName);
        ScalarVector tGrad("criterion gradient", aControl.size());
        check(pa_plasticity_problem_criterion_gradient(mProblem.get(), aControl.data(), mMultiplier, tGrad.data(),
                                                       nullptr));
        return tGrad;
    }
    bool converged() const { return mConverged; }
    // Newton iterations, Krylov iterations and final residual norm of a step
    void stepStats(int aStep, int& aNewton, int& aLinear, double& aResidual) const {
        check(pa_plasticity_problem_stats(mProblem.get(), aStep, &aNewton, &aLinear, &aResidual));
    }
    std::shared_ptr<PlasticityEvaluator> evaluator() const { return mEval; }

  private:
    void requireCriterion(const std::string& aName) const {
        if (aName != mCriterion) throw_error("PlasticityProblem: no criterion '" + aName + "'");
    }
    std::shared_ptr<Mesh> mMesh;
    std::shared_ptr<PlasticityEvaluator> mEval;
    std::shared_ptr<pa_plasticity_problem> mProblem;
    pa_plasticity_options mOptions{};
    std::string mCriterion;
    double mMultiplier = 1.0;
    bool mMultigrid = false;
    int mSmootherSweeps = 0;
    double mSmootherDamping = 0.0;
    bool mConverged = false;
    std::vector<int> mBcDofs;
    std::vector<double> mBcValues;
};

// ------------------------------------------- Plato Engine operations
namespace detail {

// Element tree of a plain XML document, the subset of Plato::InputData that
// the operation definitions use: <Operation><Function>...</Function>
// <Name>...</Name> <Output><Argument>Value</Argument>
// <ArgumentName>...</ArgumentName></Output> ...</Operation>.
struct XmlElement {
    std::string name, text;
    std::vector<XmlElement> children;

    std::vector<const XmlElement*> all(const std::string& aName) const {
        std::vector<const XmlElement*> tOut;
        for (const auto& c : children)
            if (c.name == aName) tOut.push_back(&c);
        return tOut;
    }
    const XmlElement* first(const std::string& aName) const {
        for (const auto& c : children)
            if (c.name == aName) return &c;
        return nullptr;
    }
    // the trimmed text of the first child named aName, or aDefault
    std::string get(const std::string& aName, const std::string& aDefault = "") const {
        const XmlElement* c = first(aName);
        if (!c) return aDefault;
        const size_t i = c->text.find_first_not_of(" \t\n\r"), j = c->text.find_last_not_of(" \t\n\r");
        return i == std::string::npos ? std::string() : c->text.substr(i, j - i + 1);
    }
};

// The document as a root element holding its top-level elements (comments
// and processing instructions are skipped; attributes are ignored).
inline XmlElement parseXmlElements(const std::string& aXml) {
    XmlElement tRoot;
    std::vector<XmlElement*> tStack{&tRoot};
    size_t i = 0;
    while (i < aXml.size()) {
        const size_t lt = aXml.find('<', i);
        const std::string tText = aXml.substr(i, (lt == std::string::npos ? aXml.size() : lt) - i);
        if (tStack.size() > 1) tStack.back()->text += xml_unescape(tText);
        if (lt == std::string::npos) break;
        if (aXml.compare(lt, 4, "<!--") == 0) {
            const size_t e = aXml.find("-->", lt + 4);
            if (e == std::string::npos) throw_error("XML: unterminated comment");
            i = e + 3;
            continue;
        }
        const size_t gt = aXml.find('>', lt);
        if (gt == std::string::npos) throw_error("XML: unterminated tag");
        std::string tBody = aXml.substr(lt + 1, gt - lt - 1);
        i = gt + 1;
        if (!tBody.empty() && (tBody[0] == '?' || tBody[0] == '!')) continue;
        if (!tBody.empty() && tBody[0] == '/') {
            std::string tName = tBody.substr(1);
            tName.erase(tName.find_last_not_of(" \t\n\r") + 1);
            if (tStack.size() < 2 || tStack.back()->name != tName)
                throw_error("XML: unexpected closing tag </" + tName + ">");
            tStack.pop_back();
            continue;
        }
        const bool tSelfClosing = !tBody.empty() && tBody.back() == '/';
        if (tSelfClosing) tBody.pop_back();
        const size_t n = tBody.find_first_of(" \t\n\r");
        XmlElement tElement;
        tElement.name = tBody.substr(0, n);
        if (tElement.name.empty()) throw_error("XML: empty tag name");
        tStack.back()->children.push_back(tElement);
        if (!tSelfClosing) tStack.push_back(&tStack.back()->children.back());
    }
    if (tStack.size() != 1) throw_error("XML: unterminated element <" + tStack.back()->name + ">");
    return tRoot;
}

// AnalyzeAppUtils' parse_inline (src/AnalyzeAppUtils.cpp:46-118): the target
// "[list]:[list]:name" or "...:name(i)" set to aValue.
inline void parse_inline(ParameterList& aParams, const std::string& aTarget, double aValue) {
    std::vector<std::string> tTokens;
    {
        std::istringstream s(aTarget);
        std::string t;
        while (std::getline(s, t, ':')) tTokens.push_back(t);
    }
    if (tTokens.empty()) throw_error("parameter target '" + aTarget + "' is empty");
    ParameterList* tList = &aParams;
    size_t k = 0;
    for (; k + 1 < tTokens.size() && tTokens[k].size() > 1 && tTokens[k].front() == '[' && tTokens[k].back() == ']';
         ++k) {
        const std::string tName = tTokens[k].substr(1, tTokens[k].size() - 2);
        if (!tList->isSublist(tName)) throw_error("parameter target '" + aTarget + "': no sublist '" + tName + "'");
        tList = &tList->sublist(tName);
    }
    const std::string& tLast = tTokens[k];
    const size_t p1 = tLast.find('('), p2 = tLast.find(')');
    if (p1 != std::string::npos && p2 != std::string::npos && p2 > p1) {
        const std::string tName = tLast.substr(0, p1);
        auto tVec = tList->get<std::vector<double>>(tName);
        const int tEntry = std::stoi(tLast.substr(p1 + 1, p2 - p1 - 1));
        if (tEntry < 0 || tEntry >= (int)tVec.size())
            throw_error("parameter target '" + aTarget + "': entry out of range");
        tVec[tEntry] = aValue;
        std::ostringstream o;
        o.precision(17);
        o << "{";
        for (size_t q = 0; q < tVec.size(); ++q) o << (q ? ", " : "") << tVec[q];
        o << "}";
        tList->set(tName, "Array(double)", o.str());
    } else {
        tList->set(tLast, aValue);
    }
}

// The problem an application drives, whatever its physics.
class AppProblem {
  public:
    virtual ~AppProblem() = default;
    virtual void solve(const ScalarVector& aControl) = 0;
    virtual bool solved() const = 0;
    virtual Scalar value(const ScalarVector& aControl, const std::string& aCriterion) = 0;
    virtual ScalarVector gradient(const ScalarVector& aControl, const std::string& aCriterion) = 0;
    virtual ScalarVector gradientX(const ScalarVector& aControl, const std::string& aCriterion) = 0;
    virtual std::vector<std::string> dofNames() const = 0;
    virtual ScalarVector lastState() const = 0;
};

template <class PhysicsT>
class EllipticAppProblem : public AppProblem {
  public:
    EllipticAppProblem(std::shared_ptr<Mesh> aMesh, const ParameterList& aParams) : mProblem(aMesh, aParams) {}
    void solve(const ScalarVector& aControl) override { mSolution = mProblem.solution(aControl); }
    bool solved() const override { return !mSolution.empty(); }
    Scalar value(const ScalarVector& aControl, const std::string& aCriterion) override {
        return mProblem.criterionValue(aControl, mSolution, aCriterion);
    }
    ScalarVector gradient(const ScalarVector& aControl, const std::string& aCriterion) override {
        return mProblem.criterionGradient(aControl, mSolution, aCriterion);
    }
    ScalarVector gradientX(const ScalarVector& aControl, const std::string& aCriterion) override {
        return mProblem.criterionGradientX(aControl, mSolution, aCriterion);
    }
    std::vector<std::string> dofNames() const override { return PhysicsT::dofNames(); }
    ScalarVector lastState() const override { return mSolution.get("State", mSolution.numSteps() - 1); }

  private:
    Elliptic::Problem<PhysicsT> mProblem;
    Solutions mSolution;
};

template <class PhysicsT>
class PlasticityAppProblem : public AppProblem {
  public:
    PlasticityAppProblem(std::shared_ptr<Mesh> aMesh, const ParameterList& aParams) : mProblem(aMesh, aParams) {}
    void solve(const ScalarVector& aControl) override { mSolution = mProblem.solution(aControl); }
    bool solved() const override { return !mSolution.empty(); }
    Scalar value(const ScalarVector& aControl, const std::string& aCriterion) override {
        return mProblem.criterionValue(aControl, aCriterion);
    }
    ScalarVector gradient(const ScalarVector& aControl, const std::string& aCriterion) override {
        return mProblem.criterionGradient(aControl, aCriterion);
    }
    ScalarVector gradientX(const ScalarVector&, const std::string&) override {
        throw_error("AnalyzeApp: the coordinate gradient of a plasticity criterion is not on this path");
    }
    std::vector<std::string> dofNames() const override { return PhysicsT::dofNames(); }
    ScalarVector lastState() const override { return mSolution.get("State", mSolution.numSteps() - 1); }

  private:
    PlasticityProblem<PhysicsT> mProblem;
    Solutions mSolution;
};

}  // namespace detail

// The Plato Engine application of this path (MPMD_App, src/Analyze_App.hpp:
// 64-1000, src/Analyze_App.cpp): operations named in an operations document
// run on the shared data the engine imports and exports.  Functions:
// ComputeSolution, ComputeCriterion, ComputeCriterionValue,
// ComputeCriterionGradient, ComputeCriterionX, ComputeCriterionGradientX,
// Reinitialize (with OnChange) and UpdateProblem (src/Analyze_App.cpp:
// 368-500, 700-1320); an operation's Parameters retarget the input deck as
// parse_inline does.  Physics 'Mechanical' and 'Thermal' (elliptic,
// self-adjoint criteria) and 'Plasticity' (the mixed J2 form, the plastic work
// criterion).  Other functions (output, ESP, MLS, Helmholtz, finite
// differences, mesh reload) and ProblemDefinitions other than the input deck
// throw at construction.  One rank: a scalar field's local ids are the node
// ids.
class AnalyzeApp {
  public:
    // aInput: the input deck, whose 'Plato Problem' sublist (or, without one,
    // the list itself) defines the problem; parameter targets are paths from
    // the deck's root, as in the reference ("[Plato Problem]:[...]:name")
    AnalyzeApp(std::shared_ptr<Mesh> aMesh, const ParameterList& aInput, const std::string& aOperationsXml)
        : mMesh(std::move(aMesh)), mDef(aInput.clone()) {
        const detail::XmlElement tDoc = detail::parseXmlElements(aOperationsXml);
        std::vector<const detail::XmlElement*> tOps;
        for (const auto& c : tDoc.children) {
            if (c.name == "Operation") tOps.push_back(&c);
            for (const auto* o : c.all("Operation")) tOps.push_back(o);  // operations under a wrapper element
        }
        for (const auto* o : tOps) addOperation(*o);
    }

    // MPMD_App::initialize (src/Analyze_App.cpp:338-349): the control starts at 1
    void initialize() {
        mControl = toDevice(std::vector<Scalar>((size_t)mMesh->NumNodes(), 1.0), "control");
        createProblem();
    }
    // MPMD_App::reinitialize (src/Analyze_App.cpp:511-518)
    void reinitialize() { createProblem(); }

    // MPMD_App::compute (src/Analyze_App.cpp:563-590)
    void compute(const std::string& aOperationName) {
        auto it = mOperations.find(aOperationName);
        if (it == mOperations.end()) throw_error("AnalyzeApp: no operation named '" + aOperationName + "'");
        if (!mProblem || mModified) createProblem();
        run(it->second);
    }

    // importData: a SCALAR_FIELD named "Topology" (the control, one value per node)
    void importData(const std::string& aName, const std::vector<Scalar>& aField) {
        if (aName != "Topology") throw_error("AnalyzeApp: cannot import field '" + aName + "'");
        if (aField.size() != (size_t)mMesh->NumNodes()) throw_error("AnalyzeApp: 'Topology' needs one value per node");
        mControl = toDevice(aField, "control");
    }
    // importData of a SCALAR_PARAMETER: the value of parameter aName of
    // operation aOperation (MPMD_App::importScalarParameter, src/Analyze_App.hpp:
    // 228-241); a parameter with a Target updates the input deck, and the
    // problem is recreated when an operation next runs
    void importParameter(const std::string& aName, const std::string& aOperation, Scalar aValue) {
        auto it = mOperations.find(aOperation);
        if (it == mOperations.end()) throw_error("AnalyzeApp: no operation named '" + aOperation + "'");
        auto p = it->second.parameters.find(aName);
        if (p == it->second.parameters.end())
            throw_error("Attempted to update a parameter ('" + aName + "') that wasn't defined for this operation");
        p->second.value = aValue;
        if (!p->second.target.empty()) {
            detail::parse_inline(*mDef, p->second.target, aValue);
            mModified = true;
        }
    }
    // importData of a SCALAR value (MPMD_App::importScalarValue,
    // src/Analyze_App.hpp:248-260): "Parameters" feeds Reinitialize's OnChange
    void importValue(const std::string& aName, const std::vector<Scalar>& aValues) {
        auto it = mValues.find(aName);
        if (it == mValues.end())
            throw_error("Attempted to import SharedValue ('" + aName + "') that doesn't exist.");
        it->second = aValues;
    }

    // exportData (src/Analyze_App.hpp:287-450): "Topology"; a criterion value
    // (minus its target), a gradient with respect to the control, the
    // component i of a coordinate gradient ("name@i", i = 0 by default), or a
    // component of the last solution by dof name ("displacement X", ...)
    void exportData(const std::string& aName, std::vector<Scalar>& aOut) const {
        const size_t at = aName.find('@');
        const std::string tName = aName.substr(0, at);
        const int tIndex = at == std::string::npos ? 0 : std::atoi(aName.c_str() + at + 1);
        if (tName == "Topology") {
            aOut = toHost(mControl);
            return;
        }
        if (auto v = mValueNames.find(tName); v != mValueNames.end()) {
            auto c = mCriterionValues.find(v->second);
            if (c == mCriterionValues.end())
                throw_error("Attempted to export SharedValue ('" + tName + "') that doesn't exist.");
            aOut.assign(1, c->second);
            return;
        }
        if (auto g = mGradZNames.find(tName); g != mGradZNames.end()) {
            auto c = mGradZ.find(g->second);
            if (c == mGradZ.end()) throw_error("AnalyzeApp: gradient '" + tName + "' has not been computed");
            aOut = toHost(c->second);
            return;
        }
        if (auto g = mGradXNames.find(tName); g != mGradXNames.end()) {
            auto c = mGradX.find(g->second);
            if (c == mGradX.end()) throw_error("AnalyzeApp: gradient '" + tName + "' has not been computed");
            if (tIndex < 0 || tIndex > 2) throw_error("AnalyzeApp: '" + aName + "': component out of range");
            const auto h = toHost(c->second);
            aOut.resize(h.size() / 3);
            for (size_t n = 0; n < aOut.size(); ++n) aOut[n] = h[3 * n + tIndex];
            return;
        }
        if (auto v = mValues.find(tName); v != mValues.end()) {
            aOut = v->second;
            return;
        }
        if (mProblem && mProblem->solved()) {
            const auto tDofs = mProblem->dofNames();
            const auto d = std::find(tDofs.begin(), tDofs.end(), tName);
            if (d != tDofs.end()) {
                const auto h = toHost(mProblem->lastState());
                const size_t tStride = tDofs.size(), tDof = (size_t)(d - tDofs.begin());
                aOut.resize(h.size() / tStride);
                for (size_t n = 0; n < aOut.size(); ++n) aOut[n] = h[tStride * n + tDof];
                return;
            }
        }
        throw_error("AnalyzeApp: nothing to export under the name '" + aName + "'");
    }
    // exportDataMap(SCALAR_FIELD): the node ids this rank owns (all of them)
    void exportDataMap(std::vector<int>& aLocalIDs) const {
        aLocalIDs.resize((size_t)mMesh->NumNodes());
        for (size_t n = 0; n < aLocalIDs.size(); ++n) aLocalIDs[n] = (int)n;
    }
    std::vector<std::string> operationNames() const {
        std::vector<std::string> tOut;
        for (const auto& kv : mOperations) tOut.push_back(kv.first);
        return tOut;
    }

  private:
    struct Parameter {
        std::string target;
        Scalar value = 0.0;
    };
    struct Operation {
        std::string function, criterion;
        Scalar target = 0.0;
        bool onChange = false;
        std::vector<Scalar> lastParameters;
        bool hasLast = false;
        std::map<std::string, Parameter> parameters;
    };

    // getArgumentName (src/Analyze_App.cpp:80-110): the ArgumentName of the
    // Output whose Argument is aArgument
    static std::string argumentName(const detail::XmlElement& aOp, const std::string& aArgument) {
        for (const auto* o : aOp.all("Output")) {
            if (o->get("Argument") != aArgument) continue;
            const std::string n = o->get("ArgumentName");
            if (n.empty()) throw_error("Parsing operation: 'ArgumentName' missing");
            return n;
        }
        return "";
    }
    // addUnique (src/Analyze_App.cpp:33-60)
    static void addUnique(std::map<std::string, std::string>& aMap, const std::string& aKey,
                          const std::string& aCriterion) {
        if (aKey.empty()) return;
        auto it = aMap.find(aKey);
        if (it != aMap.end() && it->second != aCriterion)
            throw_error("ArgumentName '" + aKey + "' is already associated with Criterion '" + it->second +
                        "' and cannot be reassigned to Criterion '" + aCriterion + "'");
        aMap[aKey] = aCriterion;
    }

    void addOperation(const detail::XmlElement& aNode) {
        Operation op;
        op.function = aNode.get("Function");
        const std::string tName = aNode.get("Name");
        if (tName.empty()) throw_error("AnalyzeApp: an Operation has no Name");
        if (mOperations.count(tName)) throw_error("AnalyzeApp: operation '" + tName + "' is defined twice");
        const std::string tProblem = aNode.get("ProblemDefinition");
        if (!tProblem.empty())
            throw_error("AnalyzeApp: operation '" + tName + "': ProblemDefinition '" + tProblem +
                        "' (only the input deck's problem is on this path)");
        for (const auto* pn : aNode.all("Parameter")) {
            const std::string n = pn->get("ArgumentName");
            if (op.parameters.count(n)) throw_error("ArgumentNames must be unique.");
            Parameter p;
            p.target = pn->get("Target");
            p.value = std::atof(pn->get("InitialValue", "0").c_str());
            op.parameters[n] = p;
        }
        const std::string& f = op.function;
        const bool tCriterionOp = f == "ComputeCriterion" || f == "ComputeCriterionValue" ||
                                  f == "ComputeCriterionGradient" || f == "ComputeCriterionX" ||
                                  f == "ComputeCriterionGradientX";
        if (tCriterionOp) {
            // CriterionOp (src/Analyze_App.cpp:621-633)
            op.criterion = aNode.get("Criterion");
            if (op.criterion.empty()) throw_error("Required keyword ('Criterion') missing.");
            op.target = std::atof(aNode.get("Target", "0").c_str());
            const std::string v = argumentName(aNode, "Value"), g = argumentName(aNode, "Gradient");
            if (f == "ComputeCriterion" || f == "ComputeCriterionValue" || f == "ComputeCriterionX")
                addUnique(mValueNames, v, op.criterion);
            if (f == "ComputeCriterion" || f == "ComputeCriterionGradient") addUnique(mGradZNames, g, op.criterion);
            if (f == "ComputeCriterionX" || f == "ComputeCriterionGradientX") addUnique(mGradXNames, g, op.criterion);
            if (f == "ComputeCriterionX" && !aNode.get("OutputFile").empty())
                throw_error("AnalyzeApp: ComputeCriterionX 'OutputFile' (HDF5) is not on this path");
        } else if (f == "ComputeSolution") {
            if (aNode.first("WriteOutput"))
                throw_error("AnalyzeApp: ComputeSolution 'WriteOutput' (visualization files) is not on this path");
        } else if (f == "Reinitialize") {
            // OnChangeOp (src/Analyze_App.cpp:636-643)
            op.onChange = aNode.get("OnChange", "false") == "true";
            mValues["Parameters"];
        } else if (f != "UpdateProblem") {
            throw_error("AnalyzeApp: operation function '" + f + "' is not on this path");
        }
        mOperations[tName] = op;
    }

    void createProblem() {
        const ParameterList& tDef = mDef->isSublist("Plato Problem") ? mDef->sublist("Plato Problem") : *mDef;
        const std::string tPhysics = tDef.get<std::string>("Physics", std::string(""));
        if (tPhysics == "Mechanical")
            mProblem = std::make_unique<detail::EllipticAppProblem<Mechanics<3>>>(mMesh, tDef);
        else if (tPhysics == "Thermal")
            mProblem = std::make_unique<detail::EllipticAppProblem<Thermal<3>>>(mMesh, tDef);
        else if (tPhysics == "Plasticity")
            mProblem = std::make_unique<detail::PlasticityAppProblem<InfinitesimalStrainPlasticity<3>>>(mMesh, tDef);
        else
            throw_error("AnalyzeApp: Physics '" + tPhysics + "' is not on this path");
        if (mControl.size() == 0)
            mControl = toDevice(std::vector<Scalar>((size_t)mMesh->NumNodes(), 1.0), "control");
        mModified = false;
    }

    void run(Operation& op) {
        const std::string& f = op.function;
        if (f == "ComputeSolution") {
            mProblem->solve(mControl);
        } else if (f == "ComputeCriterion") {
            // src/Analyze_App.cpp:727-745: solve, value minus target, gradient
            mProblem->solve(mControl);
            mCriterionValues[op.criterion] = mProblem->value(mControl, op.criterion) - op.target;
            mGradZ[op.criterion] = mProblem->gradient(mControl, op.criterion);
        } else if (f == "ComputeCriterionValue") {
            mProblem->solve(mControl);
            mCriterionValues[op.criterion] = mProblem->value(mControl, op.criterion) - op.target;
        } else if (f == "ComputeCriterionGradient") {
            // at the last solution (src/Analyze_App.cpp:996-1003)
            if (!mProblem->solved()) mProblem->solve(mControl);
            mGradZ[op.criterion] = mProblem->gradient(mControl, op.criterion);
        } else if (f == "ComputeCriterionX") {
            mProblem->solve(mControl);
            mCriterionValues[op.criterion] = mProblem->value(mControl, op.criterion) - op.target;
            mGradX[op.criterion] = mProblem->gradientX(mControl, op.criterion);
        } else if (f == "ComputeCriterionGradientX") {
            if (!mProblem->solved()) mProblem->solve(mControl);
            mGradX[op.criterion] = mProblem->gradientX(mControl, op.criterion);
        } else if (f == "Reinitialize") {
            // src/Analyze_App.cpp:1214-1235: with OnChange, only when the
            // imported "Parameters" differ from the last call's
            const auto& tParams = mValues["Parameters"];
            const bool tChanged = !op.hasLast || tParams != op.lastParameters;
            op.lastParameters = tParams;
            op.hasLast = true;
            if (!op.onChange || tChanged) createProblem();
        }
        // UpdateProblem (src/Analyze_App.cpp:1310-1318) updates the criteria's
        // continuation parameters; the criteria on this path have none
    }

    std::shared_ptr<Mesh> mMesh;
    std::shared_ptr<ParameterList> mDef;
    bool mModified = false;
    std::unique_ptr<detail::AppProblem> mProblem;
    ScalarVector mControl;
    std::map<std::string, Operation> mOperations;
    std::map<std::string, std::string> mValueNames, mGradZNames, mGradXNames;
    std::map<std::string, Scalar> mCriterionValues;
    std::map<std::string, ScalarVector> mGradZ, mGradX;
    std::map<std::string, std::vector<Scalar>> mValues;
};


}  // namespace amd
}  // namespace Plato

#endif  // PLATOANALYZE_AMD_HPP
