// CPU FP64 oracle for the Tet4 elliptic assembly path (SIMP elastostatics and
// steady thermal).  TEST INFRASTRUCTURE ONLY: this file is the checker that the
// HIP library is compared against.  Only tests/, the smoke check and the CPU
// baseline leg of bench.py load it; the product library never links or calls it.
//
// It restates Plato Analyze's algorithm with the reference's own structure:
//   * three separate gather passes into cell worksets (state, control, config),
//     src/Assembly.hpp:630-1204, src/WorksetBase.hpp:178-405;
//   * per-cell evaluation in forward-mode dual numbers (the Sacado SFad
//     equivalent, src/SimplexFadTypes.hpp:13-25): width 4*dpn for gradient_u,
//     width 4 for gradient_z;
//   * the block-CRS graph rebuilt inside every gradient call
//     (src/elliptic/VectorFunction.hpp:541, src/ImplicitFunctors.hpp:932-1005);
//   * scatter by linear search of the CRS row plus atomic add
//     (src/ImplicitFunctors.hpp:730-766, src/Assembly.hpp:1222-1521).
// Every function below cites the reference lines it follows.  With OpenMP the
// atomics are real (omp atomic), so the same code is the multi-core CPU
// baseline.
//
// Pinning (what fixes each function independently of this file):
//   mesh, graph, gradients, strain, stress, residual, Jacobian, gradient_z,
//   gradient_x, energies -> unit_tests/DerivativeTests.cpp goldens (tests/golden/);
//   Jacobian  -> also the analytic B^T C B closed form and central FD;
//   gradient_z, gradient_x and criterion gradients -> central finite
//   differences (gradient_x also: translation invariance);
//   Dirichlet -> symmetric elimination invariants (row/col zero, unit diag).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <utility>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Forward-mode dual number of fixed width (SFad<double,W> equivalent).
// ---------------------------------------------------------------------------
template <int W>
struct Fad {
    double v;
    double d[W];
    Fad() : v(0.0) { for (int i = 0; i < W; ++i) d[i] = 0.0; }
    Fad(double x) : v(x) { for (int i = 0; i < W; ++i) d[i] = 0.0; }
    Fad& operator=(double x) { v = x; for (int i = 0; i < W; ++i) d[i] = 0.0; return *this; }
    Fad& operator+=(const Fad& o) { v += o.v; for (int i = 0; i < W; ++i) d[i] += o.d[i]; return *this; }
    Fad& operator-=(const Fad& o) { v -= o.v; for (int i = 0; i < W; ++i) d[i] -= o.d[i]; return *this; }
    Fad& operator*=(const Fad& o) {
        for (int i = 0; i < W; ++i) d[i] = d[i] * o.v + v * o.d[i];
        v *= o.v; return *this;
    }
    Fad& operator*=(double s) { v *= s; for (int i = 0; i < W; ++i) d[i] *= s; return *this; }
    Fad& operator/=(double s) { v /= s; for (int i = 0; i < W; ++i) d[i] /= s; return *this; }
};
template <int W> Fad<W> operator+(const Fad<W>& a, const Fad<W>& b) { Fad<W> r(a); r += b; return r; }
template <int W> Fad<W> operator-(const Fad<W>& a, const Fad<W>& b) { Fad<W> r(a); r -= b; return r; }
template <int W> Fad<W> operator*(const Fad<W>& a, const Fad<W>& b) { Fad<W> r(a); r *= b; return r; }
template <int W> Fad<W> operator*(const Fad<W>& a, double s) { Fad<W> r(a); r *= s; return r; }
template <int W> Fad<W> operator*(double s, const Fad<W>& a) { Fad<W> r(a); r *= s; return r; }
template <int W> Fad<W> operator+(const Fad<W>& a, double s) { Fad<W> r(a); r.v += s; return r; }
template <int W> Fad<W> operator+(double s, const Fad<W>& a) { Fad<W> r(a); r.v += s; return r; }
template <int W> Fad<W> operator-(const Fad<W>& a, double s) { Fad<W> r(a); r.v -= s; return r; }
template <int W> Fad<W> operator/(const Fad<W>& a, double s) { Fad<W> r(a); r /= s; return r; }
template <int W> Fad<W> operator-(const Fad<W>& a) { Fad<W> r(a); r *= -1.0; return r; }
template <int W> Fad<W> operator/(const Fad<W>& a, const Fad<W>& b) {
    Fad<W> r;
    r.v = a.v / b.v;
    for (int i = 0; i < W; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v);
    return r;
}
inline double absval(double x) { return std::fabs(x); }
template <int W> Fad<W> absval(const Fad<W>& a) { return a.v < 0.0 ? -a : a; }
inline double fsqrt(double x) { return std::sqrt(x); }
template <int W> Fad<W> fsqrt(const Fad<W>& a) {
    Fad<W> r;
    r.v = std::sqrt(a.v);
    for (int i = 0; i < W; ++i) r.d[i] = a.d[i] / (2.0 * r.v);
    return r;
}
template <int W> Fad<W> fpow(const Fad<W>& a, double p) {
    Fad<W> r;
    r.v = std::pow(a.v, p);
    const double dp = (a.v == 0.0) ? 0.0 : p * std::pow(a.v, p - 1.0);
    for (int i = 0; i < W; ++i) r.d[i] = dp * a.d[i];
    return r;
}
inline double fpow(double a, double p) { return std::pow(a, p); }

inline double val(double x) { return x; }
template <int W> inline double val(const Fad<W>& x) { return x.v; }

// FAD seeding of the gathered state / control entries.
inline void seed_state(double&, int) {}
template <int W> inline void seed_state(Fad<W>& s, int k) { s.d[k] = 1.0; }
inline void seed_control(double&, int) {}
template <int W> inline void seed_control(Fad<W>& s, int k) { s.d[k] = 1.0; }

template <class T> inline void atomic_add(double* dst, T x) {
#ifdef _OPENMP
#pragma omp atomic
#endif
    *dst += x;
}

// ---------------------------------------------------------------------------
// Problem description shared with the Python side (ctypes mirror).
// ---------------------------------------------------------------------------
}  // namespace

extern "C" {

typedef struct {
    int num_nodes;
    int num_cells;
    const double* coords;     // N*3, interleaved
    const int* conn;          // C*4
    double youngs, poisson;   // Isotropic Linear Elastic
    double conductivity;      // Thermal Conduction
    double penalty_exponent;  // SIMP Exponent
    double penalty_min;       // SIMP Minimum Value
    int num_faces;            // side set faces carrying a uniform load
    const int* face_cells;    // F
    const int* face_lnodes;   // F*3 local node ordinals in the cell
    double flux[3];           // uniform traction (dpn 3) or flux (dpn 1)
    int has_body;             // constant body load/source present
    int body_dof;
    double body_value;
    const double* body_cells;  // body load given per cell (the function at the centroid), or null
    int body_cells_dof;
} orc_problem;

}  // extern "C"

namespace {

// ---------------------------------------------------------------------------
// Geometry: ComputeGradientWorkset<3>, src/ImplicitFunctors.hpp:169-221.
// cfg[n][d]; J = [x_a - x_3] with a = column; grads are the rows of J^-1.
// ---------------------------------------------------------------------------
// The config type K is double, or a config FAD for gradient_x.
template <class K>
inline void compute_gradient(const K cfg[4][3], K g[4][3], K& vol) {
    K j11 = cfg[0][0] - cfg[3][0], j12 = cfg[1][0] - cfg[3][0], j13 = cfg[2][0] - cfg[3][0];
    K j21 = cfg[0][1] - cfg[3][1], j22 = cfg[1][1] - cfg[3][1], j23 = cfg[2][1] - cfg[3][1];
    K j31 = cfg[0][2] - cfg[3][2], j32 = cfg[1][2] - cfg[3][2], j33 = cfg[2][2] - cfg[3][2];
    K detj = j11 * j22 * j33 + j12 * j23 * j31 + j13 * j21 * j32
           - j11 * j23 * j32 - j12 * j21 * j33 - j13 * j22 * j31;
    K i11 = (j22 * j33 - j23 * j32) / detj, i12 = (j13 * j32 - j12 * j33) / detj, i13 = (j12 * j23 - j13 * j22) / detj;
    K i21 = (j23 * j31 - j21 * j33) / detj, i22 = (j11 * j33 - j13 * j31) / detj, i23 = (j13 * j21 - j11 * j23) / detj;
    K i31 = (j21 * j32 - j22 * j31) / detj, i32 = (j12 * j31 - j11 * j32) / detj, i33 = (j11 * j22 - j12 * j21) / detj;
    vol = absval(detj);
    g[0][0] = i11; g[0][1] = i12; g[0][2] = i13;
    g[1][0] = i21; g[1][1] = i22; g[1][2] = i23;
    g[2][0] = i31; g[2][1] = i32; g[2][2] = i33;
    g[3][0] = -(i11 + i21 + i31); g[3][1] = -(i12 + i22 + i32); g[3][2] = -(i13 + i23 + i33);
}

// Isotropic stiffness, src/IsotropicLinearElasticMaterial.cpp:86-107.
inline void isotropic_c(double E, double nu, double C[6][6]) {
    for (int i = 0; i < 6; ++i) for (int j = 0; j < 6; ++j) C[i][j] = 0.0;
    double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    C[0][0] = c * (1.0 - nu); C[0][1] = c * nu; C[0][2] = c * nu;
    C[1][0] = c * nu; C[1][1] = c * (1.0 - nu); C[1][2] = c * nu;
    C[2][0] = c * nu; C[2][1] = c * nu; C[2][2] = c * (1.0 - nu);
    C[3][3] = 1.0 / 2.0 * c * (1.0 - 2.0 * nu);
    C[4][4] = 1.0 / 2.0 * c * (1.0 - 2.0 * nu);
    C[5][5] = 1.0 / 2.0 * c * (1.0 - 2.0 * nu);
}

// Voigt map of StressDivergence<3>, src/StressDivergence.hpp:33-48:
// diag 0,1,2; then (1,2)->3, (0,2)->4, (0,1)->5.
const int kVoigt[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};

// MSIMP, src/Simp.hpp:130-134.
template <class T> inline T msimp(const T& x, double p, double emin) {
    return emin + (1.0 - emin) * fpow(x, p);
}

// ---------------------------------------------------------------------------
// Gathers (three separate passes): src/Assembly.hpp:630-757 (control),
// :775-924 (state), :1068-1204 (config).  FAD seeding: derivative slot of
// local dof = node*dpn + d (src/Assembly.hpp:794).
// ---------------------------------------------------------------------------
void workset_config(const orc_problem* P, std::vector<double>& ws) {
    const int C = P->num_cells;
    ws.assign((size_t)C * 12, 0.0);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c)
        for (int n = 0; n < 4; ++n)
            for (int d = 0; d < 3; ++d)
                ws[(size_t)c * 12 + n * 3 + d] = P->coords[(size_t)P->conn[(size_t)c * 4 + n] * 3 + d];
}

// workset_config_fad, src/Assembly.hpp:1130-1165: config FAD slot node*3 + dim.
void workset_config_fad(const orc_problem* P, std::vector<Fad<12>>& ws) {
    const int C = P->num_cells;
    ws.assign((size_t)C * 12, Fad<12>(0.0));
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c)
        for (int n = 0; n < 4; ++n)
            for (int d = 0; d < 3; ++d) {
                Fad<12> x(P->coords[(size_t)P->conn[(size_t)c * 4 + n] * 3 + d]);
                x.d[n * 3 + d] = 1.0;
                ws[(size_t)c * 12 + n * 3 + d] = x;
            }
}

template <class S, int DPN>
void workset_state(const orc_problem* P, const double* u, std::vector<S>& ws) {
    const int C = P->num_cells;
    ws.assign((size_t)C * 4 * DPN, S(0.0));
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c)
        for (int n = 0; n < 4; ++n)
            for (int d = 0; d < DPN; ++d) {
                S s(u[(size_t)P->conn[(size_t)c * 4 + n] * DPN + d]);
                seed_state(s, n * DPN + d);
                ws[(size_t)c * 4 * DPN + n * DPN + d] = s;
            }
}

template <class Z>
void workset_control(const orc_problem* P, const double* z, std::vector<Z>& ws) {
    const int C = P->num_cells;
    ws.assign((size_t)C * 4, Z(0.0));
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c)
        for (int n = 0; n < 4; ++n) {
            Z s(z[P->conn[(size_t)c * 4 + n]]);
            seed_control(s, n);
            ws[(size_t)c * 4 + n] = s;
        }
}


// ---------------------------------------------------------------------------
// Graph: NodeNodeGraph (Omega_h ask_star of vertices, sorted, no diagonal;
// src/OmegaHMesh.cpp:137-149) then CreateBlockMatrix inserts the diagonal in
// sorted position (src/ImplicitFunctors.hpp:932-1005).
// ---------------------------------------------------------------------------
void build_graph(int N, int C, const int* conn, std::vector<int>& rowmap, std::vector<int>& colidx) {
    std::vector<std::vector<int>> star(N);
    for (int c = 0; c < C; ++c)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                if (a != b) star[conn[(size_t)c * 4 + a]].push_back(conn[(size_t)c * 4 + b]);
    rowmap.assign(N + 1, 0);
    for (int i = 0; i < N; ++i) {
        auto& s = star[i];
        std::sort(s.begin(), s.end());
        s.erase(std::unique(s.begin(), s.end()), s.end());
        rowmap[i + 1] = rowmap[i] + (int)s.size() + 1;
    }
    colidx.assign(rowmap[N], 0);
    for (int i = 0; i < N; ++i) {
        // merge the diagonal into the sorted star, as CreateBlockMatrix does
        int w = rowmap[i];
        bool inserted = false;
        for (int j : star[i]) {
            if (!inserted && j > i) { colidx[w++] = i; inserted = true; }
            colidx[w++] = j;
        }
        if (!inserted) colidx[w++] = i;
    }
}

// BlockMatrixEntryOrdinal, src/ImplicitFunctors.hpp:730-766 (linear search).
inline long entry_ordinal(const int* rowmap, const int* colidx, int row, int col) {
    for (int e = rowmap[row]; e < rowmap[row + 1]; ++e)
        if (colidx[e] == col) return e;
    return -1;
}

// ---------------------------------------------------------------------------
// Elastostatic residual evaluate, src/elliptic/ElastostaticResidual.hpp:165-234
// with Strain (src/Strain.hpp:56-95), LinearStress (src/LinearStress.hpp:102-120),
// ApplyWeighting+MSIMP (src/ApplyWeighting.hpp:48-65), StressDivergence
// (src/StressDivergence.hpp:70-90).  Cubature weight 1/6
// (src/LinearTetCubRuleDegreeOne.hpp:72-78).
// ---------------------------------------------------------------------------
// BodyLoad::get with scale -1 for a load given per cell
// (src/BodyLoads.hpp:54-122; the function evaluated at the one-point
// cubature's point, the centroid, src/PlatoMeshExpr.hpp:16-50): weight
// -1 * (1/6) |det J|, times f_c, basis 1/4, unpenalised mean cell density,
// on dof body_cells_dof of each node.  f_c enters as a constant, as in the
// reference, whose cubature points are mapped in plain doubles.
template <class K, class Z, class R>
void body_cell_load(const orc_problem* P, int c, const K& detj_abs, const std::vector<Z>& ctrlWS, int dpn, R* out) {
    Z dens = 0.0;
    for (int n = 0; n < 4; ++n) dens += ctrlWS[(size_t)c * 4 + n];
    dens = dens / 4.0;
    K w = -1.0 * (1.0 / 6.0) * detj_abs;
    for (int n = 0; n < 4; ++n) out[n * dpn + P->body_cells_dof] += w * P->body_cells[c] * 0.25 * dens;
}

template <class S, class Z, class R, class K>
void elastic_evaluate(const orc_problem* P, const std::vector<K>& cfgWS, const std::vector<S>& stateWS,
                      const std::vector<Z>& ctrlWS, std::vector<R>& result) {
    using E = decltype(std::declval<S>() * std::declval<K>());  // strain type
    const int C = P->num_cells;
    double Cm[6][6];
    isotropic_c(P->youngs, P->poisson, Cm);
    const double qw = 1.0 / 6.0;
    result.assign((size_t)C * 12, R(0.0));
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c) {
        K cfg[4][3], g[4][3], vol;
        for (int n = 0; n < 4; ++n) for (int d = 0; d < 3; ++d) cfg[n][d] = cfgWS[(size_t)c * 12 + n * 3 + d];
        compute_gradient(cfg, g, vol);
        K detj_abs = vol;
        vol *= qw;
        const S* u = &stateWS[(size_t)c * 12];
        // strain (StrainScalarType: the state type, or the config type for gradient_x)
        E eps[6];
        int v = 0;
        for (int d = 0; d < 3; ++d) {
            eps[v] = 0.0;
            for (int n = 0; n < 4; ++n) eps[v] += u[n * 3 + d] * g[n][d];
            ++v;
        }
        for (int dj = 2; dj >= 1; --dj)
            for (int di = dj - 1; di >= 0; --di) {
                eps[v] = 0.0;
                for (int n = 0; n < 4; ++n) eps[v] += (u[n * 3 + dj] * g[n][di] + u[n * 3 + di] * g[n][dj]);
                ++v;
            }
        // stress
        R sig[6];
        for (int i = 0; i < 6; ++i) {
            sig[i] = 0.0;
            for (int j = 0; j < 6; ++j) sig[i] += (eps[j] - 0.0) * Cm[i][j];
        }
        // weighting by penalised mean nodal density
        Z rho = 0.0;
        for (int n = 0; n < 4; ++n) rho += ctrlWS[(size_t)c * 4 + n];
        rho = rho / 4.0;
        Z s = msimp(rho, P->penalty_exponent, P->penalty_min);
        for (int i = 0; i < 6; ++i) sig[i] *= s;
        // divergence
        R* out = &result[(size_t)c * 12];
        for (int di = 0; di < 3; ++di)
            for (int n = 0; n < 4; ++n)
                for (int dj = 0; dj < 3; ++dj)
                    out[n * 3 + di] += 1.0 * vol * sig[kVoigt[di][dj]] * g[n][dj];
        if (P->body_cells) body_cell_load(P, c, detj_abs, ctrlWS, 3, out);
    }
}

// Thermostatic residual evaluate, src/elliptic/ThermostaticResidual.hpp:128-196:
// ScalarGrad (src/ScalarGrad.hpp:58-74), linear ThermalFlux q = -k grad T
// (src/ThermalFlux.hpp:42-52), FluxDivergence with scale -1
// (src/FluxDivergence.hpp:31-48), then body load (src/BodyLoads.hpp:54-122).
template <class S, class Z, class R, class K>
void thermal_evaluate(const orc_problem* P, const std::vector<K>& cfgWS, const std::vector<S>& stateWS,
                      const std::vector<Z>& ctrlWS, std::vector<R>& result) {
    using E = decltype(std::declval<S>() * std::declval<K>());  // temperature-gradient type
    const int C = P->num_cells;
    const double qw = 1.0 / 6.0;
    const double k = P->conductivity;
    result.assign((size_t)C * 4, R(0.0));
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c) {
        K cfg[4][3], g[4][3], vol;
        for (int n = 0; n < 4; ++n) for (int d = 0; d < 3; ++d) cfg[n][d] = cfgWS[(size_t)c * 12 + n * 3 + d];
        compute_gradient(cfg, g, vol);
        K detj_abs = vol;
        vol *= qw;
        const S* T = &stateWS[(size_t)c * 4];
        E grad[3];
        for (int d = 0; d < 3; ++d) {
            grad[d] = 0.0;
            for (int n = 0; n < 4; ++n) grad[d] += T[n] * g[n][d];
        }
        R flux[3];
        for (int i = 0; i < 3; ++i) {
            flux[i] = 0.0;
            for (int j = 0; j < 3; ++j) flux[i] -= grad[j] * (i == j ? k : 0.0);
        }
        Z rho = 0.0;
        for (int n = 0; n < 4; ++n) rho += ctrlWS[(size_t)c * 4 + n];
        rho = rho / 4.0;
        Z s = msimp(rho, P->penalty_exponent, P->penalty_min);
        for (int i = 0; i < 3; ++i) flux[i] *= s;
        R* out = &result[(size_t)c * 4];
        for (int n = 0; n < 4; ++n)
            for (int d = 0; d < 3; ++d) out[n] += -1.0 * flux[d] * g[n][d] * vol;
        if (P->has_body) {
            // BodyLoad::get with scale -1: weight = -1 * (1/6) * |detJ|,
            // basis value 1/4 at the centroid, times unpenalised cell density
            // (src/BodyLoads.hpp:103-121, src/Plato_TopOptFunctors.hpp:50-62).
            Z dens = 0.0;
            for (int n = 0; n < 4; ++n) dens += ctrlWS[(size_t)c * 4 + n];
            dens = dens / 4.0;
            K w = -1.0 * qw * detj_abs;
            for (int n = 0; n < 4; ++n) out[n] += w * P->body_value * 0.25 * dens;
        }
        if (P->body_cells) body_cell_load(P, c, detj_abs, ctrlWS, 1, out);
    }
}

// The constant elastic body load is not used by the elastic workloads; both
// physics take a load given per cell (body_cell_load).

// Uniform natural BC on a side set, src/SurfaceLoadIntegral.hpp:97-149 with
// CalculateSurfaceJacobians/Area<3> (src/SurfaceIntegralUtilities.hpp): the
// face Jacobian rows are x_{n0}-x_{n2}, x_{n1}-x_{n2}; area weight 1/2; basis 1/3.
void boundary_evaluate(const orc_problem* P, int dpn, double scale, std::vector<double>& result) {
    const int C = P->num_cells;
    result.assign((size_t)C * 4 * dpn, 0.0);
    const double cubw = 0.5, basis = 1.0 / 3.0;
    for (int f = 0; f < P->num_faces; ++f) {
        int c = P->face_cells[f];
        int ln[3] = {P->face_lnodes[f * 3], P->face_lnodes[f * 3 + 1], P->face_lnodes[f * 3 + 2]};
        double J[2][3];
        for (int a = 0; a < 2; ++a)
            for (int d = 0; d < 3; ++d)
                J[a][d] = P->coords[(size_t)P->conn[(size_t)c * 4 + ln[a]] * 3 + d] -
                          P->coords[(size_t)P->conn[(size_t)c * 4 + ln[2]] * 3 + d];
        double j23 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
        double j31 = J[0][2] * J[1][0] - J[0][0] * J[1][2];
        double j12 = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        double area = (scale * cubw) * std::sqrt(j23 * j23 + j31 * j31 + j12 * j12);
        for (int n = 0; n < 3; ++n)
            for (int d = 0; d < dpn; ++d)
                result[(size_t)c * 4 * dpn + ln[n] * dpn + d] += basis * P->flux[d] * area;
    }
}

void assemble_residual(const orc_problem* P, int dpn, const std::vector<double>& res, double* R) {
    const int C = P->num_cells;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c)
        for (int n = 0; n < 4; ++n)
            for (int d = 0; d < dpn; ++d)
                atomic_add(&R[(size_t)P->conn[(size_t)c * 4 + n] * dpn + d], res[(size_t)c * 4 * dpn + n * dpn + d]);
}

template <int DPN>
void residual_impl(const orc_problem* P, const double* u, const double* z, double* R) {
    std::vector<double> cfgWS, stateWS, ctrlWS, res, bres;
    const size_t ndof = (size_t)P->num_nodes * DPN;
    std::memset(R, 0, ndof * sizeof(double));
    workset_state<double, DPN>(P, u, stateWS);
    workset_control<double>(P, z, ctrlWS);
    workset_config(P, cfgWS);
    if (DPN == 3) elastic_evaluate(P, cfgWS, stateWS, ctrlWS, res);
    else thermal_evaluate(P, cfgWS, stateWS, ctrlWS, res);
    assemble_residual(P, DPN, res, R);
    // evaluate_boundary: elastic scale -1 (ElastostaticResidual.hpp:248-262),
    // thermal scale +1 (ThermostaticResidual.hpp:207-220).
    boundary_evaluate(P, DPN, DPN == 3 ? -1.0 : 1.0, bres);
    assemble_residual(P, DPN, bres, R);
}

// gradient_u (transpose=false) / gradient_u_T (transpose=true):
// src/elliptic/VectorFunction.hpp:431-615; scatter assemble_jacobian_fad
// (src/Assembly.hpp:1374-1399) with BlockMatrix[Transpose]EntryOrdinal.
template <int DPN>
void jacobian_impl(const orc_problem* P, const double* u, const double* z, double* values, bool transpose) {
    constexpr int W = 4 * DPN;
    std::vector<int> rowmap, colidx;
    build_graph(P->num_nodes, P->num_cells, P->conn, rowmap, colidx);  // rebuilt per call
    std::memset(values, 0, (size_t)rowmap[P->num_nodes] * DPN * DPN * sizeof(double));
    std::vector<double> cfgWS, ctrlWS;
    std::vector<Fad<W>> stateWS, res;
    workset_config(P, cfgWS);
    workset_state<Fad<W>, DPN>(P, u, stateWS);
    workset_control<double>(P, z, ctrlWS);
    if (DPN == 3) elastic_evaluate(P, cfgWS, stateWS, ctrlWS, res);
    else thermal_evaluate(P, cfgWS, stateWS, ctrlWS, res);
    const int C = P->num_cells;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c)
        for (int r = 0; r < W; ++r)
            for (int col = 0; col < W; ++col) {
                int in = r / DPN, id = r % DPN, jn = col / DPN, jd = col % DPN;
                int ni = P->conn[(size_t)c * 4 + in], nj = P->conn[(size_t)c * 4 + jn];
                long e;
                long off;
                if (!transpose) { e = entry_ordinal(rowmap.data(), colidx.data(), ni, nj); off = e * DPN * DPN + id * DPN + jd; }
                else { e = entry_ordinal(rowmap.data(), colidx.data(), nj, ni); off = e * DPN * DPN + jd * DPN + id; }
                atomic_add(&values[off], res[(size_t)c * W + r].d[col]);
            }
    // the boundary (uniform traction/flux) term does not depend on the state.
}

// gradient_z: src/elliptic/VectorFunction.hpp:617-710; block matrix with
// 1 x dpn blocks, row = control node, scatter assemble_transpose_jacobian
// (src/Assembly.hpp:1455-1487) through BlockMatrixEntryOrdinal<3,1,dpn>.
template <int DPN>
void gradz_impl(const orc_problem* P, const double* u, const double* z, double* values) {
    std::vector<int> rowmap, colidx;
    build_graph(P->num_nodes, P->num_cells, P->conn, rowmap, colidx);
    std::memset(values, 0, (size_t)rowmap[P->num_nodes] * DPN * sizeof(double));
    std::vector<double> cfgWS, stateWS;
    std::vector<Fad<4>> ctrlWS, res;
    workset_config(P, cfgWS);
    workset_control<Fad<4>>(P, z, ctrlWS);
    workset_state<double, DPN>(P, u, stateWS);
    if (DPN == 3) elastic_evaluate(P, cfgWS, stateWS, ctrlWS, res);
    else thermal_evaluate(P, cfgWS, stateWS, ctrlWS, res);
    const int C = P->num_cells;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c)
        for (int r = 0; r < 4 * DPN; ++r)
            for (int col = 0; col < 4; ++col) {
                // entry(cell, icellDof=col (control node), jcellDof=r)
                int ni = P->conn[(size_t)c * 4 + col];
                int jn = r / DPN, jd = r % DPN;
                int nj = P->conn[(size_t)c * 4 + jn];
                long e = entry_ordinal(rowmap.data(), colidx.data(), ni, nj);
                atomic_add(&values[e * DPN + jd], res[(size_t)c * 4 * DPN + r].d[col]);
            }
}

// ---------------------------------------------------------------------------
// Criteria.  InternalElasticEnergy, src/elliptic/InternalElasticEnergy.hpp:132-199
// (0.5 * weighted stress . strain * V, ScalarProduct src/ScalarProduct.hpp:22-41);
// InternalThermalEnergy, src/elliptic/InternalThermalEnergy.hpp:96-150
// (-1 * weighted flux . grad T * V).  Sum: local_result_sum (src/Assembly.hpp:20-33).
// ---------------------------------------------------------------------------
template <class S, class Z, class R, int DPN, class K = double>
void energy_evaluate(const orc_problem* P, const std::vector<K>& cfgWS, const std::vector<S>& stateWS,
                     const std::vector<Z>& ctrlWS, std::vector<R>& result) {
    using E = decltype(std::declval<S>() * std::declval<K>());
    const int C = P->num_cells;
    const double qw = 1.0 / 6.0;
    double Cm[6][6];
    isotropic_c(P->youngs, P->poisson, Cm);
    result.assign((size_t)C, R(0.0));
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c) {
        K cfg[4][3], g[4][3], vol;
        for (int n = 0; n < 4; ++n) for (int d = 0; d < 3; ++d) cfg[n][d] = cfgWS[(size_t)c * 12 + n * 3 + d];
        compute_gradient(cfg, g, vol);
        vol *= qw;
        Z rho = 0.0;
        for (int n = 0; n < 4; ++n) rho += ctrlWS[(size_t)c * 4 + n];
        rho = rho / 4.0;
        Z s = msimp(rho, P->penalty_exponent, P->penalty_min);
        if (DPN == 3) {
            const S* u = &stateWS[(size_t)c * 12];
            E eps[6];
            int v = 0;
            for (int d = 0; d < 3; ++d) {
                eps[v] = 0.0;
                for (int n = 0; n < 4; ++n) eps[v] += u[n * 3 + d] * g[n][d];
                ++v;
            }
            for (int dj = 2; dj >= 1; --dj)
                for (int di = dj - 1; di >= 0; --di) {
                    eps[v] = 0.0;
                    for (int n = 0; n < 4; ++n) eps[v] += (u[n * 3 + dj] * g[n][di] + u[n * 3 + di] * g[n][dj]);
                    ++v;
                }
            R sig[6];
            for (int i = 0; i < 6; ++i) {
                sig[i] = 0.0;
                for (int j = 0; j < 6; ++j) sig[i] += (eps[j] - 0.0) * Cm[i][j];
            }
            for (int i = 0; i < 6; ++i) sig[i] *= s;
            R inc = 0.0;
            for (int i = 0; i < 6; ++i) inc += sig[i] * eps[i];
            result[c] += 0.5 * inc * vol;
        } else {
            const S* T = &stateWS[(size_t)c * 4];
            E grad[3];
            for (int d = 0; d < 3; ++d) {
                grad[d] = 0.0;
                for (int n = 0; n < 4; ++n) grad[d] += T[n] * g[n][d];
            }
            R flux[3];
            for (int i = 0; i < 3; ++i) {
                flux[i] = 0.0;
                for (int j = 0; j < 3; ++j) flux[i] -= grad[j] * (i == j ? P->conductivity : 0.0);
            }
            for (int i = 0; i < 3; ++i) flux[i] *= s;
            R inc = 0.0;
            for (int i = 0; i < 3; ++i) inc += flux[i] * grad[i];
            result[c] += -1.0 * inc * vol;
        }
    }
}

template <int DPN>
double energy_value_impl(const orc_problem* P, const double* u, const double* z) {
    std::vector<double> cfgWS, stateWS, ctrlWS, res;
    workset_state<double, DPN>(P, u, stateWS);
    workset_control<double>(P, z, ctrlWS);
    workset_config(P, cfgWS);
    energy_evaluate<double, double, double, DPN>(P, cfgWS, stateWS, ctrlWS, res);
    double sum = 0.0;
    for (double r : res) sum += r;
    return sum;
}

// PhysicsScalarFunction::gradient_u, src/elliptic/PhysicsScalarFunction.hpp
// (state FAD, scatter through VectorEntryOrdinal).
template <int DPN>
void energy_gradu_impl(const orc_problem* P, const double* u, const double* z, double* g) {
    constexpr int W = 4 * DPN;
    std::vector<double> cfgWS, ctrlWS;
    std::vector<Fad<W>> stateWS, res;
    workset_state<Fad<W>, DPN>(P, u, stateWS);
    workset_control<double>(P, z, ctrlWS);
    workset_config(P, cfgWS);
    energy_evaluate<Fad<W>, double, Fad<W>, DPN>(P, cfgWS, stateWS, ctrlWS, res);
    std::memset(g, 0, (size_t)P->num_nodes * DPN * sizeof(double));
    for (int c = 0; c < P->num_cells; ++c)
        for (int n = 0; n < 4; ++n)
            for (int d = 0; d < DPN; ++d)
                g[(size_t)P->conn[(size_t)c * 4 + n] * DPN + d] += res[c].d[n * DPN + d];
}

// PhysicsScalarFunction::gradient_z, src/elliptic/PhysicsScalarFunction.hpp:431-494;
// scatter assemble_scalar_gradient_fad (src/Assembly.hpp:592-612).
template <int DPN>
void energy_gradz_impl(const orc_problem* P, const double* u, const double* z, double* g) {
    std::vector<double> cfgWS, stateWS;
    std::vector<Fad<4>> ctrlWS, res;
    workset_state<double, DPN>(P, u, stateWS);
    workset_control<Fad<4>>(P, z, ctrlWS);
    workset_config(P, cfgWS);
    energy_evaluate<double, Fad<4>, Fad<4>, DPN>(P, cfgWS, stateWS, ctrlWS, res);
    std::memset(g, 0, (size_t)P->num_nodes * sizeof(double));
    for (int c = 0; c < P->num_cells; ++c)
        for (int n = 0; n < 4; ++n) g[P->conn[(size_t)c * 4 + n]] += res[c].d[n];
}

// ---------------------------------------------------------------------------
// gradient_x (GradientX evaluation type: config FAD of width 12, state and
// control plain).  VectorFunction::gradient_x, src/elliptic/VectorFunction.hpp:
// 337-420: volume and boundary worksets scattered by assemble_transpose_jacobian
// (src/Assembly.hpp:1457-1521) through BlockMatrixEntryOrdinal<3, 3, dpn>
// (src/ImplicitFunctors.hpp:730-766): block (config node, state node) holds
// dR_{state node, j} / dx_{config node, i} at offset i*dpn + j.
// ---------------------------------------------------------------------------

// Boundary workset of the uniform natural BC with config FAD (the same steps
// as boundary_evaluate, the face Jacobian taken from the config workset).
void boundary_evaluate_x(const orc_problem* P, int dpn, double scale, const std::vector<Fad<12>>& cfgWS,
                         std::vector<Fad<12>>& result) {
    const int C = P->num_cells;
    result.assign((size_t)C * 4 * dpn, Fad<12>(0.0));
    const double cubw = 0.5, basis = 1.0 / 3.0;
    for (int f = 0; f < P->num_faces; ++f) {
        int c = P->face_cells[f];
        int ln[3] = {P->face_lnodes[f * 3], P->face_lnodes[f * 3 + 1], P->face_lnodes[f * 3 + 2]};
        Fad<12> J[2][3];
        for (int a = 0; a < 2; ++a)
            for (int d = 0; d < 3; ++d)
                J[a][d] = cfgWS[(size_t)c * 12 + ln[a] * 3 + d] - cfgWS[(size_t)c * 12 + ln[2] * 3 + d];
        Fad<12> j23 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
        Fad<12> j31 = J[0][2] * J[1][0] - J[0][0] * J[1][2];
        Fad<12> j12 = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        Fad<12> area = (scale * cubw) * fsqrt(j23 * j23 + j31 * j31 + j12 * j12);
        for (int n = 0; n < 3; ++n)
            for (int d = 0; d < dpn; ++d)
                result[(size_t)c * 4 * dpn + ln[n] * dpn + d] += basis * P->flux[d] * area;
    }
}

template <int DPN>
void gradx_scatter(const orc_problem* P, const int* rowmap, const int* colidx, const std::vector<Fad<12>>& res,
                   double* values) {
    const int C = P->num_cells;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c)
        for (int r = 0; r < 4 * DPN; ++r)
            for (int col = 0; col < 12; ++col) {
                // entry(cell, icellDof = col (config dof), jcellDof = r (residual dof))
                const int ni = P->conn[(size_t)c * 4 + col / 3], id = col % 3;
                const int nj = P->conn[(size_t)c * 4 + r / DPN], jd = r % DPN;
                const long e = entry_ordinal(rowmap, colidx, ni, nj);
                atomic_add(&values[e * 3 * DPN + id * DPN + jd], res[(size_t)c * 4 * DPN + r].d[col]);
            }
}

template <int DPN>
void gradx_impl(const orc_problem* P, const double* u, const double* z, double* values) {
    std::vector<int> rowmap, colidx;
    build_graph(P->num_nodes, P->num_cells, P->conn, rowmap, colidx);
    std::memset(values, 0, (size_t)rowmap[P->num_nodes] * 3 * DPN * sizeof(double));
    std::vector<double> stateWS, ctrlWS;
    std::vector<Fad<12>> cfgWS, res, bres;
    workset_config_fad(P, cfgWS);
    workset_state<double, DPN>(P, u, stateWS);
    workset_control<double>(P, z, ctrlWS);
    if (DPN == 3) elastic_evaluate(P, cfgWS, stateWS, ctrlWS, res);
    else thermal_evaluate(P, cfgWS, stateWS, ctrlWS, res);
    gradx_scatter<DPN>(P, rowmap.data(), colidx.data(), res, values);
    boundary_evaluate_x(P, DPN, DPN == 3 ? -1.0 : 1.0, cfgWS, bres);
    gradx_scatter<DPN>(P, rowmap.data(), colidx.data(), bres, values);
}

// PhysicsScalarFunction::gradient_x, src/elliptic/PhysicsScalarFunction.hpp:293-355;
// scatter assemble_vector_gradient_fad (VectorEntryOrdinal: node*3 + dim).
template <int DPN>
void energy_gradx_impl(const orc_problem* P, const double* u, const double* z, double* g) {
    std::vector<double> stateWS, ctrlWS;
    std::vector<Fad<12>> cfgWS, res;
    workset_state<double, DPN>(P, u, stateWS);
    workset_control<double>(P, z, ctrlWS);
    workset_config_fad(P, cfgWS);
    energy_evaluate<double, double, Fad<12>, DPN, Fad<12>>(P, cfgWS, stateWS, ctrlWS, res);
    std::memset(g, 0, (size_t)P->num_nodes * 3 * sizeof(double));
    for (int c = 0; c < P->num_cells; ++c)
        for (int n = 0; n < 4; ++n)
            for (int d = 0; d < 3; ++d) g[(size_t)P->conn[(size_t)c * 4 + n] * 3 + d] += res[c].d[n * 3 + d];
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI of the oracle (loaded with ctypes by the tests and the CPU baseline).
// ---------------------------------------------------------------------------
extern "C" {

int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

// Structured box split into 6 Kuhn tets per hex around the v000->v111
// diagonal.  Node id = ix*(ny+1)*(nz+1) + iy*(nz+1) + iz; hex id =
// (ix*ny + iy)*nz + iz; tets of a hex are contiguous.  Recovered from the
// DerivativeTests gradient/strain goldens (unit_tests/DerivativeTests.cpp:155-189).
void orc_box_mesh(int nx, int ny, int nz, double lx, double ly, double lz, double* coords, int* conn) {
    const int NY = ny + 1, NZ = nz + 1;
    for (int ix = 0; ix <= nx; ++ix)
        for (int iy = 0; iy <= ny; ++iy)
            for (int iz = 0; iz <= nz; ++iz) {
                long n = ((long)ix * NY + iy) * NZ + iz;
                coords[n * 3 + 0] = lx * ix / nx;
                coords[n * 3 + 1] = ly * iy / ny;
                coords[n * 3 + 2] = lz * iz / nz;
            }
    // offsets as (dx,dy,dz)
    static const int pat[6][4][3] = {
        {{0, 0, 0}, {1, 1, 0}, {0, 1, 0}, {1, 1, 1}},
        {{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {1, 1, 1}},
        {{0, 0, 0}, {0, 1, 1}, {0, 0, 1}, {1, 1, 1}},
        {{0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}},
        {{0, 0, 0}, {1, 0, 1}, {1, 0, 0}, {1, 1, 1}},
        {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {1, 1, 1}},
    };
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz) {
                long h = ((long)ix * ny + iy) * nz + iz;
                for (int t = 0; t < 6; ++t)
                    for (int a = 0; a < 4; ++a) {
                        long n = ((long)(ix + pat[t][a][0]) * NY + (iy + pat[t][a][1])) * NZ + (iz + pat[t][a][2]);
                        conn[(h * 6 + t) * 4 + a] = (int)n;
                    }
            }
}

long orc_graph_nnz(int N, int C, const int* conn) {
    std::vector<int> r, c;
    build_graph(N, C, conn, r, c);
    return r[N];
}

void orc_graph_fill(int N, int C, const int* conn, int* rowmap, int* colidx) {
    std::vector<int> r, c;
    build_graph(N, C, conn, r, c);
    std::memcpy(rowmap, r.data(), (N + 1) * sizeof(int));
    std::memcpy(colidx, c.data(), c.size() * sizeof(int));
}

// Per-cell functor chain of the DerivativeTests 3D test
// (unit_tests/DerivativeTests.cpp:46-248): gradients, strain, stress (no
// weighting), and StressDivergence result per cell.
void orc_elastic_cell_chain(const orc_problem* P, const double* u, double* grads, double* strain, double* stress,
                            double* result) {
    std::vector<double> cfgWS, stateWS;
    workset_config(P, cfgWS);
    workset_state<double, 3>(P, u, stateWS);
    double Cm[6][6];
    isotropic_c(P->youngs, P->poisson, Cm);
    for (int c = 0; c < P->num_cells; ++c) {
        double cfg[4][3], g[4][3], vol;
        for (int n = 0; n < 4; ++n) for (int d = 0; d < 3; ++d) cfg[n][d] = cfgWS[(size_t)c * 12 + n * 3 + d];
        compute_gradient(cfg, g, vol);
        vol *= 1.0 / 6.0;
        for (int n = 0; n < 4; ++n) for (int d = 0; d < 3; ++d) grads[(size_t)c * 12 + n * 3 + d] = g[n][d];
        const double* uu = &stateWS[(size_t)c * 12];
        double eps[6];
        int v = 0;
        for (int d = 0; d < 3; ++d) { eps[v] = 0; for (int n = 0; n < 4; ++n) eps[v] += uu[n * 3 + d] * g[n][d]; ++v; }
        for (int dj = 2; dj >= 1; --dj)
            for (int di = dj - 1; di >= 0; --di) {
                eps[v] = 0;
                for (int n = 0; n < 4; ++n) eps[v] += uu[n * 3 + dj] * g[n][di] + uu[n * 3 + di] * g[n][dj];
                ++v;
            }
        double sig[6];
        for (int i = 0; i < 6; ++i) { sig[i] = 0; for (int j = 0; j < 6; ++j) sig[i] += eps[j] * Cm[i][j]; }
        for (int i = 0; i < 6; ++i) { strain[(size_t)c * 6 + i] = eps[i]; stress[(size_t)c * 6 + i] = sig[i]; }
        double* out = &result[(size_t)c * 12];
        for (int i = 0; i < 12; ++i) out[i] = 0.0;
        for (int di = 0; di < 3; ++di)
            for (int n = 0; n < 4; ++n)
                for (int dj = 0; dj < 3; ++dj) out[n * 3 + di] += vol * sig[kVoigt[di][dj]] * g[n][dj];
    }
}

void orc_elastic_value(const orc_problem* P, const double* u, const double* z, double* R) { residual_impl<3>(P, u, z, R); }
void orc_thermal_value(const orc_problem* P, const double* u, const double* z, double* R) { residual_impl<1>(P, u, z, R); }
void orc_elastic_gradient_u(const orc_problem* P, const double* u, const double* z, double* v) { jacobian_impl<3>(P, u, z, v, false); }
void orc_thermal_gradient_u(const orc_problem* P, const double* u, const double* z, double* v) { jacobian_impl<1>(P, u, z, v, false); }
void orc_elastic_gradient_u_T(const orc_problem* P, const double* u, const double* z, double* v) { jacobian_impl<3>(P, u, z, v, true); }
void orc_thermal_gradient_u_T(const orc_problem* P, const double* u, const double* z, double* v) { jacobian_impl<1>(P, u, z, v, true); }
void orc_elastic_gradient_z(const orc_problem* P, const double* u, const double* z, double* v) { gradz_impl<3>(P, u, z, v); }
void orc_thermal_gradient_z(const orc_problem* P, const double* u, const double* z, double* v) { gradz_impl<1>(P, u, z, v); }
void orc_elastic_gradient_x(const orc_problem* P, const double* u, const double* z, double* v) { gradx_impl<3>(P, u, z, v); }
void orc_thermal_gradient_x(const orc_problem* P, const double* u, const double* z, double* v) { gradx_impl<1>(P, u, z, v); }

double orc_elastic_energy_value(const orc_problem* P, const double* u, const double* z) { return energy_value_impl<3>(P, u, z); }
double orc_thermal_energy_value(const orc_problem* P, const double* u, const double* z) { return energy_value_impl<1>(P, u, z); }
void orc_elastic_energy_gradient_u(const orc_problem* P, const double* u, const double* z, double* g) { energy_gradu_impl<3>(P, u, z, g); }
void orc_thermal_energy_gradient_u(const orc_problem* P, const double* u, const double* z, double* g) { energy_gradu_impl<1>(P, u, z, g); }
void orc_elastic_energy_gradient_z(const orc_problem* P, const double* u, const double* z, double* g) { energy_gradz_impl<3>(P, u, z, g); }
void orc_thermal_energy_gradient_z(const orc_problem* P, const double* u, const double* z, double* g) { energy_gradz_impl<1>(P, u, z, g); }
void orc_elastic_energy_gradient_x(const orc_problem* P, const double* u, const double* z, double* g) { energy_gradx_impl<3>(P, u, z, g); }
void orc_thermal_energy_gradient_x(const orc_problem* P, const double* u, const double* z, double* g) { energy_gradx_impl<1>(P, u, z, g); }

// applyBlockConstraints, src/ApplyConstraints.hpp:40-112 (symmetric
// elimination: zero row and column, unit diagonal, RHS lift b -= A g, b_bc = g).
void orc_apply_block_constraints(int dpn, int N, const int* rowmap, const int* colidx, double* values, double* rhs,
                                 int nbc, const int* bc_dofs, const double* bc_vals, double scale) {
    (void)N;
    for (int b = 0; b < nbc; ++b) {
        int rdof = bc_dofs[b];
        double val = scale * bc_vals[b];
        int rnode = rdof / dpn, lr = rdof % dpn;
        for (int off = rowmap[rnode]; off < rowmap[rnode + 1]; ++off)
            for (int lc = 0; lc < dpn; ++lc) {
                int cnode = colidx[off];
                long ent = (long)dpn * dpn * off + dpn * lr + lc;
                int cdof = dpn * cnode + lc;
                if (cdof == rdof) {
                    values[ent] = 1.0;
                } else {
                    rhs[cdof] -= values[ent] * val;
                    values[ent] = 0.0;
                    for (int o2 = rowmap[cnode]; o2 < rowmap[cnode + 1]; ++o2) {
                        int crn = colidx[o2];
                        long e2 = (long)dpn * dpn * o2 + dpn * lc + lr;
                        if (dpn * crn + lr == rdof) values[e2] = 0.0;
                    }
                }
            }
    }
    for (int b = 0; b < nbc; ++b) rhs[bc_dofs[b]] = scale * bc_vals[b];
}

// Block CRS matrix-vector product y = A x (reference: src/PlatoMathHelpers.hpp:79-136).
void orc_block_spmv(int dpn_r, int dpn_c, int N, const int* rowmap, const int* colidx, const double* values,
                    const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i)
        for (int a = 0; a < dpn_r; ++a) {
            double s = 0.0;
            for (int e = rowmap[i]; e < rowmap[i + 1]; ++e)
                for (int b = 0; b < dpn_c; ++b) s += values[(long)e * dpn_r * dpn_c + a * dpn_c + b] * x[(long)colidx[e] * dpn_c + b];
            y[(long)i * dpn_r + a] = s;
        }
}

}  // extern "C"
