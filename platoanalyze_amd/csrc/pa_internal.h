// Internal types shared by the host code and the HIP kernels of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/platoanalyze_amd.h"

namespace pa {

// Error carried to the C ABI as (status, message).
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

// A 'Function' string compiled once (pa_expr.cpp): numbers, x y z t, pi,
// + - * / ^, parentheses, the common functions, leading 'name = expr;'.
class Expr {
  public:
    struct Inst {
        int op, i;
        double v;
    };
    explicit Expr(const std::string& text);
    double eval(double x, double y, double z, double t) const;
    const std::string& text() const { return text_; }

  private:
    std::string text_;
    std::vector<Inst> prog_;
    int nslots_ = 0, depth_ = 0;
};

#define PA_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            throw ::pa::Error(PA_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define PA_REQUIRE(cond, code, msg)                         \
    do {                                                    \
        if (!(cond)) throw ::pa::Error((code), (msg));      \
    } while (0)

// Runs f, converting exceptions to a status code and the thread's last error.
template <class F>
int guarded(F&& f) {
    try {
        f();
        return PA_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return PA_ERR_STATE;
    } catch (...) {
        set_last_error("unknown error");
        return PA_ERR_STATE;
    }
}

// Owning device buffer.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        if (count) PA_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void upload(const T* h, size_t count) {
        alloc(count);
        if (count) PA_HIP(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
    }
};

// Side set: faces given as (cell, three local node ordinals), plus the
// derived boundary-node -> face lists used by the node-ordered load kernel.
struct SideSet {
    std::vector<int> cells, lnodes;    // host, F and F*3
    DevBuf<int> d_cells, d_lnodes;     // device copies
    DevBuf<int> d_bnodes;              // distinct nodes touched, sorted
    DevBuf<int> d_bface_off;           // nb+1
    DevBuf<int> d_bface;               // entries: face*4 + position in face (0..2)
    int num_bnodes = 0;
};

// Cell tiles (shared by the schedule builder and the assembly kernels).
constexpr int kTileX = 2, kTileY = 4, kTileZ = 4;  // rows per tile along the node grid
constexpr int kTileHalo = (kTileX + 2) * (kTileY + 2) * (kTileZ + 2);  // nodes a tile's cells reach
constexpr int kTileStride = 13;     // doubles per cell in LDS: sqrt(w) g_n for n = 0..3, one pad
constexpr int kTileCellsMax = 512;  // cells per tile the kernels' LDS is sized for
constexpr int kTilePatternsMax = 4096;
constexpr int kTileUnitEntries = 8;   // entry slots per work unit (contributions to one block)
constexpr int kTileDiagParts = 4;     // units per diagonal block, summed across adjacent lanes
constexpr int kTileThreads = 256;     // threads per tile workgroup
constexpr int kTileUnitsMax = 3 * kTileThreads;  // work units per tile: three per thread

}  // namespace pa

// Mesh with the assembly graph and the node -> cell incidence used by the
// node-ordered assembly kernels.
struct pa_mesh {
    int N = 0, C = 0;
    long long nnzb = 0;
    int max_row_len = 0;
    int max_cells_per_node = 0;
    std::vector<int> h_rowmap, h_colidx;
    pa::DevBuf<double> d_coords;   // N*3 as given
    pa::DevBuf<double> d_coords4;  // N*4 (x, y, z, 0) for 32-byte gathers
    pa::DevBuf<int> d_conn;        // C*4
    pa::DevBuf<int> d_rowmap;      // N+1
    pa::DevBuf<int> d_colidx;      // nnzb
    pa::DevBuf<int> d_n2c_off;     // N+1
    pa::DevBuf<int2> d_pairs;      // 4C: x = cell<<2 | local index of the row node, y = 4 packed uint8 column slots
    // Row patterns: rows whose (incident-cell local index, column slots)
    // sequences coincide share one gather/contraction schedule.
    bool has_patterns = false;      // every pattern fits the pattern kernels' limits
    int num_patterns = 0;
    pa::DevBuf<int> d_pcell;        // 4C: cell id of each (row, incident cell) pair
    pa::DevBuf<int> d_row_pat;      // N: pattern id of each row
    pa::DevBuf<int4> d_pat_info;    // 2 per pattern: {npairs, rowlen, nitems, diag_slot}, {a_off, ent_off, item_off, first_diag_item}
    pa::DevBuf<unsigned char> d_pat_a;    // per pattern pair: local index of the row node in the cell
    pa::DevBuf<int4> d_pat_delta;         // per pattern pair: node id - row id of the cell's four nodes
    pa::DevBuf<uint4> d_pat_item2;        // work items with their (<= 8) entries: slot | count << 8, entry bytes
    pa::DevBuf<unsigned char> d_pat_ent;  // contributions (pair << 2 | local node b), grouped by column slot
    pa::DevBuf<unsigned> d_pat_item;      // work items: slot | start << 8 | count << 20
    // Cell tiles: rows grouped into blocks of the lexicographic node grid; a
    // block's incident cells are evaluated once each and its rows summed from
    // LDS.  Blocks with the same structure relative to their first row share
    // one schedule (tile pattern).
    bool has_tiles = false;
    int num_tiles = 0, num_tile_patterns = 0, tile_cells_max = 0, tile_rows_max = 0;
    int tile_unit_entries = 0;        // entries per work unit the kernels evaluate (6 or 8)
    int tile_blocks_max = 0;          // output blocks of a tile (its staging size)
    int grid_dims[3] = {0, 0, 0};     // detected node grid (x, y, z node counts)
    // Kuhn grid: nodes a lexicographic grid (z fastest) and cells exactly the
    // six tets of every hex around its v0 -> v7 diagonal; the streaming grid
    // kernels apply
    bool has_grid = false;
    int kgrid[3] = {0, 0, 0};
    int side_set_version = 0;         // bumped by every side-set definition
    pa::DevBuf<int4> d_tile;          // 3 per tile: {first row, row_off, cell_off, unit_off},
                                      // {rows, cells, units, runs}, {run_off, blocks, 0, 0}
    pa::DevBuf<int4> d_tp_row;        // per pattern row: {row - first row, incident cells, pair offset,
                                      // staging offset}
    pa::DevBuf<int4> d_tp_run;        // per pattern output run (rows with consecutive ids): {first row index,
                                      // staging offset, blocks, 0}
    pa::DevBuf<int4> d_tp_cell;       // per pattern cell: its four node ids - first row
    pa::DevBuf<uint2> d_tp_cslot;     // per pattern cell: residual pair slot (row * 32 + k) of each node
    pa::DevBuf<unsigned> d_tp_hslot;  // per tile cell: halo slots of its 4 nodes
                                      // (16 bits each, 0xffff for a node that is not a tile row)
    pa::DevBuf<unsigned short> d_tp_pair;  // per (row, incident cell): tile cell << 2 | local index (host checks)
    pa::DevBuf<uint2> d_tp_unit;      // (row, column slot) work units: {row | slot << 8 | mirror row << 16 |
                                      // mirror slot << 24, entry offset | count << 22 | diagonal << 26 |
                                      // part << 27 | mirror << 29 | wave shape << 30}
    pa::DevBuf<unsigned> d_tp_ent;    // unit entries (kTileUnitEntries per unit): LDS offsets of g_a and g_b
    std::map<std::string, pa::SideSet> side_sets;
    std::map<std::string, std::vector<int>> node_sets;
};

struct NaturalBC {
    std::string side_set;
    double values[3];
};

struct pa_vector_function {
    pa_mesh* mesh = nullptr;
    int physics = PA_MECHANICS;
    int dpn = 3;
    double material[2] = {0, 0};
    double penalty[2] = {3.0, 1e-9};
    std::vector<NaturalBC> natural_bcs;
    bool has_body = false;
    int body_dof = 0;
    double body_value = 0.0;
    // body loads given per cell: the function's value at each cell's
    // centroid (the one-point cubature of Tet4), for any dof
    struct CellLoad {
        int dof = 0;
        std::string function;
        pa::DevBuf<double> f;  // C values
    };
    std::vector<std::unique_ptr<CellLoad>> cell_loads;
    // the natural BCs' nodal load (it depends on the coordinates and the
    // load values only), built on first use: valid for fext_nbc loads and
    // the mesh's side sets at fext_version
    mutable pa::DevBuf<double> fext;
    mutable size_t fext_nbc = 0;
    mutable int fext_version = -1;
};

namespace pa {
struct Material;
// Streaming kernels on Kuhn grids (pa_grid.hip); false when the mesh is not one.
// Residual rows R = scale * (assembled + body source) + fext (fext may be null).
bool grid_residual(const pa_mesh* m, int dpn, const Material& mt, const double* u, const double* z, double body,
                   double scale, const double* fext, double* R, hipStream_t st);
// Thermal (1 x 1 block) Jacobian values in the mesh's CRS order.
template <class VT>
bool grid_jacobian1(const pa_mesh* m, const Material& mt, const double* z, VT* vals, hipStream_t st);
// Elastic (3 x 3 block) Jacobian values in the mesh's CRS order.
template <class VT>
bool grid_jacobian3(const pa_mesh* m, const Material& mt, const double* z, VT* vals, hipStream_t st);
// The elastic residual R = K u + fext (fext may be null) and the Jacobian
// values in one pass.
template <class VT>
bool grid_value_jacobian3(const pa_mesh* m, const Material& mt, const double* u, const double* z,
                          const double* fext, double* R, VT* vals, hipStream_t st);
// Criterion gradient_x (N*3 doubles), scaled (the self-adjoint total
// gradient wrt the coordinates is -1 times it).
void criterion_gradient_x(const pa_scalar_function* sf, const double* u, const double* z, double scale, double* g,
                          hipStream_t st);
// Fused self-adjoint criterion gradient (lambda = -u), defined with the
// assembly kernels; used by the problem driver.
void self_adjoint_gradient(const pa_vector_function* vf, const pa_scalar_function* sf, const double* u,
                           const double* z, double* out, hipStream_t st);
// applyBlockConstraints in a fixed summation order, with caller-owned work
// vectors: dense (N * dpn doubles) and touched (N bytes).
void apply_block_constraints(const pa_mesh* m, int dpn, void* vals, int precision, double* rhs, int nbc,
                             const int* dofs, const double* bcv, double scale, double* dense, unsigned char* touched,
                             hipStream_t st);
// z = M^-1 r by one multigrid V-cycle (pa_multigrid.hip)
void multigrid_apply(pa_multigrid* mg, const double* r, double* z, hipStream_t st);
// the operators given to pa_multigrid_setup are symmetric positive definite
// (the elliptic problems): the fine level keeps half of its block stencil and
// the default smoother is Chebyshev
void multigrid_set_elliptic(pa_multigrid* mg);
}  // namespace pa

struct pa_scalar_function {
    pa_mesh* mesh = nullptr;
    int type = PA_INTERNAL_ELASTIC_ENERGY;
    int dpn = 3;
    double material[2] = {0, 0};
    double penalty[2] = {3.0, 1e-9};
    pa::DevBuf<double> d_partials;  // per-block partial sums of the value reduction
};
