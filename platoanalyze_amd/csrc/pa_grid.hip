// Streaming assembly on Kuhn-split hex grids (gfx950): the residual of SIMP
// elastostatics and steady thermal diffusion, and the thermal Jacobian.
//
// Applies to a mesh whose nodes are a lexicographic grid (z fastest) and
// whose cells are exactly the 6 tets of every hex around its v0 -> v7
// diagonal (the box meshes; checked once per mesh on the host).  Hex corner
// v = dx << 2 | dy << 1 | dz; the tets are (0, p, q, 7) with (p, q) = (6, 2),
// (2, 3), (3, 1), (1, 5), (5, 4), (4, 6), the vertex order of the box mesh.
//
// Work decomposition.  A workgroup of W waves owns node rows (i, j, k) in a
// block of lines: lane l of wave w sits on hex column (j, k) = (hj0 + w,
// hk0 + l), and the workgroup walks the x layers of hexes.  At layer h every
// lane evaluates its hex once (six tets, one reciprocal each) and sends the
// hex's contributions to the rows of its eight corners: corners with
// (dy, dz) = (0, 0) stay in the lane, (0, 1) go to lane l + 1 by a shuffle,
// (1, 1) to lane l + 1 by a shuffle and on with the (1, 0) corners to wave
// w + 1 through a double-buffered LDS slot.  Corners at dx = 0 complete the
// rows of node layer h; those at dx = 1 are carried in registers to layer
// h + 1.  Wave 0 (past the first line block) and lane 0 (past the first z
// block) only compute the halo hexes of their neighbours, so each hex is
// evaluated about (W / (W - 1)) (64 / 63) times, and every row is summed in
// registers in a fixed order: no atomics, deterministic.  Rows of one line
// are consecutive node ids, so a wave's rows form one contiguous CRS range,
// which the Jacobian stages in LDS and stores with lane-contiguous stores.
//
// Per tet (nodes x_0..x_3 in the cell's order, J = [x_a - x_3]), with the
// cofactor vectors c_a (c_0 = e_1 x e_2, c_1 = e_2 x e_0, c_2 = e_0 x e_1,
// e_a = x_a - x_3; c_3 = -(c_0 + c_1 + c_2)), grad N_a = c_a / det J and
// V = |det J| / 6:
//   thermal Jacobian   K_ab = s(rho) k V grad N_a . grad N_b
//                            = s k / (6 |det|) c_a . c_b
//   thermal residual   R_a  = s k / (6 |det|) sum_b T_b c_b . c_a
//                             - V f rho / 4                (body source)
//   elastic residual   R_a  = B(c_a)^T [s / (6 |det|) C eps(u; c)]
// the reference's chain (src/ImplicitFunctors.hpp:169-221, src/Strain.hpp:56-95,
// src/LinearStress.hpp:102-120, src/Simp.hpp:130-134, src/StressDivergence.hpp:70-90,
// src/ScalarGrad.hpp:58-74, src/ThermalFlux.hpp:42-52, src/FluxDivergence.hpp:31-48,
// src/BodyLoads.hpp:103-121) with one division per tet.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <utility>

#include "pa_cell.h"
#include "pa_internal.h"

namespace pa {
namespace {

constexpr int kGridWaves = 4;  // waves per workgroup (one halo line)

// Matrix values are written once and never read back by the kernel that
// writes them: non-temporal stores (measured at 128x64x64: FP64
// value_gradient_u 0.2015 -> 0.1817 ms, gradient_u 0.1795 -> 0.165 ms).
template <class T>
__device__ inline void store_nt(T* p, T v) {
    __builtin_nontemporal_store(v, p);
}
typedef double nt_double2 __attribute__((ext_vector_type(2)));
__device__ inline void store_nt2(double* p, double a, double b) {
    const nt_double2 v = {a, b};
    __builtin_nontemporal_store(v, reinterpret_cast<nt_double2*>(p));
}

// Compile-time loop: f(std::integral_constant<int, i>) for i = 0..N-1, so
// table lookups on i are constant expressions and the hex value arrays stay
// in registers.
template <class F, int... Is>
__device__ inline void sfor_impl(F&& f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__device__ inline void sfor(F&& f) {
    sfor_impl(f, std::make_integer_sequence<int, N>{});
}
#define PA_CI(x) decltype(x)::value

// The 15 neighbour offsets of a node of the Kuhn grid, in the order of their
// node ids (z fastest, so the order of the CRS columns of a row).
__host__ __device__ constexpr int nb_di(int s) { return s < 4 ? -1 : (s < 11 ? 0 : 1); }
__host__ __device__ constexpr int nb_dj(int s) {
    constexpr int t[15] = {-1, -1, 0, 0, -1, -1, 0, 0, 0, 1, 1, 0, 0, 1, 1};
    return t[s];
}
__host__ __device__ constexpr int nb_dk(int s) {
    constexpr int t[15] = {-1, 0, -1, 0, -1, 0, -1, 0, 1, 0, 1, 0, 1, 0, 1};
    return t[s];
}
constexpr int kDiagSlot = 7;

// Edges of the split (12 hex edges, 6 face diagonals, the body diagonal) as
// corner pairs a < b; index in the hex value array is 8 + edge.
__host__ __device__ constexpr int edge_of(int a, int b) {
    constexpr int ea[19] = {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 4, 4, 4, 5, 6};
    constexpr int eb[19] = {1, 2, 3, 4, 5, 6, 7, 3, 5, 7, 3, 6, 7, 7, 5, 6, 7, 7, 7};
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    for (int e = 0; e < 19; ++e)
        if (ea[e] == lo && eb[e] == hi) return e;
    return -1;
}

// Index in the hex's 27 values (8 diagonals, 19 edges) of what the hex adds
// to slot s of the row of its corner v, or -1.
__host__ __device__ constexpr int jac_entry(int v, int s) {
    if (s == kDiagSlot) return v;
    const int x = (v >> 2) + nb_di(s), y = ((v >> 1) & 1) + nb_dj(s), z = (v & 1) + nb_dk(s);
    if (x < 0 || x > 1 || y < 0 || y > 1 || z < 0 || z > 1) return -1;
    const int e = edge_of(v, x << 2 | y << 1 | z);
    return e < 0 ? -1 : 8 + e;
}

// Kuhn tets (0, p, q, 7)
__host__ __device__ constexpr int tet_p(int t) {
    constexpr int p[6] = {6, 2, 3, 1, 5, 4};
    return p[t];
}
__host__ __device__ constexpr int tet_q(int t) {
    constexpr int q[6] = {2, 3, 1, 5, 4, 6};
    return q[t];
}

struct GridArgs {
    int d0, d1, d2;  // node grid (x, y, z counts)
    int nj, nk;      // line blocks and z blocks
    int lx;          // node layers per x block
    const double4* __restrict__ x4;
    const double* __restrict__ z;
};

struct Corner {
    double x, y, z, r;
};

__device__ inline void cross(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ inline double dot3(const double a[3], const double b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// Edge vectors and cross products shared by the six tets of a hex:
// e[p] = x_p - x_7 (p = 0..6), F[p] = e[p] x e[0] (p = 1..6).  Tet (0, p, q, 7)
// (local nodes 0..3 = corners 0, p, q, 7) has det J = e[0] . (e[p] x e[q]) and
// cofactor vectors c_0 = e[p] x e[q], c_1 = e[q] x e[0] = F[q],
// c_2 = e[0] x e[p] = -F[p], c_3 = -(c_0 + c_1 + c_2).
struct HexGeom {
    double e[7][3];
    double F[7][3];  // 1..6
};

template <bool WITH_F = true>
__device__ inline void hex_geom(const Corner X[8], HexGeom& h) {
#pragma unroll
    for (int p = 0; p < 7; ++p) {
        h.e[p][0] = X[p].x - X[7].x;
        h.e[p][1] = X[p].y - X[7].y;
        h.e[p][2] = X[p].z - X[7].z;
    }
    if constexpr (WITH_F) {
#pragma unroll
        for (int p = 1; p < 7; ++p) cross(h.e[p], h.e[0], h.F[p]);
    }
}

// Mean nodal density of tet (0, p, q, 7), in the cell's node order.
__device__ inline double tet_rho(const Corner X[8], int p, int q) {
    return (((X[0].r + X[p].r) + X[q].r) + X[7].r) * 0.25;
}

// 1/x: the hardware estimate refined by two Newton steps (each one squares
// the relative error), well inside an ulp-level bar.
__device__ inline double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// ------------------------------------------------------------------ blocks
// Line block jc: lines [J0, J1) owned, hex lines from hj0; z block kc: rows
// [K0, K1), hexes from hk0; the last z block's last lane also owns the top
// row when it lies one past the lanes (top).
struct Block {
    int J0, J1, hj0, K0, K1, hk0, X0, X1;
    bool top;
};

__device__ inline Block grid_block(const GridArgs& g, int W) {
    const int b = xcd_block();
    const int kc = b % g.nk, jc = (b / g.nk) % g.nj, xc = b / (g.nk * g.nj);
    Block B;
    B.J0 = jc == 0 ? 0 : W + (W - 1) * (jc - 1);
    B.J1 = min(B.J0 + (jc == 0 ? W : W - 1), g.d1);
    B.hj0 = jc == 0 ? 0 : B.J0 - 1;
    B.K0 = kc == 0 ? 0 : 64 + 63 * (kc - 1);
    B.K1 = min(B.K0 + (kc == 0 ? 64 : 63), g.d2);
    B.hk0 = kc == 0 ? 0 : B.K0 - 1;
    B.top = B.K1 == g.d2 - 1 && B.hk0 + 64 == g.d2 - 1;
    if (B.top) B.K1 = g.d2;
    B.X0 = xc * g.lx;
    B.X1 = min(B.X0 + g.lx, g.d0);
    return B;
}

// 15-bit mask of the neighbours of node (i, j, k) inside the grid.
__device__ inline unsigned row_exists(const GridArgs& g, int i, int j, int k) {
    unsigned m = 0x7fffu;
    if (i == 0) m &= ~0x000fu;
    if (i == g.d0 - 1) m &= ~0x7800u;
    if (j == 0) m &= ~0x0033u;
    if (j == g.d1 - 1) m &= ~0x6600u;
    if (k == 0) m &= ~0x0055u;
    if (k == g.d2 - 1) m &= ~0x5500u;
    return m;
}

__device__ inline void load_corner(const GridArgs& g, long node, Corner& c) {
    const double4 p = g.x4[node];
    c.x = p.x;
    c.y = p.y;
    c.z = p.z;
    c.r = g.z[node];
}

// The four corners (dy, dz) of hex column (j, k) at node layer i.
__device__ inline void load_layer(const GridArgs& g, int i, int j, int k, Corner c[4]) {
    const long n = ((long)i * g.d1 + j) * g.d2 + k;
    load_corner(g, n, c[0]);
    load_corner(g, n + 1, c[1]);
    load_corner(g, n + g.d2, c[2]);
    load_corner(g, n + g.d2 + 1, c[3]);
}

// The value of lane l - 1 (0 on lane 0): a DPP wave shift right by one lane,
// whose out-of-range source lane reads as the zero `old` operand.
__device__ inline double shfl_prev(double v, int) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffull), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Orders a wave's LDS writes before its own lanes' later reads.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ----------------------------------------------------- thermal Jacobian
// Live slots of the row of corner v (the hex adds something there).
__host__ __device__ constexpr unsigned jac_live(int v) {
    unsigned m = 0;
    for (int s = 0; s < 15; ++s)
        if (jac_entry(v, s) >= 0) m |= 1u << s;
    return m;
}

// Slots sent to the next wave: corners 2 (layer h) and 6 (layer h + 1) of the
// own hex plus corners 3 and 7 of the previous lane's hex (same node lines).
__host__ __device__ constexpr unsigned jx_live(int lev) { return jac_live(lev << 2 | 2) | jac_live(lev << 2 | 3); }
__host__ __device__ constexpr int popc15(unsigned m) {
    int n = 0;
    for (int s = 0; s < 15; ++s) n += (m >> s) & 1u;
    return n;
}
__host__ __device__ constexpr int jx_index(int lev, int s) {
    // position of (lev, s) among the exported slots, layer h first
    int n = lev == 0 ? 0 : popc15(jx_live(0));
    const unsigned m = jx_live(lev);
    for (int t = 0; t < s; ++t) n += (m >> t) & 1u;
    return n;
}
constexpr int kJX = popc15(jx_live(0)) + popc15(jx_live(1));  // 16
// corners 3, 7 of the last lane's hex go to the next wave's top row
__host__ __device__ constexpr int jt_index(int lev, int s) {
    int n = lev == 0 ? 0 : popc15(jac_live(3));
    const unsigned m = jac_live(lev << 2 | 3);
    for (int t = 0; t < s; ++t) n += (m >> t) & 1u;
    return n;
}
constexpr int kJT = popc15(jac_live(3)) + popc15(jac_live(7));  // 13
// carried slots: corners at dx = 1
__host__ __device__ constexpr unsigned carry_live() {
    return jac_live(4) | jac_live(5) | jac_live(6) | jac_live(7);
}
constexpr int kStage = 65 * 15;  // staged entries of one wave's rows (64 rows + the top row)
static_assert(kJX * 64 >= kStage + 1, "a wave's staging (and its spare entry) fits its exchange slot");

// The 27 values of one hex: w-scaled products c_a . c_b summed over its tets.
template <int PK>
__device__ inline void hex_jac1(const Corner X[8], const Material& mt, double H[27]) {
#pragma unroll
    for (int q = 0; q < 27; ++q) H[q] = 0.0;
    HexGeom hg;
    hex_geom(X, hg);
    sfor<6>([&](auto T) {
        constexpr int t = PA_CI(T);
        constexpr int p = tet_p(t), q = tet_q(t);
        constexpr int v[4] = {0, p, q, 7};
        double G[3];  // c_0
        cross(hg.e[p], hg.e[q], G);
        const double det = dot3(hg.e[0], G);
        // products c_a . c_b (a <= b); those with c_3 from c_3 = -(c_0 + c_1 + c_2)
        double P[4][4];
        P[0][0] = dot3(G, G);
        P[0][1] = dot3(G, hg.F[q]);
        P[0][2] = -dot3(G, hg.F[p]);
        P[1][1] = dot3(hg.F[q], hg.F[q]);
        P[1][2] = -dot3(hg.F[q], hg.F[p]);
        P[2][2] = dot3(hg.F[p], hg.F[p]);
        P[0][3] = -((P[0][0] + P[0][1]) + P[0][2]);
        P[1][3] = -((P[0][1] + P[1][1]) + P[1][2]);
        P[2][3] = -((P[0][2] + P[1][2]) + P[2][2]);
        P[3][3] = -((P[0][3] + P[1][3]) + P[2][3]);
        // s(rho) k V grad N_a . grad N_b = s k / (6 |det|) c_a . c_b
        const double coef = (penal_k<PK>(tet_rho(X, p, q), mt.p, mt.emin) * mt.k) * recip(6.0 * fabs(det));
        sfor<4>([&](auto A) {
            constexpr int a = PA_CI(A);
            sfor<4>([&](auto Bq) {
                constexpr int b = PA_CI(Bq);
                if constexpr (b >= a) {
                    constexpr int idx = a == b ? v[a] : 8 + edge_of(v[a], v[b]);
                    H[idx] = fma(coef, P[a][b], H[idx]);
                }
            });
        });
        // one tet at a time: interleaving the six tets would keep all their
        // temporaries live at once
        __builtin_amdgcn_sched_barrier(0);
    });
}

// A row's entries of the neighbours inside the grid, packed in column order
// at stage[off...]; the slots of missing neighbours (zero) go to a spare
// entry at stage[trash] instead, so there are no branches.
template <class VT>
__device__ inline void store_row(VT* stage, int off, int trash, unsigned exist, const double row[15]) {
    int pos = off;
#pragma unroll
    for (int s = 0; s < 15; ++s) {
        const bool in = (exist >> s) & 1u;
        stage[in ? pos : trash] = (VT)row[s];
        pos += in ? 1 : 0;
    }
}

template <class VT, int W, int PK>
__global__ __launch_bounds__(W * 64, 2) void k_grid_jacobian1(GridArgs g, Material mt, const int* __restrict__ rowmap,
                                                              VT* __restrict__ vals) {
    // xch[buf][w][slot][lane]: wave w's exports to wave w + 1; the region of
    // wave w - 1 (wave 0: the unused region of wave W - 1) then stages wave
    // w's rows
    __shared__ double xch[2][W][kJX][64];
    __shared__ double xtop[2][W][kJT];     // last lane's corners 3, 7 for the next wave's top row
    __shared__ double topc[W][15];         // carried top row of each wave
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const Block B = grid_block(g, W);
    const int j = B.hj0 + w, k = B.hk0 + l;
    const bool line_owned = j >= B.J0 && j < B.J1;
    const bool row_owned = line_owned && k >= B.K0 && k < B.K1;
    const bool top_lane = B.top && l == 63;  // also owns row d2 - 1
    const bool hexcol = j < g.d1 - 1 && k < g.d2 - 1;
    const int hs = B.X0 == 0 ? 0 : B.X0 - 1;
    const int he = B.X1 == g.d0 ? g.d0 : B.X1;  // layer d0 - 1 has no hexes: its rows are the carry
    constexpr unsigned CL = carry_live();
    double carry[15];
#pragma unroll
    for (int s = 0; s < 15; ++s) carry[s] = 0.0;
    if (top_lane) {
#pragma unroll
        for (int s = 0; s < 15; ++s) topc[w][s] = 0.0;
    }
    // corners of node layers h (lo) and h + 1 (hi); the next layer's are
    // loaded one step ahead, with this step's row offsets, so their latency
    // hides behind the hex arithmetic
    Corner lo[4], hi[4];
    if (hexcol) {
        load_layer(g, hs, j, k, lo);
        if (hs < g.d0 - 1) load_layer(g, hs + 1, j, k, hi);
    }
    for (int h = hs; h < he; ++h) {
        const int buf = h & 1;
        const bool hex = hexcol && h < g.d0 - 1;
        Corner nx[4];
        if (hexcol && h + 1 < g.d0 - 1 && h + 1 < he) load_layer(g, h + 2, j, k, nx);
        const bool out = line_owned && h >= B.X0;
        const long line = ((long)h * g.d1 + j) * g.d2;
        int rbase = 0, rend = 0, rrow = 0, rtop = 0;
        if (out) {
            rbase = rowmap[line + B.K0];
            rend = rowmap[line + B.K1];
            if (row_owned) rrow = rowmap[line + k];
            if (top_lane) rtop = rowmap[line + g.d2 - 1];
        }
        double H[27];
        if (hex) {
            Corner X[8];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                X[q] = lo[q];
                X[4 + q] = hi[q];
            }
            hex_jac1<PK>(X, mt, H);
        } else {
#pragma unroll
            for (int q = 0; q < 27; ++q) H[q] = 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            lo[q] = hi[q];
            hi[q] = nx[q];
        }
        // corners 1, 5 of the previous lane's hex, and its corners 3, 7
        double rk[2][15], rjk[2][15];
        sfor<2>([&](auto LV) {
            constexpr int lev = PA_CI(LV);
            sfor<15>([&](auto S) {
                constexpr int s = PA_CI(S);
                constexpr int e1 = jac_entry(lev << 2 | 1, s), e3 = jac_entry(lev << 2 | 3, s);
                if constexpr (e1 >= 0) rk[lev][s] = shfl_prev(H[e1], l);
                else rk[lev][s] = 0.0;
                if constexpr (e3 >= 0) rjk[lev][s] = shfl_prev(H[e3], l);
                else rjk[lev][s] = 0.0;
            });
        });
        if (w < W - 1) {
            sfor<2>([&](auto LV) {
                constexpr int lev = PA_CI(LV);
                sfor<15>([&](auto S) {
                    constexpr int s = PA_CI(S);
                    if constexpr ((jx_live(lev) >> s) & 1u) {
                        constexpr int e2 = jac_entry(lev << 2 | 2, s);
                        double v = rjk[lev][s];
                        if constexpr (e2 >= 0) v += H[e2];
                        xch[buf][w][jx_index(lev, s)][l] = v;
                    }
                });
            });
            if (top_lane) {
                sfor<2>([&](auto LV) {
                    constexpr int lev = PA_CI(LV);
                    sfor<15>([&](auto S) {
                        constexpr int s = PA_CI(S);
                        constexpr int e3 = jac_entry(lev << 2 | 3, s);
                        if constexpr (e3 >= 0) xtop[buf][w][jt_index(lev, s)] = H[e3];
                    });
                });
            }
        }
        __syncthreads();
        double row[15];
        sfor<2>([&](auto LV) {
            constexpr int lev = PA_CI(LV);
            sfor<15>([&](auto S) {
                constexpr int s = PA_CI(S);
                constexpr int e0 = jac_entry(lev << 2, s);
                double v = rk[lev][s];
                if constexpr (e0 >= 0) v = H[e0] + v;
                if constexpr ((jx_live(lev) >> s) & 1u) v += w > 0 ? xch[buf][w - 1][jx_index(lev, s)][l] : 0.0;
                if constexpr (lev == 0) row[s] = carry[s] + v;
                else carry[s] = ((CL >> s) & 1u) ? v : 0.0;
            });
        });
        double trow[15];
        if (top_lane) {
            sfor<2>([&](auto LV) {
                constexpr int lev = PA_CI(LV);
                sfor<15>([&](auto S) {
                    constexpr int s = PA_CI(S);
                    constexpr int e1 = jac_entry(lev << 2 | 1, s), e3 = jac_entry(lev << 2 | 3, s);
                    double v = 0.0;
                    if constexpr (e1 >= 0) v = H[e1];
                    if constexpr (e3 >= 0) v += w > 0 ? xtop[buf][w - 1][jt_index(lev, s)] : 0.0;
                    if constexpr (lev == 0) trow[s] = topc[w][s] + v;
                    else topc[w][s] = v;
                });
            });
        }
        if (out) {
            // the wave's rows [K0, K1) of line (h, j) are one CRS range
            double* region = &xch[buf][w == 0 ? W - 1 : w - 1][0][0];
            VT* stage = reinterpret_cast<VT*>(region);
            wave_lds_sync();  // the exchange reads above come first
            if (row_owned) store_row(stage, rrow - rbase, kStage, row_exists(g, h, j, k), row);
            if (top_lane) store_row(stage, rtop - rbase, kStage, row_exists(g, h, j, g.d2 - 1), trow);
            wave_lds_sync();
            VT* dst = vals + rbase;
            const int n = rend - rbase;
            // four staged entries per lane read before any is stored
            for (int t = l; t < n; t += 256) {
                VT a[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) a[q] = t + 64 * q < n ? stage[t + 64 * q] : VT(0);
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (t + 64 * q < n) store_nt(dst + t + 64 * q, a[q]);
            }
        }
    }
}

// -------------------------------------------------------------- residual
// Per corner NV = DPN values; the hex's eight corner sums.
template <int DPN, int PK>
__device__ inline void hex_residual(const Corner X[8], const double U[8][DPN], const Material& mt, double body,
                                   double H[8][DPN]) {
#pragma unroll
    for (int v = 0; v < 8; ++v)
#pragma unroll
        for (int d = 0; d < DPN; ++d) H[v][d] = 0.0;
    HexGeom hg;
    hex_geom<DPN == 1>(X, hg);
    // state differences to corner 7: sum_b u_b (x) c_b = sum_{b < 3} (u_b - u_3) (x) c_b
    double du[7][DPN];
#pragma unroll
    for (int p = 0; p < 7; ++p)
#pragma unroll
        for (int d = 0; d < DPN; ++d) du[p][d] = U[p][d] - U[7][d];
    sfor<6>([&](auto T) {
        constexpr int t = PA_CI(T);
        constexpr int p = tet_p(t), q = tet_q(t);
        constexpr int v[4] = {0, p, q, 7};
        double c[3][3];  // c_0, c_1, c_2
        cross(hg.e[p], hg.e[q], c[0]);
        if constexpr (DPN == 3) {
            // the elastic kernel has no registers to keep the shared F[p]
            cross(hg.e[q], hg.e[0], c[1]);
            cross(hg.e[0], hg.e[p], c[2]);
        } else {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                c[1][d] = hg.F[q][d];
                c[2][d] = -hg.F[p][d];
            }
        }
        const double det = dot3(hg.e[0], c[0]);
        const double* dl[3] = {du[0], du[p], du[q]};
        const double rho = tet_rho(X, p, q);
        const double s = penal_k<PK>(rho, mt.p, mt.emin);
        double f[3][DPN];  // nodal residuals of local nodes 0..2; node 3's is minus their sum
        if constexpr (DPN == 3) {
            // eps(u; c) in Voigt order xx, yy, zz, yz, xz, xy (engineering shear)
            double e[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double* u = dl[a];
                const double* ca = c[a];
                e[0] += u[0] * ca[0];
                e[1] += u[1] * ca[1];
                e[2] += u[2] * ca[2];
                e[3] += u[2] * ca[1] + u[1] * ca[2];
                e[4] += u[2] * ca[0] + u[0] * ca[2];
                e[5] += u[1] * ca[0] + u[0] * ca[1];
            }
            const double coef = s * recip(6.0 * fabs(det));
            const double tr = (e[0] + e[1]) + e[2];
            const double cl = coef * mt.lambda, c2m = coef * 2.0 * mt.mu, cm = coef * mt.mu;
            const double sg[6] = {cl * tr + c2m * e[0], cl * tr + c2m * e[1], cl * tr + c2m * e[2],
                                  cm * e[3], cm * e[4], cm * e[5]};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double* ca = c[a];
                f[a][0] = (sg[0] * ca[0] + sg[5] * ca[1]) + sg[4] * ca[2];
                f[a][1] = (sg[5] * ca[0] + sg[1] * ca[1]) + sg[3] * ca[2];
                f[a][2] = (sg[4] * ca[0] + sg[3] * ca[1]) + sg[2] * ca[2];
            }
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int d = 0; d < 3; ++d) H[v[a]][d] += f[a][d];
#pragma unroll
            for (int d = 0; d < 3; ++d) H[7][d] -= (f[0][d] + f[1][d]) + f[2][d];
        } else {
            double gt[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) gt[d] = (dl[0][0] * c[0][d] + dl[1][0] * c[1][d]) + dl[2][0] * c[2][d];
            const double coef = (s * mt.k) * recip(6.0 * fabs(det));
            // body source: -(cubature weight 1/6)|det J| f (1/4) rho (unpenalised mean density)
            const double load = -1.0 * ((1.0 / 6.0) * fabs(det)) * body * 0.25 * rho;
#pragma unroll
            for (int a = 0; a < 3; ++a) f[a][0] = coef * dot3(gt, c[a]);
#pragma unroll
            for (int a = 0; a < 3; ++a) H[v[a]][0] += f[a][0] + load;
            H[7][0] += load - ((f[0][0] + f[1][0]) + f[2][0]);
        }
        __builtin_amdgcn_sched_barrier(0);
    });
}

template <int DPN>
__device__ inline void load_state(const double* __restrict__ u, int i, int j, int k, const GridArgs& g,
                                  double U[4][DPN]) {
    const long n = ((long)i * g.d1 + j) * g.d2 + k;
    const long nn[4] = {n, n + 1, n + g.d2, n + g.d2 + 1};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int d = 0; d < DPN; ++d) U[q][d] = u[nn[q] * DPN + d];
}

template <int DPN, int W, int PK>
__global__ __launch_bounds__(W * 64, 2) void k_grid_residual(GridArgs g, Material mt, const double* __restrict__ u,
                                                             double body, double scale,
                                                             const double* __restrict__ fext,
                                                             double* __restrict__ R) {
    constexpr int NX = 2 * DPN;
    __shared__ double xch[2][W][NX][64];
    __shared__ double xtop[2][W][NX];
    __shared__ double topc[W][DPN];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const Block B = grid_block(g, W);
    const int j = B.hj0 + w, k = B.hk0 + l;
    const bool line_owned = j >= B.J0 && j < B.J1;
    const bool row_owned = line_owned && k >= B.K0 && k < B.K1;
    const bool top_lane = B.top && l == 63;
    const bool hexcol = j < g.d1 - 1 && k < g.d2 - 1;
    const int hs = B.X0 == 0 ? 0 : B.X0 - 1;
    const int he = B.X1 == g.d0 ? g.d0 : B.X1;
    double carry[DPN];
#pragma unroll
    for (int d = 0; d < DPN; ++d) carry[d] = 0.0;
    if (top_lane) {
#pragma unroll
        for (int d = 0; d < DPN; ++d) topc[w][d] = 0.0;
    }
    // corners and states of node layers h (lo) and h + 1 (hi).  The thermal
    // kernel loads the next layer's one step ahead (its latency hides behind
    // the hex arithmetic); the elastic one has no registers to spare for it
    // and loads layer h + 1 in the step.
    constexpr bool AHEAD = DPN == 1;
    Corner lo[4], hi[4];
    double ulo[4][DPN], uhi[4][DPN];
    if (hexcol) {
        load_layer(g, hs, j, k, lo);
        load_state<DPN>(u, hs, j, k, g, ulo);
        if (AHEAD && hs < g.d0 - 1) {
            load_layer(g, hs + 1, j, k, hi);
            load_state<DPN>(u, hs + 1, j, k, g, uhi);
        }
    }
    for (int h = hs; h < he; ++h) {
        const int buf = h & 1;
        const bool hex = hexcol && h < g.d0 - 1;
        Corner nx[4];
        double unx[4][DPN];
        if (AHEAD && hexcol && h + 1 < g.d0 - 1 && h + 1 < he) {
            load_layer(g, h + 2, j, k, nx);
            load_state<DPN>(u, h + 2, j, k, g, unx);
        }
        if (!AHEAD && hex) {
            load_layer(g, h + 1, j, k, hi);
            load_state<DPN>(u, h + 1, j, k, g, uhi);
        }
        double H[8][DPN];
        if (hex) {
            Corner X[8];
            double U[8][DPN];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                X[q] = lo[q];
                X[4 + q] = hi[q];
#pragma unroll
                for (int d = 0; d < DPN; ++d) {
                    U[q][d] = ulo[q][d];
                    U[4 + q][d] = uhi[q][d];
                }
            }
            hex_residual<DPN, PK>(X, U, mt, body, H);
        } else {
#pragma unroll
            for (int v = 0; v < 8; ++v)
#pragma unroll
                for (int d = 0; d < DPN; ++d) H[v][d] = 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            lo[q] = hi[q];
#pragma unroll
            for (int d = 0; d < DPN; ++d) ulo[q][d] = uhi[q][d];
            if constexpr (AHEAD) {
                hi[q] = nx[q];
#pragma unroll
                for (int d = 0; d < DPN; ++d) uhi[q][d] = unx[q][d];
            }
        }
        double rk[2][DPN], rjk[2][DPN];
#pragma unroll
        for (int lev = 0; lev < 2; ++lev)
#pragma unroll
            for (int d = 0; d < DPN; ++d) {
                rk[lev][d] = shfl_prev(H[lev << 2 | 1][d], l);
                rjk[lev][d] = shfl_prev(H[lev << 2 | 3][d], l);
            }
        if (w < W - 1) {
#pragma unroll
            for (int lev = 0; lev < 2; ++lev)
#pragma unroll
                for (int d = 0; d < DPN; ++d) xch[buf][w][lev * DPN + d][l] = H[lev << 2 | 2][d] + rjk[lev][d];
            if (top_lane) {
#pragma unroll
                for (int lev = 0; lev < 2; ++lev)
#pragma unroll
                    for (int d = 0; d < DPN; ++d) xtop[buf][w][lev * DPN + d] = H[lev << 2 | 3][d];
            }
        }
        __syncthreads();
        double row[DPN], trow[DPN];
#pragma unroll
        for (int d = 0; d < DPN; ++d) {
            const double j0 = w > 0 ? xch[buf][w - 1][d][l] : 0.0;
            const double j1 = w > 0 ? xch[buf][w - 1][DPN + d][l] : 0.0;
            row[d] = carry[d] + ((H[0][d] + rk[0][d]) + j0);
            carry[d] = (H[4][d] + rk[1][d]) + j1;
        }
        if (top_lane) {
#pragma unroll
            for (int d = 0; d < DPN; ++d) {
                const double t0 = w > 0 ? xtop[buf][w - 1][d] : 0.0;
                const double t1 = w > 0 ? xtop[buf][w - 1][DPN + d] : 0.0;
                trow[d] = topc[w][d] + (H[1][d] + t0);
                topc[w][d] = H[5][d] + t1;
            }
        }
        if (line_owned && h >= B.X0) {
            const long line = ((long)h * g.d1 + j) * g.d2;
            if (row_owned) {
                const long o = (line + k) * DPN;
#pragma unroll
                for (int d = 0; d < DPN; ++d) R[o + d] = scale * row[d] + (fext ? fext[o + d] : 0.0);
            }
            if (top_lane) {
                const long o = (line + g.d2 - 1) * DPN;
#pragma unroll
                for (int d = 0; d < DPN; ++d) R[o + d] = scale * trow[d] + (fext ? fext[o + d] : 0.0);
            }
        }
    }
}

// ----------------------------------------------------- elastic Jacobian
// 3 x 3 blocks of SIMP elastostatics on Kuhn grids.  A workgroup owns a brick
// of 4 x 4 x 4 node rows (1024 threads, 16 per row).  It stages the brick's
// node halo (coordinates and densities) in LDS, evaluates each tet of the
// 5 x 5 x 5 hexes around the rows once, one per thread, and parks the four
// cofactor vectors of the tet scaled by sqrt(s), s = s(rho) / (6 |det J|), so
// that the stiffness block of nodes a, b in the tet is
// lambda A + mu A^T + mu tr(A) I with A = u_a (x) u_b.  Then lane `slot` of a
// row takes the neighbour of CRS slot `slot` and sums A over the tets the two
// nodes share (at most six, in the order of their hexes and tets, which is the
// same for the block (m, n) as for (n, m), so the two are transposes bit for
// bit).  The diagonal block is minus the symmetric part of the sum of the
// row's other blocks (a rigid translation is in the null space of every tet's
// stiffness), summed over the row's 16 lanes by DPP.  Blocks are staged in LDS
// in CRS order (the tet data is dead by then) and each line of rows, one
// contiguous CRS range, is stored by one wave with contiguous stores.  No
// atomics: deterministic.  Measured on the 128 x 64 x 64 grid: bricks of
// 2 x 3 x 4 rows (4 workgroups per CU) 0.24 ms, 2 x 4 x 4 0.21 ms, 2 x 6 x 4
// 0.19 ms, 4 x 4 x 4 (2 workgroups, 8 waves per SIMD) 0.19 ms; 8 lanes per
// row (two blocks each) 0.21 ms.  With the per-lane tables below (no index
// arithmetic) 0.179 ms FP64, 0.139 ms FP32.  The FP64 kernel is bound by its
// load-stage-store structure, not by arithmetic: with the tet and entry work
// skipped it still takes 0.175 ms (0.144 ms without the node-halo loads),
// while a plain fill of the 577 MB writes in 0.092 ms.  Measured slower:
// persistent workgroups that prefetch the next brick's halo into registers
// (0.221 ms, 64 VGPRs), an L2 prefetch of a later brick's halo (0.193 ms),
// FP64 blocks stored directly by each lane (0.255 ms), the diagonal sums
// exchanged through ds_bpermute instead of DPP (0.226 ms), bricks of
// 3 x 5 x 4 rows (960 threads; 6.6% fewer threads than 4 x 4 x 4 on the
// 129 x 65 x 65 node grid, whose 4k + 1 sides leave one-row bricks) 0.216 ms.
constexpr int kJ3Slots = 16;  // lanes per row: the 15 CRS slots and one idle
constexpr int kJ3MaxEnt = 6;  // tets shared by a node and a neighbour

struct J3Entry {
    signed char hx, hy, hz;  // hex origin - row node, each -1 or 0
    signed char t, a, b;     // tet; local indices of the row node and the neighbour
};
struct J3Table {
    J3Entry e[kJ3Slots][kJ3MaxEnt];
    int n[kJ3Slots];
};

// Local index of hex corner v in tet (0, p, q, 7), or -1.
__host__ __device__ constexpr int tet_local(int t, int v) {
    return v == 0 ? 0 : v == tet_p(t) ? 1 : v == tet_q(t) ? 2 : v == 7 ? 3 : -1;
}

__host__ __device__ constexpr J3Table make_j3_table() {
    J3Table T{};
    for (int s = 0; s < kJ3Slots; ++s) {
        T.n[s] = 0;
        if (s >= 15 || s == kDiagSlot) continue;
        for (int hx = -1; hx <= 0; ++hx)
            for (int hy = -1; hy <= 0; ++hy)
                for (int hz = -1; hz <= 0; ++hz) {
                    const int cx = -hx, cy = -hy, cz = -hz;  // the row node's corner
                    const int mx = cx + nb_di(s), my = cy + nb_dj(s), mz = cz + nb_dk(s);
                    if (mx < 0 || mx > 1 || my < 0 || my > 1 || mz < 0 || mz > 1) continue;
                    const int vn = cx << 2 | cy << 1 | cz, vm = mx << 2 | my << 1 | mz;
                    for (int t = 0; t < 6; ++t) {
                        const int a = tet_local(t, vn), b = tet_local(t, vm);
                        if (a < 0 || b < 0) continue;
                        T.e[s][T.n[s]] = J3Entry{(signed char)hx, (signed char)hy, (signed char)hz, (signed char)t,
                                                 (signed char)a, (signed char)b};
                        ++T.n[s];
                    }
                }
    }
    return T;
}

__host__ __device__ constexpr int j3_count(int s) { return make_j3_table().n[s]; }
// axis edges and the body diagonal: 6 tets; face diagonals: 4
static_assert(j3_count(0) == 6 && j3_count(1) == 4 && j3_count(2) == 4 && j3_count(3) == 6, "Kuhn edge tets");
static_assert(j3_count(4) == 4 && j3_count(5) == 6 && j3_count(6) == 6 && j3_count(7) == 0, "Kuhn edge tets");
static_assert(j3_count(8) == 6 && j3_count(14) == 6 && j3_count(15) == 0, "Kuhn edge tets");

// v from the lane given by a DPP row control (both halves of the double).
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)(b & 0xffffffffull), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over the 16 lanes of a DPP row, every lane getting it, in a fixed
// order: lane pairs (xor 1), quads (xor 2), half rows (mirror), rows (mirror).
__device__ inline double row16_sum(double v) {
    v += dpp_f64<0xb1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_f64<0x4e>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_f64<0x141>(v);  // row_half_mirror
    v += dpp_f64<0x140>(v);  // row_mirror
    return v;
}

template <int BX, int BY, int BZ>
struct J3Brick {
    static constexpr int HX = BX + 1, HY = BY + 1, HZ = BZ + 1;  // hexes
    static constexpr int NX = BX + 2, NY = BY + 2, NZ = BZ + 2;  // nodes
    static constexpr int rows = BX * BY * BZ, lines = BX * BY, threads = rows * kJ3Slots, waves = threads / 64;
    static constexpr int tets = HX * HY * HZ * 6, nodes = NX * NY * NZ;
    static constexpr int hex_stride = 6 * 12, tet_doubles = HX * HY * HZ * hex_stride;  // LDS doubles
    static constexpr int line_blocks = BZ * 15;  // staged blocks per line
};

// Per-lane tables of a brick, resolved at compile time so that the kernel
// spends no instructions on index arithmetic:
// - entries: for thread tid (row tid / 16, slot tid % 16) the LDS offsets
//   (doubles) of u_a and u_b of each of its six entries, packed as two
//   uint16; unused entries read the zero vector after the tet data;
// - tets: for thread tid < tets, the tet it evaluates (tet tid % 6 of hex
//   tid / 6): the LDS node index of the hex's corner 0 (bits 0-7), the node
//   offsets of the tet's corners p (8-13) and q (14-19), and the hex's
//   position in the brick (3 bits each from bit 20).
template <int BX, int BY, int BZ>
struct J3LaneTables {
    uint2 ent[J3Brick<BX, BY, BZ>::threads][3];
    unsigned tet[J3Brick<BX, BY, BZ>::threads];
};
template <int BX, int BY, int BZ>
__host__ __device__ constexpr J3LaneTables<BX, BY, BZ> make_j3_lanes() {
    using B = J3Brick<BX, BY, BZ>;
    constexpr J3Table T = make_j3_table();
    constexpr int zero = B::tet_doubles;
    static_assert(zero + 3 < 65536, "16-bit LDS offsets");
    J3LaneTables<BX, BY, BZ> L{};
    for (int tid = 0; tid < B::threads; ++tid) {
        const int slot = tid % kJ3Slots, r = tid / kJ3Slots;
        const int rc = r % BZ, rb = (r / BZ) % BY, ra = r / (BZ * BY);
        const int rowbase = (((ra + 1) * B::HY + (rb + 1)) * B::HZ + (rc + 1)) * B::hex_stride;
        unsigned w[6] = {};
        for (int e = 0; e < kJ3MaxEnt; ++e) {
            int oa = zero, ob = zero;
            if (e < T.n[slot]) {
                const J3Entry E = T.e[slot][e];
                const int h = (E.hx * B::HY + E.hy) * B::HZ + E.hz;
                oa = rowbase + h * B::hex_stride + E.t * 12 + 3 * E.a;
                ob = rowbase + h * B::hex_stride + E.t * 12 + 3 * E.b;
            }
            w[e] = (unsigned)oa | (unsigned)ob << 16;
        }
        for (int k = 0; k < 3; ++k) L.ent[tid][k] = uint2{w[2 * k], w[2 * k + 1]};
        L.tet[tid] = 0;
        if (tid < B::tets) {
            const int h = tid / 6, t = tid % 6;
            const int hc = h % B::HZ, hb = (h / B::HZ) % B::HY, ha = h / (B::HZ * B::HY);
            const int n0 = (ha * B::NY + hb) * B::NZ + hc;
            auto at = [](int v) { return ((v >> 2) * B::NY + ((v >> 1) & 1)) * B::NZ + (v & 1); };
            L.tet[tid] = (unsigned)n0 | (unsigned)at(tet_p(t)) << 8 | (unsigned)at(tet_q(t)) << 14 |
                         (unsigned)ha << 20 | (unsigned)hb << 23 | (unsigned)hc << 26;
        }
    }
    return L;
}
template <int BX, int BY, int BZ>
__constant__ J3LaneTables<BX, BY, BZ> kJ3Lanes = make_j3_lanes<BX, BY, BZ>();

// sqrt(s) for a positive normal s: the hardware reciprocal square root
// estimate refined by a Goldschmidt step and two Newton corrections of the
// root (within an ulp, without the range scaling a general sqrt needs).
__device__ inline double sqrt_pos(double s) {
    const double y = __builtin_amdgcn_rsq(s);
    double g = s * y, h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, s);
    g = fma(d, h, g);
    d = fma(-g, g, s);
    return fma(d, h, g);
}

// Residual with the Jacobian (WR): linear elastostatics has R(u) = K u + f,
// so the lanes of a row multiply their blocks by the neighbours'
// displacements (the diagonal lane adds the natural-BC load f) and, after
// the stores are issued, sum the products over the row's 16 lanes by DPP
// (fixed order); lane 7 writes the row of R.  This costs a few instructions
// per lane instead of a second pass over the tets.
// (launch bound: 8 waves per SIMD, two 16-wave workgroups per CU)
template <class VT, int BX, int BY, int BZ, int PK, bool WR>
__global__ __launch_bounds__(BX * BY * BZ * kJ3Slots) void k_grid_jacobian3(GridArgs g, Material mt,
                                                                             const int* __restrict__ rowmap,
                                                                             VT* __restrict__ vals,
                                                                             const double* __restrict__ u,
                                                                             const double* __restrict__ fext,
                                                                             double* __restrict__ R) {
    using B = J3Brick<BX, BY, BZ>;
    static_assert(B::nodes + B::rows + B::lines + 4 <= B::threads, "one load per thread");
    static_assert(B::lines * B::line_blocks * 9 * sizeof(VT) <= B::tet_doubles * sizeof(double), "stage fits");
    constexpr int kZero = B::tet_doubles;      // a zero vector after the tets
    __shared__ double tet[B::tet_doubles + 4];  // sqrt(s) c_a, a = 0..3, by hex; later the staged blocks
    __shared__ double4 node[B::nodes];        // x, y, z, density
    __shared__ int rstart[B::rows], lstart[B::lines], llen[B::lines];
    const int tid = threadIdx.x;
    const int nbz = (g.d2 + BZ - 1) / BZ, nby = (g.d1 + BY - 1) / BY;
    const int bk = xcd_block();
    const int X0 = (bk / (nbz * nby)) * BX, Y0 = ((bk / nbz) % nby) * BY, Z0 = (bk % nbz) * BZ;
    // this lane's row and its entries' LDS offsets of u_a (low 16 bits) and
    // u_b (high 16 bits)
    const int slot = tid & (kJ3Slots - 1), r = tid / kJ3Slots;
    const int rc = r % BZ, rb = (r / BZ) % BY, ra = r / (BZ * BY);
    const uint2 w0 = kJ3Lanes<BX, BY, BZ>.ent[tid][0], w1 = kJ3Lanes<BX, BY, BZ>.ent[tid][1],
                w2 = kJ3Lanes<BX, BY, BZ>.ent[tid][2];
    const unsigned off[kJ3MaxEnt] = {w0.x, w0.y, w1.x, w1.y, w2.x, w2.y};
    const unsigned tw = kJ3Lanes<BX, BY, BZ>.tet[tid];
    const int ri = X0 + ra, rj = Y0 + rb, rk = Z0 + rc;
    // stage 0: node halo, row offsets, line ranges, the zero vector
    if (tid < B::nodes) {
        const int nc = tid % B::NZ, nb = (tid / B::NZ) % B::NY, na = tid / (B::NZ * B::NY);
        const int i = X0 - 1 + na, j = Y0 - 1 + nb, k = Z0 - 1 + nc;
        double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
        if (i >= 0 && i < g.d0 && j >= 0 && j < g.d1 && k >= 0 && k < g.d2) {
            const long n = ((long)i * g.d1 + j) * g.d2 + k;
            p = g.x4[n];
            p.w = g.z[n];
        }
        node[tid] = p;
    } else if (tid < B::nodes + B::rows) {
        const int q = tid - B::nodes;
        const int qc = q % BZ, qb = (q / BZ) % BY, qa = q / (BZ * BY);
        const int i = X0 + qa, j = Y0 + qb, k = Z0 + qc;
        rstart[q] = (i < g.d0 && j < g.d1 && k < g.d2) ? rowmap[((long)i * g.d1 + j) * g.d2 + k] : 0;
    } else if (tid < B::nodes + B::rows + B::lines) {
        const int q = tid - B::nodes - B::rows;
        const int i = X0 + q / BY, j = Y0 + q % BY;
        int s0 = 0, len = 0;
        if (i < g.d0 && j < g.d1) {
            const long n0 = ((long)i * g.d1 + j) * g.d2 + Z0;
            const int kl = min(Z0 + BZ, g.d2) - 1;
            s0 = rowmap[n0];
            len = rowmap[n0 + (kl - Z0) + 1] - s0;
        }
        lstart[q] = s0;
        llen[q] = len;
    } else if (tid < B::nodes + B::rows + B::lines + 4) {
        tet[kZero + tid - (B::nodes + B::rows + B::lines)] = 0.0;
    }
    __syncthreads();
    // stage 1: the tets, one per thread (tet tid % 6 of hex tid / 6)
    static_assert(B::tets <= B::threads, "one tet per thread");
    if (tid < B::tets) {
        const int ha = (tw >> 20) & 7, hb = (tw >> 23) & 7, hc = (tw >> 26) & 7;
        const int i = X0 - 1 + ha, j = Y0 - 1 + hb, k = Z0 - 1 + hc;
        // branch-free: a hex outside the grid (its corners read as zeros or
        // as nodes of the grid) gets sqrt(s) = 0
        const bool in = i >= 0 && i < g.d0 - 1 && j >= 0 && j < g.d1 - 1 && k >= 0 && k < g.d2 - 1;
        const int n0 = tw & 255;
        const double4 x0 = node[n0], xp = node[n0 + ((tw >> 8) & 63)], xq = node[n0 + ((tw >> 14) & 63)];
        const double4 x7 = node[n0 + (B::NY + 1) * B::NZ + 1];
        const double e0[3] = {x0.x - x7.x, x0.y - x7.y, x0.z - x7.z};
        const double e1[3] = {xp.x - x7.x, xp.y - x7.y, xp.z - x7.z};
        const double e2[3] = {xq.x - x7.x, xq.y - x7.y, xq.z - x7.z};
        double u[4][3];
        cross(e1, e2, u[0]);
        cross(e2, e0, u[1]);
        cross(e0, e1, u[2]);
        const double det = dot3(e0, u[0]);
        const double rho = (((x0.w + xp.w) + xq.w) + x7.w) * 0.25;
        const double s = penal_k<PK>(rho, mt.p, mt.emin) * recip(6.0 * fabs(det));
        const double sw = in ? sqrt_pos(s) : 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            u[3][d] = -((u[0][d] + u[1][d]) + u[2][d]) * sw;
            u[0][d] *= sw;
            u[1][d] *= sw;
            u[2][d] *= sw;
        }
        double2* dst = reinterpret_cast<double2*>(tet + tid * 12);
#pragma unroll
        for (int a = 0; a < 4; a += 2) {
            dst[3 * (a / 2) + 0] = make_double2(u[a][0], u[a][1]);
            dst[3 * (a / 2) + 1] = make_double2(u[a][2], u[a + 1][0]);
            dst[3 * (a / 2) + 2] = make_double2(u[a + 1][1], u[a + 1][2]);
        }
    }
    __syncthreads();
    // WR: the displacement of the lane's neighbour and, on the diagonal lane,
    // the row's load (in flight during stage 2)
    double un[3] = {0.0, 0.0, 0.0}, fr[3] = {0.0, 0.0, 0.0};
    if constexpr (WR) {
        const bool in = ri < g.d0 && rj < g.d1 && rk < g.d2;
        if (in && slot < 15 && ((row_exists(g, ri, rj, rk) >> slot) & 1u)) {
            const long n = ((long)(ri + nb_di(slot)) * g.d1 + (rj + nb_dj(slot))) * g.d2 + (rk + nb_dk(slot));
            un[0] = u[3 * n];
            un[1] = u[3 * n + 1];
            un[2] = u[3 * n + 2];
            if (slot == kDiagSlot && fext) {
                fr[0] = fext[3 * n];
                fr[1] = fext[3 * n + 1];
                fr[2] = fext[3 * n + 2];
            }
        }
    }
    // stage 2: the lane's block from its (padded) entries
    double A[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) A[q] = 0.0;
#pragma unroll
    for (int e = 0; e < kJ3MaxEnt; ++e) {
        const double* pa = tet + (off[e] & 0xffffu);
        const double* pb = tet + (off[e] >> 16);
        const double a0 = pa[0], a1 = pa[1], a2 = pa[2];
        const double b0 = pb[0], b1 = pb[1], b2 = pb[2];
        A[0] = fma(a0, b0, A[0]); A[1] = fma(a0, b1, A[1]); A[2] = fma(a0, b2, A[2]);
        A[3] = fma(a1, b0, A[3]); A[4] = fma(a1, b1, A[4]); A[5] = fma(a1, b2, A[5]);
        A[6] = fma(a2, b0, A[6]); A[7] = fma(a2, b1, A[7]); A[8] = fma(a2, b2, A[8]);
    }
    double K[9];
    const double tr = (A[0] + A[4]) + A[8];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int i = q / 3, n = q % 3;
        K[q] = mt.lambda * A[i * 3 + n] + mt.mu * A[n * 3 + i];
        if (i == n) K[q] += mt.mu * tr;
    }
    // diagonal block: minus the symmetric part of the sum of the others,
    // summed over the row's 16 lanes by DPP (slots 7 and 15 hold zero)
    const double S00 = row16_sum(K[0]), S11 = row16_sum(K[4]), S22 = row16_sum(K[8]);
    const double S01 = row16_sum(K[1] + K[3]), S02 = row16_sum(K[2] + K[6]), S12 = row16_sum(K[5] + K[7]);
    if (slot == kDiagSlot) {
        K[0] = -S00;
        K[4] = -S11;
        K[8] = -S22;
        K[1] = K[3] = -0.5 * S01;
        K[2] = K[6] = -0.5 * S02;
        K[5] = K[7] = -0.5 * S12;
    }
    // WR: the lane's part of row m of K u + f (the load on the diagonal lane)
    double q[3] = {0.0, 0.0, 0.0};
    if constexpr (WR) {
#pragma unroll
        for (int i = 0; i < 3; ++i) q[i] = fma(K[3 * i + 2], un[2], fma(K[3 * i + 1], un[1], fma(K[3 * i], un[0], fr[i])));
    }
    if constexpr (sizeof(VT) == 4) {
        // FP32 values: each lane stores its 36-byte block where it belongs (a
        // row's blocks are contiguous in CRS, a wave's rows consecutive);
        // measured faster than staging, unlike the 72-byte FP64 blocks
        if (ri < g.d0 && rj < g.d1 && rk < g.d2 && slot < 15) {
            const unsigned m = row_exists(g, ri, rj, rk);
            if ((m >> slot) & 1u) {
                VT* dst = vals + ((long)rstart[r] + __popc(m & ((1u << slot) - 1u))) * 9;
#pragma unroll
                for (int q = 0; q < 9; ++q) store_nt(dst + q, (VT)K[q]);
            }
        }
    } else {
        __syncthreads();
        // stage 3: blocks into the stage, line by line in CRS order
        VT* stage = reinterpret_cast<VT*>(tet);
        if (ri < g.d0 && rj < g.d1 && rk < g.d2 && slot < 15) {
            const unsigned m = row_exists(g, ri, rj, rk);
            if ((m >> slot) & 1u) {
                const int line = ra * BY + rb;
                const int pos = __popc(m & ((1u << slot) - 1u));
                VT* dst = stage + (line * B::line_blocks + (rstart[r] - lstart[line]) + pos) * 9;
#pragma unroll
                for (int q = 0; q < 9; ++q) dst[q] = (VT)K[q];
            }
        }
        if constexpr (WR && BZ * kJ3Slots == 64) {
            // a line's rows are one wave's rows: the wave stores the line it
            // staged, so only the wave waits for its LDS writes (measured:
            // value_gradient_u 0.2060 -> 0.2005 ms; gradient_u alone, with no
            // row sums after the stores, 0.1795 -> 0.1900 ms, so it keeps the
            // workgroup barrier)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            __syncthreads();
        }
        // stage 4: wave w stores lines w, w + waves, ...: each one contiguous CRS
        // range, in 16-byte pairs (a single value first when the range starts
        // at an odd double, and last when one is left over)
        static_assert(B::line_blocks * 9 % 2 == 0, "16-byte aligned lines in the stage");
        const int wv = tid >> 6, lane = tid & 63;
        for (int line = wv; line < B::lines; line += B::waves) {
            const int n = llen[line] * 9;
            const VT* src = stage + line * B::line_blocks * 9;
            VT* dstg = vals + (long)lstart[line] * 9;
            const int h = (int)(lstart[line] & 1);  // 9 * lstart is odd iff lstart is
            if (lane == 0 && h && n > 0) store_nt(dstg, src[0]);
            for (int e = h + 2 * lane; e + 1 < n; e += 128) store_nt2(dstg + e, src[e], src[e + 1]);
            if (lane == 0 && n > h && ((n - h) & 1)) store_nt(dstg + n - 1, src[n - 1]);
        }
    }
    if constexpr (WR) {
        // the row of R = K u + f, summed over the row's lanes while the
        // stores drain
#pragma unroll
        for (int i = 0; i < 3; ++i) q[i] = row16_sum(q[i]);
        // lanes 0, 1, 2 of the row write its three components: a wave's
        // four rows are consecutive nodes, so one store of 96 bytes
        if (slot < 3 && ri < g.d0 && rj < g.d1 && rk < g.d2) {
            const long o = 3 * (((long)ri * g.d1 + rj) * g.d2 + rk);
            R[o + slot] = slot == 0 ? q[0] : (slot == 1 ? q[1] : q[2]);
        }
    }
}

// Launch geometry: line blocks, z blocks, x blocks (about 8 workgroups per
// CU in all, each walking at least 4 layers).
GridArgs grid_args(const pa_mesh* m, int W) {
    GridArgs g{};
    g.d0 = m->kgrid[0];
    g.d1 = m->kgrid[1];
    g.d2 = m->kgrid[2];
    g.nj = 1;
    while (W + (W - 1) * (g.nj - 1) < g.d1) ++g.nj;
    g.nk = 1;
    for (;;) {
        const int K0 = g.nk == 1 ? 0 : 64 + 63 * (g.nk - 2);
        int K1 = std::min(K0 + (g.nk == 1 ? 64 : 63), g.d2);
        const int hk0 = g.nk == 1 ? 0 : K0 - 1;
        if (K1 == g.d2 - 1 && hk0 + 64 == g.d2 - 1) K1 = g.d2;
        if (K1 >= g.d2) break;
        ++g.nk;
    }
    const long lines = (long)g.nj * g.nk;
    const long target = 2048;
    int nx = (int)std::max<long>(1, std::min<long>((target + lines - 1) / lines, (g.d0 + 3) / 4));
    g.lx = (g.d0 + nx - 1) / nx;
    g.x4 = reinterpret_cast<const double4*>(m->d_coords4.p);
    return g;
}

int grid_blocks(const GridArgs& g) { return g.nj * g.nk * ((g.d0 + g.lx - 1) / g.lx); }

// Waves per workgroup of the grid kernels: kGridWaves, or PA_GRID_WAVES=8
// (fewer halo lines, one workgroup per CU).
int grid_waves() {
    static const int w = [] {
        const char* e = std::getenv("PA_GRID_WAVES");
        return e && std::atoi(e) == 8 ? 8 : kGridWaves;
    }();
    return w;
}

// Exponent classes with their own code: 3 (the configurations), 1 (the
// reference's unit tests), any other through pow.
template <class Run, class K3, class K1, class K0>
void by_penalty(double p, Run&& run, K3 k3, K1 k1, K0 k0) {
    const int pk = penal_kind(p);
    if (pk == 3) run(k3);
    else if (pk == 1) run(k1);
    else run(k0);
}

template <int W>
void grid_residual_w(const pa_mesh* m, int dpn, const Material& mt, const double* u, const double* z, double body,
                     double scale, const double* fext, double* R, hipStream_t st) {
    GridArgs g = grid_args(m, W);
    g.z = z;
    const int nb = grid_blocks(g);
    auto run = [&](auto kernel) { kernel<<<nb, W * 64, 0, st>>>(g, mt, u, body, scale, fext, R); };
    if (dpn == 3)
        by_penalty(mt.p, run, k_grid_residual<3, W, 3>, k_grid_residual<3, W, 1>, k_grid_residual<3, W, 0>);
    else
        by_penalty(mt.p, run, k_grid_residual<1, W, 3>, k_grid_residual<1, W, 1>, k_grid_residual<1, W, 0>);
}

template <class VT, int W>
void grid_jacobian1_w(const pa_mesh* m, const Material& mt, const double* z, VT* vals, hipStream_t st) {
    GridArgs g = grid_args(m, W);
    g.z = z;
    const int nb = grid_blocks(g);
    auto run = [&](auto kernel) { kernel<<<nb, W * 64, 0, st>>>(g, mt, m->d_rowmap.p, vals); };
    by_penalty(mt.p, run, k_grid_jacobian1<VT, W, 3>, k_grid_jacobian1<VT, W, 1>, k_grid_jacobian1<VT, W, 0>);
}

template <class VT, int BX, int BY, int BZ>
void grid_jacobian3_brick(const pa_mesh* m, const GridArgs& g, const Material& mt, VT* vals, const double* u,
                          const double* fext, double* R, hipStream_t st) {
    const long nb = (long)((g.d0 + BX - 1) / BX) * ((g.d1 + BY - 1) / BY) * ((g.d2 + BZ - 1) / BZ);
    PA_REQUIRE(nb < (1L << 31), PA_ERR_LIMIT, "gradient_u: grid too large for the brick launch");
    constexpr int nt = J3Brick<BX, BY, BZ>::threads;
    auto run = [&](auto kernel) { kernel<<<(int)nb, nt, 0, st>>>(g, mt, m->d_rowmap.p, vals, u, fext, R); };
    if (R)
        by_penalty(mt.p, run, k_grid_jacobian3<VT, BX, BY, BZ, 3, true>, k_grid_jacobian3<VT, BX, BY, BZ, 1, true>,
                   k_grid_jacobian3<VT, BX, BY, BZ, 0, true>);
    else
        by_penalty(mt.p, run, k_grid_jacobian3<VT, BX, BY, BZ, 3, false>, k_grid_jacobian3<VT, BX, BY, BZ, 1, false>,
                   k_grid_jacobian3<VT, BX, BY, BZ, 0, false>);
}

template <class VT>
void grid_jacobian3_run(const pa_mesh* m, const Material& mt, const double* z, VT* vals, const double* u,
                        const double* fext, double* R, hipStream_t st) {
    GridArgs g{};
    g.d0 = m->kgrid[0];
    g.d1 = m->kgrid[1];
    g.d2 = m->kgrid[2];
    g.x4 = reinterpret_cast<const double4*>(m->d_coords4.p);
    g.z = z;
    grid_jacobian3_brick<VT, 4, 4, 4>(m, g, mt, vals, u, fext, R, st);
}

}  // namespace

bool grid_residual(const pa_mesh* m, int dpn, const Material& mt, const double* u, const double* z, double body,
                   double scale, const double* fext, double* R, hipStream_t st) {
    if (!m->has_grid) return false;
    if (grid_waves() == 8) grid_residual_w<8>(m, dpn, mt, u, z, body, scale, fext, R, st);
    else grid_residual_w<kGridWaves>(m, dpn, mt, u, z, body, scale, fext, R, st);
    PA_HIP(hipGetLastError());
    return true;
}

template <class VT>
bool grid_jacobian1(const pa_mesh* m, const Material& mt, const double* z, VT* vals, hipStream_t st) {
    if (!m->has_grid) return false;
    if (grid_waves() == 8) grid_jacobian1_w<VT, 8>(m, mt, z, vals, st);
    else grid_jacobian1_w<VT, kGridWaves>(m, mt, z, vals, st);
    PA_HIP(hipGetLastError());
    return true;
}

// PA_GRID_J3=0 (read once) sends the elastic Jacobian of grid meshes to the
// tile kernels instead.
template <class VT>
bool grid_jacobian3(const pa_mesh* m, const Material& mt, const double* z, VT* vals, hipStream_t st) {
    static const bool off = [] {
        const char* e = std::getenv("PA_GRID_J3");
        return e != nullptr && std::atoi(e) == 0;
    }();
    if (!m->has_grid || off) return false;
    grid_jacobian3_run<VT>(m, mt, z, vals, nullptr, nullptr, nullptr, st);
    PA_HIP(hipGetLastError());
    return true;
}

// Elastic value and gradient_u on Kuhn grids in one pass: the residual from
// the blocks (R = K u + f, f the natural-BC load or null).
template <class VT>
bool grid_value_jacobian3(const pa_mesh* m, const Material& mt, const double* u, const double* z,
                          const double* fext, double* R, VT* vals, hipStream_t st) {
    static const bool off = [] {
        const char* e = std::getenv("PA_GRID_J3");
        return e != nullptr && std::atoi(e) == 0;
    }();
    if (!m->has_grid || off) return false;
    grid_jacobian3_run<VT>(m, mt, z, vals, u, fext, R, st);
    PA_HIP(hipGetLastError());
    return true;
}

template bool grid_jacobian1<double>(const pa_mesh*, const Material&, const double*, double*, hipStream_t);
template bool grid_jacobian1<float>(const pa_mesh*, const Material&, const double*, float*, hipStream_t);
template bool grid_jacobian3<double>(const pa_mesh*, const Material&, const double*, double*, hipStream_t);
template bool grid_value_jacobian3<double>(const pa_mesh*, const Material&, const double*, const double*,
                                           const double*, double*, double*, hipStream_t);
template bool grid_value_jacobian3<float>(const pa_mesh*, const Material&, const double*, const double*,
                                          const double*, double*, float*, hipStream_t);
template bool grid_jacobian3<float>(const pa_mesh*, const Material&, const double*, float*, hipStream_t);

}  // namespace pa
