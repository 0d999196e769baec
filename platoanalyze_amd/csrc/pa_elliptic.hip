// Node-ordered residual / Jacobian / dR/dz assembly on linear tets for SIMP
// elastostatics and steady thermal diffusion (gfx950).
//
// Work decomposition: one group of G lanes per graph row (= mesh node).  The
// lanes of a group walk the cells incident to the row node (the node -> cell
// incidence built once per mesh), each lane evaluating one cell: gather the
// four nodes' coordinates, densities and states, the Tet4 gradients and
// volume, and the row node's 3x3 (or 1x1) blocks against the cell's four
// nodes.  Blocks are summed into an LDS copy of the row (slots precomputed per
// (row, cell) pair), so every matrix value is written to HBM exactly once,
// with contiguous stores, and without global atomics.  The arithmetic
// restates, per cell, the reference's chain:
//   ComputeGradientWorkset<3>   src/ImplicitFunctors.hpp:169-221
//   Strain / LinearStress       src/Strain.hpp:56-95, src/LinearStress.hpp:102-120
//   ApplyWeighting + MSIMP      src/ApplyWeighting.hpp:48-65, src/Simp.hpp:130-134
//   StressDivergence            src/StressDivergence.hpp:70-90
//   ScalarGrad / ThermalFlux / FluxDivergence  src/ScalarGrad.hpp:58-74,
//                               src/ThermalFlux.hpp:42-52, src/FluxDivergence.hpp:31-48
//   BodyLoad (x mean density)   src/BodyLoads.hpp:103-121
// with the isotropic B^T C B contraction in closed form (lambda, mu).
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "pa_cell.h"
#include "pa_internal.h"

// Partial sums of pa_scalar_function_value_noalloc: one per workgroup of the
// cell kernel (at most kNoAllocPartials), kept on each GPU until the process
// ends.  Host code uses it only through the address hipGetSymbolAddress
// returns for the current device.
constexpr int kNoAllocPartials = 1024;
static __device__ __align__(256) double scratch[kNoAllocPartials];

namespace pa {
namespace {

constexpr int kBlock = 256;

struct RowMesh {
    int N;
    const double4* __restrict__ x4;  // coords padded to 32 bytes
    const int4* __restrict__ conn;
    const int* __restrict__ rowmap;
    const int* __restrict__ n2c_off;
    const int2* __restrict__ pairs;
    // row patterns
    const int* __restrict__ pcell;
    const int* __restrict__ row_pat;
    const int4* __restrict__ pat_info;
    const unsigned char* __restrict__ pat_a;
    const unsigned char* __restrict__ pat_ent;
    const unsigned* __restrict__ pat_item;
    const int4* __restrict__ pat_delta;
    const uint4* __restrict__ pat_item2;
};

// Tet4 shape-function gradients and |det J| (ComputeGradientWorkset<3>).
// One correctly rounded reciprocal of det J replaces the reference's nine
// divisions (each gradient entry differs from cofactor/det by at most 1 ulp).
__device__ inline double tet_gradients(const double4 x[4], double g[4][3]) {
    const double j11 = x[0].x - x[3].x, j12 = x[1].x - x[3].x, j13 = x[2].x - x[3].x;
    const double j21 = x[0].y - x[3].y, j22 = x[1].y - x[3].y, j23 = x[2].y - x[3].y;
    const double j31 = x[0].z - x[3].z, j32 = x[1].z - x[3].z, j33 = x[2].z - x[3].z;
    const double det = j11 * j22 * j33 + j12 * j23 * j31 + j13 * j21 * j32 - j11 * j23 * j32 - j12 * j21 * j33 -
                       j13 * j22 * j31;
    const double id = 1.0 / det;
    g[0][0] = (j22 * j33 - j23 * j32) * id;
    g[0][1] = (j13 * j32 - j12 * j33) * id;
    g[0][2] = (j12 * j23 - j13 * j22) * id;
    g[1][0] = (j23 * j31 - j21 * j33) * id;
    g[1][1] = (j11 * j33 - j13 * j31) * id;
    g[1][2] = (j13 * j21 - j11 * j23) * id;
    g[2][0] = (j21 * j32 - j22 * j31) * id;
    g[2][1] = (j12 * j31 - j11 * j32) * id;
    g[2][2] = (j11 * j22 - j12 * j21) * id;
    g[3][0] = -(g[0][0] + g[1][0] + g[2][0]);
    g[3][1] = -(g[0][1] + g[1][1] + g[2][1]);
    g[3][2] = -(g[0][2] + g[1][2] + g[2][2]);
    return fabs(det);
}

// Everything one lane needs about one incident cell.
struct CellEval {
    int nodes[4];
    int a;          // local index of the row node
    unsigned slots; // packed column slots of the four nodes in the row
    double g[4][3];
    double absdet;  // |det J| = 6 V
    double rho;     // mean nodal density
};

__device__ inline void load_cell(const RowMesh& m, const double* __restrict__ z, int pair, CellEval& c) {
    const int2 pr = m.pairs[pair];
    const int cell = pr.x >> 2;
    c.a = pr.x & 3;
    c.slots = (unsigned)pr.y;
    const int4 nd = m.conn[cell];
    c.nodes[0] = nd.x;
    c.nodes[1] = nd.y;
    c.nodes[2] = nd.z;
    c.nodes[3] = nd.w;
    double4 x[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) x[n] = m.x4[c.nodes[n]];
    c.absdet = tet_gradients(x, c.g);
    c.rho = (((z[c.nodes[0]] + z[c.nodes[1]]) + z[c.nodes[2]]) + z[c.nodes[3]]) / 4.0;
}

// Voigt strain (xx, yy, zz, yz, xz, xy; engineering shear) of a cell.
__device__ inline void cell_strain(const CellEval& c, const double* __restrict__ u, double eps[6]) {
    double un[4][3];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        un[n][0] = u[3L * c.nodes[n] + 0];
        un[n][1] = u[3L * c.nodes[n] + 1];
        un[n][2] = u[3L * c.nodes[n] + 2];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double s = 0.0;
#pragma unroll
        for (int n = 0; n < 4; ++n) s += un[n][d] * c.g[n][d];
        eps[d] = s;
    }
    // (1,2) -> 3, (0,2) -> 4, (0,1) -> 5
    const int pi[3] = {1, 0, 0}, pj[3] = {2, 2, 1};
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const int di = pi[v], dj = pj[v];
        double s = 0.0;
#pragma unroll
        for (int n = 0; n < 4; ++n) s += (un[n][dj] * c.g[n][di] + un[n][di] * c.g[n][dj]);
        eps[3 + v] = s;
    }
}

__device__ inline void iso_stress(const Material& mt, const double eps[6], double sig[6]) {
    const double tr = eps[0] + eps[1] + eps[2];
    const double twomu = 2.0 * mt.mu;
    sig[0] = mt.lambda * tr + twomu * eps[0];
    sig[1] = mt.lambda * tr + twomu * eps[1];
    sig[2] = mt.lambda * tr + twomu * eps[2];
    sig[3] = mt.mu * eps[3];
    sig[4] = mt.mu * eps[4];
    sig[5] = mt.mu * eps[5];
}

// B_b^T sigma: row i gets sum_j sigma(voigt(i,j)) * g_b[j].
__device__ inline void btsig(const double sig[6], const double gb[3], double out[3]) {
    out[0] = sig[0] * gb[0] + sig[5] * gb[1] + sig[4] * gb[2];
    out[1] = sig[5] * gb[0] + sig[1] * gb[1] + sig[3] * gb[2];
    out[2] = sig[4] * gb[0] + sig[3] * gb[1] + sig[2] * gb[2];
}

// 3x3 stiffness block w * B_a^T C B_b of an isotropic material.
__device__ inline void iso_block(const Material& mt, double w, const double ga[3], const double gb[3], double K[9]) {
    const double wl = w * mt.lambda, wm = w * mt.mu;
    const double dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int n = 0; n < 3; ++n) K[i * 3 + n] = wl * ga[i] * gb[n] + wm * ga[n] * gb[i];
    K[0] += wm * dot;
    K[4] += wm * dot;
    K[8] += wm * dot;
}

// Row-node gradient g[a] with a known only at run time.  Bit masks rather
// than a select chain: the optimiser folds selects between elements of one
// array back into a dynamic index, which forces the array into scratch.
__device__ inline void pick_row(const double g[4][3], int a, double out[3]) {
    const unsigned long long m0 = 0ULL - (unsigned long long)(a == 0), m1 = 0ULL - (unsigned long long)(a == 1);
    const unsigned long long m2 = 0ULL - (unsigned long long)(a == 2), m3 = 0ULL - (unsigned long long)(a == 3);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const unsigned long long b = (__double_as_longlong(g[0][d]) & m0) | (__double_as_longlong(g[1][d]) & m1) |
                                     (__double_as_longlong(g[2][d]) & m2) | (__double_as_longlong(g[3][d]) & m3);
        out[d] = __longlong_as_double((long long)b);
    }
}

template <int G>
__device__ inline double group_sum(double v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}

__device__ inline void lds_add(double* p, double v) { atomicAdd(p, v); }

// ---------------------------------------------------------------------------
// Residual rows: R[row] = sum over incident cells of the row node's cell
// residual (+ thermal body load).  Also serves criterion gradient_u (scale 2
// for the thermal energy, no loads).
// ---------------------------------------------------------------------------
template <int DPN, int G>
__global__ __launch_bounds__(kBlock) void k_residual_rows(RowMesh m, Material mt, const double* __restrict__ u,
                                                          const double* __restrict__ z, double body, double scale,
                                                          double* __restrict__ R) {
    constexpr int ROWS = kBlock / G;
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    const int row = xcd_block() * ROWS + grp;
    if (row >= m.N) return;  // whole group exits together
    const int p0 = m.n2c_off[row], p1 = m.n2c_off[row + 1];
    double r[DPN];
#pragma unroll
    for (int i = 0; i < DPN; ++i) r[i] = 0.0;
    for (int p = p0 + lane; p < p1; p += G) {
        CellEval c;
        load_cell(m, z, p, c);
        const double s = penal(c.rho, mt.p, mt.emin);
        const double w = s * (c.absdet / 6.0);
        double ga[3];
        pick_row(c.g, c.a, ga);
        if constexpr (DPN == 3) {
            double eps[6], sig[6], out[3];
            cell_strain(c, u, eps);
            iso_stress(mt, eps, sig);
            btsig(sig, ga, out);
            r[0] += w * out[0];
            r[1] += w * out[1];
            r[2] += w * out[2];
        } else {
            double gT[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                const double T = u[c.nodes[n]];
                gT[0] += T * c.g[n][0];
                gT[1] += T * c.g[n][1];
                gT[2] += T * c.g[n][2];
            }
            r[0] += w * mt.k * (gT[0] * ga[0] + gT[1] * ga[1] + gT[2] * ga[2]);
            // body load, scale -1, weight (1/6)|detJ|, basis 1/4, unpenalised density
            r[0] += -1.0 * (c.absdet / 6.0) * body * 0.25 * c.rho;
        }
    }
#pragma unroll
    for (int i = 0; i < DPN; ++i) r[i] = group_sum<G>(r[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < DPN; ++i) R[(long)row * DPN + i] = scale * r[i];
    }
}

// ---------------------------------------------------------------------------
// Jacobian rows (optionally with the residual in the same pass).
// TR: write the transposed operator (block (i,j) = K_ji^T).
// ---------------------------------------------------------------------------
template <int DPN, int G, int KMAX, class VT, bool RES, bool TR>
__global__ __launch_bounds__(kBlock) void k_jacobian_rows(RowMesh m, Material mt, const double* __restrict__ u,
                                                          const double* __restrict__ z, double body,
                                                          VT* __restrict__ vals, double* __restrict__ R) {
    constexpr int BS = DPN * DPN;
    constexpr int ROWS = kBlock / G;
    __shared__ double acc[ROWS][KMAX * BS];
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    const int row = xcd_block() * ROWS + grp;
    const bool active = row < m.N;
    int rs = 0, rl = 0, p0 = 0, p1 = 0;
    if (active) {
        rs = m.rowmap[row];
        rl = m.rowmap[row + 1] - rs;
        p0 = m.n2c_off[row];
        p1 = m.n2c_off[row + 1];
    }
    for (int t = lane; t < rl * BS; t += G) acc[grp][t] = 0.0;
    __syncthreads();
    double r[DPN];
#pragma unroll
    for (int i = 0; i < DPN; ++i) r[i] = 0.0;
    for (int p = p0 + lane; p < p1; p += G) {
        CellEval c;
        load_cell(m, z, p, c);
        const double s = penal(c.rho, mt.p, mt.emin);
        const double w = s * (c.absdet / 6.0);
        double ga[3];
        pick_row(c.g, c.a, ga);
        double* racc = acc[grp];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int slot = (c.slots >> (8 * b)) & 0xff;
            if constexpr (DPN == 3) {
                double K[9];
                if constexpr (!TR) {
                    iso_block(mt, w, ga, c.g[b], K);
#pragma unroll
                    for (int q = 0; q < 9; ++q) lds_add(&racc[slot * 9 + q], K[q]);
                } else {
                    iso_block(mt, w, c.g[b], ga, K);
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int n = 0; n < 3; ++n) lds_add(&racc[slot * 9 + i * 3 + n], K[n * 3 + i]);
                }
            } else {
                const double* gb = c.g[b];
                const double kv = TR ? w * mt.k * (gb[0] * ga[0] + gb[1] * ga[1] + gb[2] * ga[2])
                                     : w * mt.k * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
                lds_add(&racc[slot], kv);
            }
        }
        if constexpr (RES) {
            if constexpr (DPN == 3) {
                double eps[6], sig[6], out[3];
                cell_strain(c, u, eps);
                iso_stress(mt, eps, sig);
                btsig(sig, ga, out);
                r[0] += w * out[0];
                r[1] += w * out[1];
                r[2] += w * out[2];
            } else {
                double gT[3] = {0.0, 0.0, 0.0};
#pragma unroll
                for (int n = 0; n < 4; ++n) {
                    const double T = u[c.nodes[n]];
                    gT[0] += T * c.g[n][0];
                    gT[1] += T * c.g[n][1];
                    gT[2] += T * c.g[n][2];
                }
                r[0] += w * mt.k * (gT[0] * ga[0] + gT[1] * ga[1] + gT[2] * ga[2]);
                r[0] += -1.0 * (c.absdet / 6.0) * body * 0.25 * c.rho;
            }
        }
    }
    __syncthreads();
    if (active) {
        VT* dst = vals + (long)rs * BS;
        for (int t = lane; t < rl * BS; t += G) dst[t] = (VT)acc[grp][t];
    }
    if constexpr (RES) {
#pragma unroll
        for (int i = 0; i < DPN; ++i) r[i] = group_sum<G>(r[i]);
        if (active && lane == 0) {
#pragma unroll
            for (int i = 0; i < DPN; ++i) R[(long)row * DPN + i] = r[i];
        }
    }
}

// ---------------------------------------------------------------------------
// Jacobian rows through row patterns, without atomics (32 lanes per row).
//  phase 1: lane p < npairs evaluates incident cell p: gradients g, weight
//           w = s(rho) |det J| / 6; parks (w g_a, g_0..g_3) in LDS; with RES
//           also the row node's residual (group sum).
//  phase 2: lane t < nitems sums M = sum (w g_a) (x) g_b over its contribution
//           entries; an off-diagonal slot is then final and stored as
//           K = lambda M + mu M^T + mu tr(M) I (thermal: k tr(M)); diagonal
//           chunks park M in LDS.
//  phase 3: lanes q < block size add the diagonal chunks and store entry q.
// ---------------------------------------------------------------------------
template <int DPN>
__device__ inline double block_entry(const Material& mt, const double* M, int q) {
    if constexpr (DPN == 3) {
        const int i = q / 3, n = q % 3;
        double v = mt.lambda * M[i * 3 + n] + mt.mu * M[n * 3 + i];
        if (i == n) v += mt.mu * ((M[0] + M[4]) + M[8]);
        return v;
    } else {
        return mt.k * M[0];
    }
}

// (x, y, z, density) of a node: the padded coordinates with the density.
__device__ inline double4 node_rec(const double4* __restrict__ x4, const double* __restrict__ z, int n) {
    double4 r = x4[n];
    r.w = z[n];
    return r;
}

// Residual rows through row patterns: one incident cell per lane (32 lanes per
// row), nodes from the pattern offsets; also serves criterion gradient_u (scale 2 for the thermal
// energy, no loads).
template <int DPN>
__global__ __launch_bounds__(kBlock) void k_residual_pattern(RowMesh m, const double* __restrict__ z, Material mt,
                                                             const double* __restrict__ u, double body, double scale,
                                                             double* __restrict__ R) {
    constexpr int G = 32, ROWS = kBlock / G;
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    const int row = xcd_block() * ROWS + grp;
    if (row >= m.N) return;  // whole group exits together
    const int pid = m.row_pat[row];
    const int npairs = m.pat_info[2 * pid].x;
    const int a_off = m.pat_info[2 * pid + 1].x;
    double r[DPN];
#pragma unroll
    for (int i = 0; i < DPN; ++i) r[i] = 0.0;
    if (lane < npairs) {
        CellEval c;
        const int4 dl = m.pat_delta[a_off + lane];
        c.a = m.pat_a[a_off + lane];
        c.nodes[0] = row + dl.x;
        c.nodes[1] = row + dl.y;
        c.nodes[2] = row + dl.z;
        c.nodes[3] = row + dl.w;
        double4 x[4];
#pragma unroll
        for (int n = 0; n < 4; ++n) x[n] = node_rec(m.x4, z, c.nodes[n]);
        c.absdet = tet_gradients(x, c.g);
        c.rho = (((x[0].w + x[1].w) + x[2].w) + x[3].w) / 4.0;
        const double s = penal(c.rho, mt.p, mt.emin);
        const double w = s * (c.absdet / 6.0);
        double ga[3];
        pick_row(c.g, c.a, ga);
        if constexpr (DPN == 3) {
            double eps[6], sig[6], out[3];
            cell_strain(c, u, eps);
            iso_stress(mt, eps, sig);
            btsig(sig, ga, out);
            r[0] = w * out[0];
            r[1] = w * out[1];
            r[2] = w * out[2];
        } else {
            double gT[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                const double T = u[c.nodes[n]];
                gT[0] += T * c.g[n][0];
                gT[1] += T * c.g[n][1];
                gT[2] += T * c.g[n][2];
            }
            r[0] = w * mt.k * (gT[0] * ga[0] + gT[1] * ga[1] + gT[2] * ga[2]);
            r[0] += -1.0 * (c.absdet / 6.0) * body * 0.25 * c.rho;
        }
    }
#pragma unroll
    for (int i = 0; i < DPN; ++i) r[i] = group_sum<G>(r[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < DPN; ++i) R[(long)row * DPN + i] = scale * r[i];
    }
}

template <int DPN, class VT, bool RES, int PMAX>
__global__ __launch_bounds__(kBlock) void k_jacobian_pattern(RowMesh m, const double* __restrict__ z, Material mt,
                                                             const double* __restrict__ u, double body,
                                                             VT* __restrict__ vals, double* __restrict__ R) {
    constexpr int G = 32, ROWS = kBlock / G;
    constexpr int MS = DPN == 3 ? 9 : 1;  // accumulator entries (block size)
    // per pair: w g_a (0..2), g_0..g_3 (3..14); the odd stride of 15 doubles
    // spreads lanes reading different pairs over the LDS banks
    constexpr int PD = 15;
    static_assert(PMAX * PD >= 32 * MS, "the pair buffer also stages the output row");
    __shared__ double pd[ROWS][PMAX * PD];
    __shared__ double dg[ROWS][4][MS];          // diagonal chunk partial sums
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    const int row = xcd_block() * ROWS + grp;
    const bool active = row < m.N;
    int4 i0 = make_int4(0, 0, 0, 0), i1 = make_int4(0, 0, 0, 0);
    int rs = 0, rl = 0;
    if (active) {
        const int pid = m.row_pat[row];
        rs = m.rowmap[row];
        i0 = m.pat_info[2 * pid];
        i1 = m.pat_info[2 * pid + 1];
        rl = i0.y;
    }
    const int npairs = i0.x, nitems = i0.z, diag = i0.w;
    // the work item of this lane with its contribution entries
    uint4 item = make_uint4(0u, 0u, 0u, 0u);
    if (lane < nitems) item = m.pat_item2[i1.z + lane];
    double r[DPN];
#pragma unroll
    for (int i = 0; i < DPN; ++i) r[i] = 0.0;
    if (lane < npairs) {
        CellEval c;
        // the cell's nodes are the row node plus the pattern's offsets
        const int4 dl = m.pat_delta[i1.x + lane];
        c.a = m.pat_a[i1.x + lane];
        c.nodes[0] = row + dl.x;
        c.nodes[1] = row + dl.y;
        c.nodes[2] = row + dl.z;
        c.nodes[3] = row + dl.w;
        double4 x[4];  // (x, y, z, density) records
#pragma unroll
        for (int n = 0; n < 4; ++n) x[n] = node_rec(m.x4, z, c.nodes[n]);
        c.absdet = tet_gradients(x, c.g);
        c.rho = (((x[0].w + x[1].w) + x[2].w) + x[3].w) / 4.0;
        const double s = penal(c.rho, mt.p, mt.emin);
        const double w = s * (c.absdet / 6.0);
        double ga[3];
        pick_row(c.g, c.a, ga);
        double* P = &pd[grp][lane * PD];
        P[0] = w * ga[0];
        P[1] = w * ga[1];
        P[2] = w * ga[2];
#pragma unroll
        for (int n = 0; n < 4; ++n)
#pragma unroll
            for (int d = 0; d < 3; ++d) P[3 + 3 * n + d] = c.g[n][d];
        if constexpr (RES) {
            if constexpr (DPN == 3) {
                double eps[6], sig[6], out[3];
                cell_strain(c, u, eps);
                iso_stress(mt, eps, sig);
                btsig(sig, ga, out);
                r[0] = w * out[0];
                r[1] = w * out[1];
                r[2] = w * out[2];
            } else {
                double gT[3] = {0.0, 0.0, 0.0};
#pragma unroll
                for (int n = 0; n < 4; ++n) {
                    const double T = u[c.nodes[n]];
                    gT[0] += T * c.g[n][0];
                    gT[1] += T * c.g[n][1];
                    gT[2] += T * c.g[n][2];
                }
                r[0] = w * mt.k * (gT[0] * ga[0] + gT[1] * ga[1] + gT[2] * ga[2]);
                r[0] += -1.0 * (c.absdet / 6.0) * body * 0.25 * c.rho;
            }
        }
    }
    if constexpr (RES) {
#pragma unroll
        for (int i = 0; i < DPN; ++i) r[i] = group_sum<G>(r[i]);
        if (active && lane == 0) {
#pragma unroll
            for (int i = 0; i < DPN; ++i) R[(long)row * DPN + i] = r[i];
        }
    }
    __syncthreads();
    const int slot = item.x & 0xff;
    double M[MS];
#pragma unroll
    for (int q = 0; q < MS; ++q) M[q] = 0.0;
    if (lane < nitems) {
        const int count = (item.x >> 8) & 0xff;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if (e >= count) break;
            const unsigned word = e < 4 ? item.y : item.z;
            const int pe = (word >> (8 * (e % 4))) & 0xff;
            const double* P = &pd[grp][(pe >> 2) * PD];
            const double* gb = P + 3 + 3 * (pe & 3);
            const double a0 = P[0], a1 = P[1], a2 = P[2];
            const double b0 = gb[0], b1 = gb[1], b2 = gb[2];
            if constexpr (DPN == 3) {
                M[0] += a0 * b0; M[1] += a0 * b1; M[2] += a0 * b2;
                M[3] += a1 * b0; M[4] += a1 * b1; M[5] += a1 * b2;
                M[6] += a2 * b0; M[7] += a2 * b1; M[8] += a2 * b2;
            } else {
                M[0] += (a0 * b0 + a1 * b1) + a2 * b2;
            }
        }
    }
    __syncthreads();  // the pair buffer now stages the output row
    double* kout = pd[grp];
    if (lane < nitems) {
        if (slot != diag) {
#pragma unroll
            for (int q = 0; q < MS; ++q) kout[slot * MS + q] = block_entry<DPN>(mt, M, q);
        } else {
#pragma unroll
            for (int q = 0; q < MS; ++q) dg[grp][lane - i1.w][q] = M[q];
        }
    }
    __syncthreads();
    if (active && lane < MS) {
        const int nd = npairs > 0 ? (npairs + 7) / 8 : 1;
        double Md[MS];
#pragma unroll
        for (int q = 0; q < MS; ++q) {
            double s = 0.0;
            for (int k = 0; k < nd; ++k) s += dg[grp][k][q];
            Md[q] = s;
        }
        kout[diag * MS + lane] = block_entry<DPN>(mt, Md, lane);
    }
    __syncthreads();
    if (active) {
        VT* dst = vals + (long)rs * MS;
        for (int t = lane; t < rl * MS; t += G) dst[t] = (VT)kout[t];
    }
}

// ---------------------------------------------------------------------------
// Cell-tile assembly: one workgroup per tile, a block of up to 64 rows of the
// node grid (schedules built once per mesh on the host).
//  phase 1: every distinct cell incident to the tile's rows is evaluated
//           once into LDS: the Jacobian kernel parks sqrt(w) g_n (w = s(rho)
//           |det J| / 6), so a block is M = sum g_a (x) g_b over the cells
//           of its edge; the residual kernel (which first stages the tile's
//           node halo in LDS) parks the cell's residuals of the tile's rows.
//  phase 2: work units (row, column slot) sum their block from LDS and store
//           it; 8 lanes per row sum the residual (the addition tree of a
//           32-lane butterfly over the incident cells).
// Each cell is evaluated about twice (once per tile touching it) instead of
// four times (once per node) as in the row kernels.
// ---------------------------------------------------------------------------
struct TileMesh {
    const int4* __restrict__ tile;  // 3 per tile
    const int4* __restrict__ rows;
    const int4* __restrict__ runs;
    const int4* __restrict__ cells;
    const uint2* __restrict__ cslots;
    const unsigned* __restrict__ hslots;
    const uint2* __restrict__ units;
    const uint4* __restrict__ ents;
    const int* __restrict__ rowmap;
    int d0, d1, d2;  // node grid (x, y, z counts; z fastest)
};

// A tile's first row, the offsets of its schedule's tables and their sizes.
struct TileHead {
    int base, row_off, cell_off, unit_off, nrows, ncells, nunits, nruns, run_off;
};

__device__ inline TileHead tile_head(const TileMesh& t) {
    const int b = xcd_block();
    const int4 a = t.tile[3 * b], c = t.tile[3 * b + 1], e = t.tile[3 * b + 2];
    return TileHead{a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w, e.x};
}

// The tile's halo (kTileHalo nodes: its row block grown by one node on each
// side, z fastest) into LDS: (x, y, z, density) records and, with NU > 0, the
// NU state values of each node.  Halo nodes outside the mesh stay unset; no
// cell refers to them.
template <int NU>
__device__ inline void tile_halo_load(const TileMesh& t, const TileHead& h, const double4* __restrict__ x4,
                                      const double* __restrict__ z, const double* __restrict__ u, double4* hrec,
                                      double* hu) {
    const int i0 = h.base / (t.d1 * t.d2), j0 = (h.base / t.d2) % t.d1, k0 = h.base % t.d2;
    for (int s = threadIdx.x; s < kTileHalo; s += blockDim.x) {
        const int a = s / ((kTileY + 2) * (kTileZ + 2)), b = (s / (kTileZ + 2)) % (kTileY + 2), c = s % (kTileZ + 2);
        const int i = i0 - 1 + a, j = j0 - 1 + b, k = k0 - 1 + c;
        if (i < 0 || i >= t.d0 || j < 0 || j >= t.d1 || k < 0 || k >= t.d2) continue;
        const long node = ((long)i * t.d1 + j) * t.d2 + k;
        double4 r = x4[node];
        r.w = z[node];
        hrec[s] = r;
#pragma unroll
        for (int q = 0; q < NU; ++q) hu[s * NU + q] = u[node * NU + q];
    }
}

// Node ids and (x, y, z, density) of tile cell c gathered from global memory
// (the elastic Jacobian: its block phase keeps the LDS busy, so the cell
// records come through the vector memory path instead).
__device__ inline void tile_cell_gather(const TileMesh& t, const double4* __restrict__ x4,
                                        const double* __restrict__ z, const TileHead& h, int c, CellEval& ce,
                                        double4 x[4]) {
    const int4 d = t.cells[h.cell_off + c];
    ce.nodes[0] = h.base + d.x;
    ce.nodes[1] = h.base + d.y;
    ce.nodes[2] = h.base + d.z;
    ce.nodes[3] = h.base + d.w;
#pragma unroll
    for (int n = 0; n < 4; ++n) x[n] = node_rec(x4, z, ce.nodes[n]);
}

// (x, y, z, density) of tile cell c from the halo; ce.nodes holds the halo
// slots of its nodes (the index of their state values in the halo).
__device__ inline void tile_cell_load(const TileMesh& t, const double4* hrec, const TileHead& h, int c,
                                      CellEval& ce, double4 x[4]) {
    const unsigned hs = t.hslots[h.cell_off + c];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        ce.nodes[n] = (int)((hs >> (8 * n)) & 0xffu);
        x[n] = hrec[ce.nodes[n]];
    }
}

__device__ inline void tile_cell_geometry(CellEval& ce, const double4 x[4]) {
    ce.absdet = tet_gradients(x, ce.g);
    ce.rho = (((x[0].w + x[1].w) + x[2].w) + x[3].w) / 4.0;
}

// Phase 1 over the tile's cells, two per thread (a tile has at most
// 2 * kTileThreads cells), with both cells' gathers issued before either is
// used.
template <bool HALO, class F>
__device__ inline void tile_cells(const TileMesh& t, const double4* src, const double* __restrict__ z,
                                  const TileHead& h, F&& eval) {
    static_assert(kTileCellsMax <= 2 * kTileThreads, "two cells per thread");
    const int c0 = threadIdx.x, c1 = threadIdx.x + kTileThreads;
    CellEval e0, e1;
    double4 x0[4], x1[4];
    if constexpr (HALO) {
        if (c0 < h.ncells) tile_cell_load(t, src, h, c0, e0, x0);
        if (c1 < h.ncells) tile_cell_load(t, src, h, c1, e1, x1);
    } else {
        if (c0 < h.ncells) tile_cell_gather(t, src, z, h, c0, e0, x0);
        if (c1 < h.ncells) tile_cell_gather(t, src, z, h, c1, e1, x1);
    }
    if (c0 < h.ncells) {
        tile_cell_geometry(e0, x0);
        eval(c0, e0);
    }
    if (c1 < h.ncells) {
        tile_cell_geometry(e1, x1);
        eval(c1, e1);
    }
}

// Residual tile kernel, one cell per thread (kResThreads >= kTileCellsMax).
// Phase 1 writes each cell's contributions to the tile's rows straight into
// their pair slots (pair k of row r at r * 32 + k; contributions to nodes
// outside the tile are dropped), so the LDS holds 32 * DPN doubles per row
// instead of the cells' full residuals and more workgroups fit on a CU.
// Phase 2 sums a row's slots with the 8-lane tree of the row kernels.
// The launch bound asks for 8 waves per SIMD (64 VGPRs, no spills): four
// workgroups per CU instead of three, 10% faster.
constexpr int kResThreads = 512;
static_assert(kTileCellsMax <= kResThreads && kTileX * kTileY * kTileZ * 8 <= kResThreads, "one cell per thread");

template <int DPN>
__global__ __launch_bounds__(kResThreads, 8) void k_tile_residual(TileMesh t, const double4* __restrict__ x4,
                                                               const double* __restrict__ z, Material mt,
                                                               const double* __restrict__ u, double body,
                                                               double scale, int hoff, const double* __restrict__ fext,
                                                               double* __restrict__ R) {
    // LDS: the pair slots (rows x 32 x DPN doubles), then at hoff the halo
    // records and the halo's state values
    extern __shared__ double lds[];
    double4* hrec = reinterpret_cast<double4*>(lds + hoff);
    double* hu = lds + hoff + 4 * kTileHalo;
    const TileHead h = tile_head(t);
    // schedule words fetched ahead: this lane's row of phase 2 and its cell's slots
    const int q = threadIdx.x, r = q >> 3, s = q & 7;
    int4 ri = make_int4(0, 0, 0, 0);
    if (r < h.nrows) ri = t.rows[h.row_off + r];
    const int c = threadIdx.x;
    uint2 sl = make_uint2(0u, 0u);
    if (c < h.ncells) sl = t.cslots[h.cell_off + c];
    tile_halo_load<DPN>(t, h, x4, z, u, hrec, hu);
    __syncthreads();
    if (c < h.ncells) {
        CellEval ce;
        double4 x[4];
        tile_cell_load(t, hrec, h, c, ce, x);
        tile_cell_geometry(ce, x);
        const unsigned slot[4] = {sl.x & 0xffffu, sl.x >> 16, sl.y & 0xffffu, sl.y >> 16};
        const double sp = penal(ce.rho, mt.p, mt.emin);
        const double w = sp * (ce.absdet / 6.0);
        if constexpr (DPN == 3) {
            double eps[6], sig[6], out[3];
            cell_strain(ce, hu, eps);
            iso_stress(mt, eps, sig);
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                if (slot[a] == 0xffffu) continue;
                btsig(sig, ce.g[a], out);
                double* F = lds + slot[a] * 3;
                F[0] = w * out[0];
                F[1] = w * out[1];
                F[2] = w * out[2];
            }
        } else {
            double gT[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                const double T = hu[ce.nodes[n]];
                gT[0] += T * ce.g[n][0];
                gT[1] += T * ce.g[n][1];
                gT[2] += T * ce.g[n][2];
            }
            const double load = -1.0 * (ce.absdet / 6.0) * body * 0.25 * ce.rho;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                if (slot[a] == 0xffffu) continue;
                double v = w * mt.k * (gT[0] * ce.g[a][0] + gT[1] * ce.g[a][1] + gT[2] * ce.g[a][2]);
                v += load;
                lds[slot[a]] = v;
            }
        }
    }
    __syncthreads();
    if (r >= h.nrows) return;
    // lane s: pairs s, s + 16, s + 8, s + 24 (the 32-lane butterfly's tree)
    const int ps[4] = {s, s + 16, s + 8, s + 24};
    double v[4][DPN];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < DPN; ++i) v[j][i] = ps[j] < ri.y ? lds[(r * 32 + ps[j]) * DPN + i] : 0.0;
    double b[DPN];
#pragma unroll
    for (int i = 0; i < DPN; ++i) {
        b[i] = (v[0][i] + v[1][i]) + (v[2][i] + v[3][i]);
#pragma unroll
        for (int o = 4; o > 0; o >>= 1) b[i] += __shfl_xor(b[i], o, 8);
    }
    if (s == 0) {
        const long o = (long)(h.base + ri.x) * DPN;
        if (fext) {  // plus the natural BCs' nodal load
#pragma unroll
            for (int i = 0; i < DPN; ++i) R[o + i] = scale * b[i] + fext[o + i];
        } else {
#pragma unroll
            for (int i = 0; i < DPN; ++i) R[o + i] = scale * b[i];
        }
    }
}

// One work unit (contributions to one block, or to a part of a diagonal
// block) with its entry slots, fetched ahead of use.
struct TileUnitFetch {
    uint2 unit;
    uint4 ew[kTileUnitEntries / 4];
};

__device__ inline void tile_unit_fetch(const TileMesh& t, const TileHead& h, int it, TileUnitFetch& f) {
    f.unit = t.units[h.unit_off + it];
#pragma unroll
    for (int k = 0; k < kTileUnitEntries / 4; ++k) f.ew[k] = t.ents[(f.unit.y & 0x3fffffu) + k];
}

// Slot of the product for node pair (a, b) among the ten a <= b pairs of a
// cell, in the order (0,0) (0,1) (0,2) (0,3) (1,1) (1,2) (1,3) (2,2) (2,3) (3,3).
// The thermal Jacobian parks a cell's ten products at this stride (the
// schedule's entry slots still count cells in kTileStride units).
constexpr int kThermalStride = 10;
__device__ inline unsigned thermal_pair_index(unsigned a, unsigned b) {
    const unsigned lo = a < b ? a : b, hi = a < b ? b : a;
    return lo * 4 - (lo * (lo - 1)) / 2 + (hi - lo);
}

// M = sum over a unit's first NE entry slots of g_a (x) g_b (gradients
// scaled by sqrt(w)); slots beyond the unit's count hold the zero cell, so
// the loop has no branches and its LDS reads pipeline.  SAME: the entries are
// diagonal (g_b = g_a), read once, with the six distinct products mirrored
// (a_i a_n = a_n a_i exactly).
template <int DPN, int NE, bool SAME>
__device__ inline void tile_unit_block(const double* lds, const TileUnitFetch& f, double* M) {
    constexpr int MS = DPN == 3 ? 9 : 1;
#pragma unroll
    for (int q = 0; q < MS; ++q) M[q] = 0.0;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const uint4 w4 = f.ew[e / 4];
        const unsigned en = (e & 3) == 0 ? w4.x : (e & 3) == 1 ? w4.y : (e & 3) == 2 ? w4.z : w4.w;
        if constexpr (DPN == 1) {
            // the entry's slots name g_a and g_b of one cell (cell * kTileStride
            // + 3 a); the cell parked the product of the pair instead
            const unsigned oa = en & 0xffffu, ob = en >> 16;
            const unsigned c = oa / kTileStride, a = (oa - c * kTileStride) / 3, b = (ob - c * kTileStride) / 3;
            M[0] += lds[c * kThermalStride + thermal_pair_index(a, b)];
        } else {
            const double* A = lds + (en & 0xffffu);
            const double a0 = A[0], a1 = A[1], a2 = A[2];
            if constexpr (SAME) {
                M[0] += a0 * a0; M[1] += a0 * a1; M[2] += a0 * a2;
                M[4] += a1 * a1; M[5] += a1 * a2; M[8] += a2 * a2;
            } else {
                const double* B = lds + (en >> 16);
                const double b0 = B[0], b1 = B[1], b2 = B[2];
                M[0] += a0 * b0; M[1] += a0 * b1; M[2] += a0 * b2;
                M[3] += a1 * b0; M[4] += a1 * b1; M[5] += a1 * b2;
                M[6] += a2 * b0; M[7] += a2 * b1; M[8] += a2 * b2;
            }
        }
    }
    if constexpr (SAME && DPN == 3) {
        M[3] = M[1];
        M[6] = M[2];
        M[7] = M[5];
    }
}

// Jacobian of one tile.  Phase 1 parks sqrt(w) g_n per cell (kTileStride
// doubles), from coordinates and densities gathered through the vector memory
// path (the 3 x 3 block phase keeps the LDS busy); thermal: the ten products
// w g_a . g_b per cell, from the tile's node halo staged in LDS (its block
// phase reads one product per entry and leaves the LDS room).  Phase 2: each thread evaluates
// up to three work units; the kTileDiagParts parts of a diagonal block sit on
// adjacent lanes and are summed by a butterfly; blocks are stored as
// lambda M + mu M^T + mu tr(M) I (thermal: k M).
template <int DPN, class VT, int UE>
__global__ __launch_bounds__(kTileThreads) void k_tile_jacobian(TileMesh t, const double4* __restrict__ nodes,
                                                                const double* __restrict__ z, Material mt, int cmax,
                                                                VT* __restrict__ vals) {
    // nodes: the padded coordinates, gathered per cell with z (DPN 3) or
    // read with z into the halo (DPN 1)
    constexpr bool HALO = DPN == 1;
    constexpr int MS = DPN == 3 ? 9 : 1;
    constexpr int NU = kTileUnitsMax / kTileThreads;
    static_assert(kTileStride <= kTileThreads && kTileDiagParts == 4, "zero cell by one pass; 4-lane butterfly");
    extern __shared__ double lds[];
    __shared__ int rstart[64], rstage[64];  // rowmap and staging offset of the tile's rows
    const TileHead h = tile_head(t);
    // the thread's units are fetched while the cells load
    TileUnitFetch f[NU];
#pragma unroll
    for (int k = 0; k < NU; ++k)
        if ((int)threadIdx.x + k * kTileThreads < h.nunits) tile_unit_fetch(t, h, threadIdx.x + k * kTileThreads, f[k]);
    if ((int)threadIdx.x < h.nrows) {
        const int4 ri = t.rows[h.row_off + threadIdx.x];
        rstart[threadIdx.x] = t.rowmap[h.base + ri.x];
        rstage[threadIdx.x] = ri.w;
    }
    constexpr int CS = DPN == 3 ? kTileStride : kThermalStride;  // LDS doubles per cell
    if ((int)threadIdx.x < CS) lds[h.ncells * CS + threadIdx.x] = 0.0;  // the zero cell
    double4* hrec = reinterpret_cast<double4*>(lds + ((cmax * CS + 1) & ~1));  // after the cells, 16-byte aligned
    if constexpr (HALO) {
        tile_halo_load<0>(t, h, nodes, z, nullptr, hrec, nullptr);
        __syncthreads();
    }
    tile_cells<HALO>(t, HALO ? hrec : nodes, z, h, [&](int c, const CellEval& ce) {
        const double s = penal(ce.rho, mt.p, mt.emin);
        const double w = s * (ce.absdet / 6.0);
        double* Q = lds + c * CS;
        if constexpr (DPN == 3) {
            const double sw = sqrt(w);
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                Q[3 * n + 0] = sw * ce.g[n][0];
                Q[3 * n + 1] = sw * ce.g[n][1];
                Q[3 * n + 2] = sw * ce.g[n][2];
            }
        } else {
            // 1 x 1 blocks: the ten products w g_a . g_b (a <= b), one LDS
            // read per unit entry instead of two gradients
            int q = 0;
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = a; b < 4; ++b)
                    Q[q++] = w * ((ce.g[a][0] * ce.g[b][0] + ce.g[a][1] * ce.g[b][1]) + ce.g[a][2] * ce.g[b][2]);
        }
    });
    __syncthreads();
    // blocks are kept in registers, then staged in LDS (the cell data is
    // dead by then) in CRS order and written run by run with contiguous stores
    double K[NU][MS];
    unsigned ux[NU];
    bool put[NU], mirror[NU];
#pragma unroll
    for (int k = 0; k < NU; ++k) put[k] = mirror[k] = false;
#pragma unroll
    for (int k = 0; k < NU; ++k) {
        if (k * kTileThreads >= h.nunits) break;  // uniform
        const int it = threadIdx.x + k * kTileThreads;
        const bool live = it < h.nunits;
        const uint2 u = live ? f[k].unit : make_uint2(0u, 0u);
        double M[MS];
        // the wave shape is uniform over the wave's live lanes
        switch (u.y >> 30) {
            case 1: tile_unit_block<DPN, 6, true>(lds, f[k], M); break;
            case 2: tile_unit_block<DPN, 6, false>(lds, f[k], M); break;
            case 3: tile_unit_block<DPN, 4, false>(lds, f[k], M); break;
            default:
                if (live) {
                    tile_unit_block<DPN, UE, false>(lds, f[k], M);
                } else {
#pragma unroll
                    for (int q = 0; q < MS; ++q) M[q] = 0.0;
                }
        }
        const bool diag = ((u.y >> 26) & 1u) != 0u;
        // parts of a diagonal block: lanes 4j..4j+3, summed (p0 + p1) + (p2 + p3)
#pragma unroll
        for (int q = 0; q < MS; ++q) {
            const double o1 = __shfl_xor(M[q], 1);
            if (diag) M[q] += o1;
            const double o2 = __shfl_xor(M[q], 2);
            if (diag) M[q] += o2;
        }
#pragma unroll
        for (int q = 0; q < MS; ++q) K[k][q] = block_entry<DPN>(mt, M, q);
        ux[k] = u.x;
        put[k] = live && (!diag || ((u.y >> 27) & 3u) == 0u);
        mirror[k] = ((u.y >> 29) & 1u) != 0u;
    }
    __syncthreads();
    VT* stage = reinterpret_cast<VT*>(lds);
#pragma unroll
    for (int k = 0; k < NU; ++k) {
        if (!put[k]) continue;
        VT* dst = stage + (rstage[ux[k] & 0xffu] + ((ux[k] >> 8) & 0xffu)) * MS;
#pragma unroll
        for (int q = 0; q < MS; ++q) dst[q] = (VT)K[k][q];
        if (mirror[k]) {  // the mirror block K_ji = K_ij^T
            VT* dm = stage + (rstage[(ux[k] >> 16) & 0xffu] + (ux[k] >> 24)) * MS;
            if constexpr (DPN == 3) {
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int n = 0; n < 3; ++n) dm[3 * i + n] = (VT)K[k][3 * n + i];
            } else {
                dm[0] = (VT)K[k][0];
            }
        }
    }
    __syncthreads();
    for (int j = 0; j < h.nruns; ++j) {
        const int4 run = t.runs[h.run_off + j];
        VT* dst = vals + (long)rstart[run.x] * MS;
        const VT* src = stage + run.y * MS;
        const int len = run.z * MS;
        for (int i = threadIdx.x; i < len; i += kTileThreads) dst[i] = src[i];
    }
}

// ---------------------------------------------------------------------------
// gradient_z rows: block (row = control node a, col = state node b) holds
// dR_b/dz_a (1 x DPN) = sum over cells with a, b of (1/4) s'(rho) (V) B_b^T C eps
// (+ thermal body load derivative).
// ---------------------------------------------------------------------------
template <int DPN, int G, int KMAX>
__global__ __launch_bounds__(kBlock) void k_gradz_rows(RowMesh m, Material mt, const double* __restrict__ u,
                                                       const double* __restrict__ z, double body,
                                                       double* __restrict__ vals) {
    constexpr int ROWS = kBlock / G;
    __shared__ double acc[ROWS][KMAX * DPN];
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    const int row = xcd_block() * ROWS + grp;
    const bool active = row < m.N;
    int rs = 0, rl = 0, p0 = 0, p1 = 0;
    if (active) {
        rs = m.rowmap[row];
        rl = m.rowmap[row + 1] - rs;
        p0 = m.n2c_off[row];
        p1 = m.n2c_off[row + 1];
    }
    for (int t = lane; t < rl * DPN; t += G) acc[grp][t] = 0.0;
    __syncthreads();
    for (int p = p0 + lane; p < p1; p += G) {
        CellEval c;
        load_cell(m, z, p, c);
        const double ds = dpenal(c.rho, mt.p, mt.emin) * 0.25;
        const double w = ds * (c.absdet / 6.0);
        double* racc = acc[grp];
        if constexpr (DPN == 3) {
            double eps[6], sig[6];
            cell_strain(c, u, eps);
            iso_stress(mt, eps, sig);
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int slot = (c.slots >> (8 * b)) & 0xff;
                double out[3];
                btsig(sig, c.g[b], out);
                lds_add(&racc[slot * 3 + 0], w * out[0]);
                lds_add(&racc[slot * 3 + 1], w * out[1]);
                lds_add(&racc[slot * 3 + 2], w * out[2]);
            }
        } else {
            double gT[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                const double T = u[c.nodes[n]];
                gT[0] += T * c.g[n][0];
                gT[1] += T * c.g[n][1];
                gT[2] += T * c.g[n][2];
            }
            const double bl = -1.0 * (c.absdet / 6.0) * body * 0.25 * 0.25;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int slot = (c.slots >> (8 * b)) & 0xff;
                const double* gb = c.g[b];
                lds_add(&racc[slot], w * mt.k * (gT[0] * gb[0] + gT[1] * gb[1] + gT[2] * gb[2]) + bl);
            }
        }
    }
    __syncthreads();
    if (active) {
        double* dst = vals + (long)rs * DPN;
        for (int t = lane; t < rl * DPN; t += G) dst[t] = acc[grp][t];
    }
}

// ---------------------------------------------------------------------------
// gradient_x rows (shape derivatives): block (row = node a whose position
// moves, col = state node b) holds dR_b/dx_a, 3 x DPN, entry [i][j] =
// dR_{b,j}/dx_{a,i}.  With g_c = grad N_c and V the cell volume,
//   d g_c / dx_{a,i} = -g_a g_{c,i},   d V / dx_{a,i} = V g_{a,i},
// so with H = grad u (H_dn = sum_c u_{c,d} g_{c,n}) and sigma0 = C eps(H):
//   elastic: w [ g_{a,i} (sigma0 g_b)_j - g_{b,i} (sigma0 g_a)_j
//                + (C deps_i g_b)_j ],  deps_i = sym(-H e_i (x) g_a),
//   thermal: w k [ g_{a,i} (gT.g_b) - gT_i (g_a.g_b) - g_{b,i} (gT.g_a) ]
//            - V g_{a,i} f rho / 4 (body source),
// w = s(rho) V.  The reference obtains the same numbers by forward-mode
// differentiation in the cell coordinates (GradientX evaluation type).
// ---------------------------------------------------------------------------
__device__ inline void grad_u_cell(const CellEval& c, const double* __restrict__ u, double H[3][3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int n = 0; n < 3; ++n) H[d][n] = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double u0 = u[3L * c.nodes[k] + 0], u1 = u[3L * c.nodes[k] + 1], u2 = u[3L * c.nodes[k] + 2];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            H[0][n] += u0 * c.g[k][n];
            H[1][n] += u1 * c.g[k][n];
            H[2][n] += u2 * c.g[k][n];
        }
    }
}

// Voigt strain (engineering shear) of a displacement gradient.
__device__ inline void voigt_strain(const double H[3][3], double eps[6]) {
    eps[0] = H[0][0];
    eps[1] = H[1][1];
    eps[2] = H[2][2];
    eps[3] = H[1][2] + H[2][1];
    eps[4] = H[0][2] + H[2][0];
    eps[5] = H[0][1] + H[1][0];
}

template <int DPN, int G, int KMAX>
__global__ __launch_bounds__(kBlock) void k_gradx_rows(RowMesh m, Material mt, const double* __restrict__ u,
                                                       const double* __restrict__ z, double body,
                                                       double* __restrict__ vals) {
    constexpr int ROWS = kBlock / G, BS = 3 * DPN;
    __shared__ double acc[ROWS][KMAX * BS];
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    const int row = xcd_block() * ROWS + grp;
    const bool active = row < m.N;
    int rs = 0, rl = 0, p0 = 0, p1 = 0;
    if (active) {
        rs = m.rowmap[row];
        rl = m.rowmap[row + 1] - rs;
        p0 = m.n2c_off[row];
        p1 = m.n2c_off[row + 1];
    }
    for (int t = lane; t < rl * BS; t += G) acc[grp][t] = 0.0;
    __syncthreads();
    for (int p = p0 + lane; p < p1; p += G) {
        CellEval c;
        load_cell(m, z, p, c);
        const double V = c.absdet / 6.0;
        const double w = penal(c.rho, mt.p, mt.emin) * V;
        double ga[3];
        pick_row(c.g, c.a, ga);
        double* racc = acc[grp];
        if constexpr (DPN == 3) {
            double H[3][3], eps[6], sig[6], sa[3];
            grad_u_cell(c, u, H);
            voigt_strain(H, eps);
            iso_stress(mt, eps, sig);
            btsig(sig, ga, sa);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                double dH[3][3], de[6], ds[6];
#pragma unroll
                for (int d = 0; d < 3; ++d)
#pragma unroll
                    for (int n = 0; n < 3; ++n) dH[d][n] = -H[d][i] * ga[n];
                voigt_strain(dH, de);
                iso_stress(mt, de, ds);
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int slot = (c.slots >> (8 * b)) & 0xff;
                    double sb[3], db[3];
                    btsig(sig, c.g[b], sb);
                    btsig(ds, c.g[b], db);
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        lds_add(&racc[slot * BS + i * 3 + j], w * ((ga[i] * sb[j] - c.g[b][i] * sa[j]) + db[j]));
                }
            }
        } else {
            double gT[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                const double T = u[c.nodes[n]];
                gT[0] += T * c.g[n][0];
                gT[1] += T * c.g[n][1];
                gT[2] += T * c.g[n][2];
            }
            const double wk = w * mt.k;
            const double ta = gT[0] * ga[0] + gT[1] * ga[1] + gT[2] * ga[2];
            const double bl = -1.0 * V * body * 0.25 * c.rho;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int slot = (c.slots >> (8 * b)) & 0xff;
                const double* gb = c.g[b];
                const double tb = gT[0] * gb[0] + gT[1] * gb[1] + gT[2] * gb[2];
                const double ab = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    lds_add(&racc[slot * BS + i], wk * ((ga[i] * tb - gT[i] * ab) - gb[i] * ta) + bl * ga[i]);
            }
        }
    }
    __syncthreads();
    if (active) {
        double* dst = vals + (long)rs * BS;
        for (int t = lane; t < rl * BS; t += G) dst[t] = acc[grp][t];
    }
}

// Shape derivative of a uniform natural BC, added to the gradient_x blocks of
// the boundary rows: R_b += (1/3) v scale A_f over the faces f holding b, and
// for face (A, B, C) with unit normal n of (B - A) x (C - A) the area moves as
// dA/dx_A = 1/2 (x_B - x_C) x n (cyclic for B, C).  One thread per boundary
// node a; it alone updates row a, after the volume kernel.
template <int DPN>
__global__ void k_gradx_natural_bc(int nb, const int* __restrict__ bnodes, const int* __restrict__ boff,
                                   const int* __restrict__ bface, const int* __restrict__ fcells,
                                   const int* __restrict__ flnodes, const int4* __restrict__ conn,
                                   const double4* __restrict__ x4, const int* __restrict__ rowmap,
                                   const int* __restrict__ colidx, double v0, double v1, double v2, double scale,
                                   double* __restrict__ vals) {
    constexpr int BS = 3 * DPN;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb) return;
    const int a = bnodes[t];
    const int r0 = rowmap[a], r1 = rowmap[a + 1];
    const double v[3] = {v0, v1, v2};
    const double basis = 1.0 / 3.0;
    for (int k = boff[t]; k < boff[t + 1]; ++k) {
        const int f = bface[k] >> 2, pos = bface[k] & 3;
        const int4 nd = conn[fcells[f]];
        const int cn[4] = {nd.x, nd.y, nd.z, nd.w};
        const int fn[3] = {cn[flnodes[3 * f + 0]], cn[flnodes[3 * f + 1]], cn[flnodes[3 * f + 2]]};
        const double4 A = x4[fn[0]], B = x4[fn[1]], C = x4[fn[2]];
        const double e1[3] = {B.x - A.x, B.y - A.y, B.z - A.z}, e2[3] = {C.x - A.x, C.y - A.y, C.z - A.z};
        double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const double inv = 1.0 / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        n[0] *= inv;
        n[1] *= inv;
        n[2] *= inv;
        // edge opposite to this node, oriented cyclically: x_next - x_prev
        const double4 P = pos == 0 ? B : (pos == 1 ? C : A), Q = pos == 0 ? C : (pos == 1 ? A : B);
        const double q[3] = {P.x - Q.x, P.y - Q.y, P.z - Q.z};
        const double dA[3] = {0.5 * (q[1] * n[2] - q[2] * n[1]), 0.5 * (q[2] * n[0] - q[0] * n[2]),
                              0.5 * (q[0] * n[1] - q[1] * n[0])};
        for (int b = 0; b < 3; ++b) {
            int e = r0;
            while (e < r1 && colidx[e] != fn[b]) ++e;
            if (e == r1) continue;  // not reached: face nodes share a cell
            double* blk = vals + (long)e * BS;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < DPN; ++j) blk[i * DPN + j] += basis * v[j] * scale * dA[i];
        }
    }
}

// ---------------------------------------------------------------------------
// Internal energy: per-cell values (block partial sums), nodal gradient_z.
// ---------------------------------------------------------------------------
template <int DPN>
__device__ inline double cell_energy_density(const CellEval& c, const Material& mt, const double* __restrict__ u) {
    // unpenalised energy per unit volume weight: 1/2 sigma0 : eps, or k |grad T|^2
    if constexpr (DPN == 3) {
        double eps[6], sig[6];
        cell_strain(c, u, eps);
        iso_stress(mt, eps, sig);
        double inc = 0.0;
#pragma unroll
        for (int i = 0; i < 6; ++i) inc += sig[i] * eps[i];
        return 0.5 * inc;
    } else {
        double gT[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            const double T = u[c.nodes[n]];
            gT[0] += T * c.g[n][0];
            gT[1] += T * c.g[n][1];
            gT[2] += T * c.g[n][2];
        }
        return mt.k * (gT[0] * gT[0] + gT[1] * gT[1] + gT[2] * gT[2]);
    }
}

template <int DPN>
__global__ __launch_bounds__(kBlock) void k_energy_cells(int C, const double4* __restrict__ x4,
                                                         const int4* __restrict__ conn, Material mt,
                                                         const double* __restrict__ u, const double* __restrict__ z,
                                                         double* __restrict__ partials) {
    __shared__ double red[kBlock / 64];
    double acc = 0.0;
    for (long cell = (long)blockIdx.x * kBlock + threadIdx.x; cell < C; cell += (long)gridDim.x * kBlock) {
        CellEval c;
        const int4 nd = conn[cell];
        c.nodes[0] = nd.x;
        c.nodes[1] = nd.y;
        c.nodes[2] = nd.z;
        c.nodes[3] = nd.w;
        double4 x[4];
#pragma unroll
        for (int n = 0; n < 4; ++n) x[n] = x4[c.nodes[n]];
        c.absdet = tet_gradients(x, c.g);
        c.rho = (((z[c.nodes[0]] + z[c.nodes[1]]) + z[c.nodes[2]]) + z[c.nodes[3]]) / 4.0;
        const double s = penal(c.rho, mt.p, mt.emin);
        acc += s * (c.absdet / 6.0) * cell_energy_density<DPN>(c, mt, u);
    }
    acc = group_sum<64>(acc);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kBlock / 64; ++i) t += red[i];
        partials[blockIdx.x] = t;
    }
}

__global__ void k_sum_partials(int n, const double* __restrict__ partials, double* __restrict__ out) {
    // single wave, fixed order
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 64) acc += partials[i];
    acc = group_sum<64>(acc);
    if (threadIdx.x == 0) *out = acc;
}

// Criterion gradient_x rows: for node a, sum over incident cells of
//   elastic: s V [ 1/2 g_{a,i} (sigma0 : eps) - sum_d H_{d,i} (sigma0 g_a)_d ],
//   thermal: s k V [ g_{a,i} |gT|^2 - 2 gT_i (gT . g_a) ],
// the derivatives of 1/2 s sigma0:eps V and s k |grad T|^2 V (see k_gradx_rows).
template <int DPN, int G>
__global__ __launch_bounds__(kBlock) void k_energy_gradx_rows(RowMesh m, Material mt, const double* __restrict__ u,
                                                              const double* __restrict__ z, double scale,
                                                              double* __restrict__ gx) {
    constexpr int ROWS = kBlock / G;
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    const int row = xcd_block() * ROWS + grp;
    if (row >= m.N) return;
    const int p0 = m.n2c_off[row], p1 = m.n2c_off[row + 1];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int p = p0 + lane; p < p1; p += G) {
        CellEval c;
        load_cell(m, z, p, c);
        const double w = penal(c.rho, mt.p, mt.emin) * (c.absdet / 6.0);
        double ga[3];
        pick_row(c.g, c.a, ga);
        if constexpr (DPN == 3) {
            double H[3][3], eps[6], sig[6], sa[3];
            grad_u_cell(c, u, H);
            voigt_strain(H, eps);
            iso_stress(mt, eps, sig);
            btsig(sig, ga, sa);
            double se = 0.0;
#pragma unroll
            for (int v = 0; v < 6; ++v) se += sig[v] * eps[v];
#pragma unroll
            for (int i = 0; i < 3; ++i)
                acc[i] += w * (0.5 * ga[i] * se - ((H[0][i] * sa[0] + H[1][i] * sa[1]) + H[2][i] * sa[2]));
        } else {
            double gT[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                const double T = u[c.nodes[n]];
                gT[0] += T * c.g[n][0];
                gT[1] += T * c.g[n][1];
                gT[2] += T * c.g[n][2];
            }
            const double tt = gT[0] * gT[0] + gT[1] * gT[1] + gT[2] * gT[2];
            const double ta = gT[0] * ga[0] + gT[1] * ga[1] + gT[2] * ga[2];
#pragma unroll
            for (int i = 0; i < 3; ++i) acc[i] += w * mt.k * (ga[i] * tt - 2.0 * gT[i] * ta);
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) acc[i] = group_sum<G>(acc[i]);
    if (lane == 0) {
        gx[3L * row + 0] = scale * acc[0];
        gx[3L * row + 1] = scale * acc[1];
        gx[3L * row + 2] = scale * acc[2];
    }
}

template <int DPN, int G>
__global__ __launch_bounds__(kBlock) void k_energy_gradz_rows(RowMesh m, Material mt, const double* __restrict__ u,
                                                              const double* __restrict__ z, double* __restrict__ gz) {
    constexpr int ROWS = kBlock / G;
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    const int row = xcd_block() * ROWS + grp;
    if (row >= m.N) return;
    const int p0 = m.n2c_off[row], p1 = m.n2c_off[row + 1];
    double acc = 0.0;
    for (int p = p0 + lane; p < p1; p += G) {
        CellEval c;
        load_cell(m, z, p, c);
        const double ds = dpenal(c.rho, mt.p, mt.emin) * 0.25;
        acc += ds * (c.absdet / 6.0) * cell_energy_density<DPN>(c, mt, u);
    }
    acc = group_sum<G>(acc);
    if (lane == 0) gz[row] = acc;
}

// ---------------------------------------------------------------------------
// Criterion gradient with an adjoint, fused: for row (control node) a,
//   dF/dz_a = sum over incident cells of
//     (1/4) s_F'(rho) V * f_density(u)            (criterion, its own SIMP)
//   + (1/4) s_R'(rho) V * <lambda, B^T C eps(u)>   (PDE, lambda = lam_scale u)
//   + thermal body load: -(1/6)|det J| f (1/4)(1/4) sum_b lambda_b.
// The reference forms the same sum as MatrixTimesVectorPlusVector(gradient_z,
// lambda, dfdz) (src/elliptic/Problem.hpp:421-448).
// ---------------------------------------------------------------------------
template <int DPN, int G>
__global__ __launch_bounds__(kBlock) void k_adjoint_gradz_rows(RowMesh m, Material mc, Material mp,
                                                               const double* __restrict__ u, double lam_scale,
                                                               const double* __restrict__ z, double body,
                                                               double* __restrict__ out) {
    constexpr int ROWS = kBlock / G;
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    const int row = xcd_block() * ROWS + grp;
    if (row >= m.N) return;
    const int p0 = m.n2c_off[row], p1 = m.n2c_off[row + 1];
    double acc = 0.0;
    for (int p = p0 + lane; p < p1; p += G) {
        CellEval c;
        load_cell(m, z, p, c);
        const double V = c.absdet / 6.0;
        const double dsF = dpenal(c.rho, mc.p, mc.emin) * 0.25;
        const double dsR = dpenal(c.rho, mp.p, mp.emin) * 0.25;
        if constexpr (DPN == 3) {
            double eps[6], sigF[6], sigR[6];
            cell_strain(c, u, eps);
            iso_stress(mc, eps, sigF);
            iso_stress(mp, eps, sigR);
            double fF = 0.0, fR = 0.0;
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                fF += sigF[i] * eps[i];
                fR += sigR[i] * eps[i];
            }
            // eps(lambda) = lam_scale * eps(u)
            acc += dsF * V * (0.5 * fF) + dsR * V * (lam_scale * fR);
        } else {
            double gT[3] = {0.0, 0.0, 0.0};
            double us = 0.0;
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                const double T = u[c.nodes[n]];
                us += T;
                gT[0] += T * c.g[n][0];
                gT[1] += T * c.g[n][1];
                gT[2] += T * c.g[n][2];
            }
            const double gg = gT[0] * gT[0] + gT[1] * gT[1] + gT[2] * gT[2];
            acc += dsF * V * (mc.k * gg) + dsR * V * (mp.k * lam_scale * gg);
            acc += -1.0 * (c.absdet / 6.0) * body * 0.25 * 0.25 * (lam_scale * us);
        }
    }
    acc = group_sum<G>(acc);
    if (lane == 0) out[row] = acc;
}

// ---------------------------------------------------------------------------
// Uniform natural BC on a side set, in boundary-node order:
// R[node] += sum over faces (basis 1/3) * value * scale * (1/2) * |cross|.
// ---------------------------------------------------------------------------
template <int DPN>
__global__ void k_natural_bc(int nb, const int* __restrict__ bnodes, const int* __restrict__ boff,
                             const int* __restrict__ bface, const int* __restrict__ fcells,
                             const int* __restrict__ flnodes, const int4* __restrict__ conn,
                             const double4* __restrict__ x4, double v0, double v1, double v2, double scale,
                             double* __restrict__ R) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const int node = bnodes[i];
    double sum = 0.0;
    for (int k = boff[i]; k < boff[i + 1]; ++k) {
        const int f = bface[k] >> 2;
        const int cell = fcells[f];
        const int4 nd = conn[cell];
        const int cn[4] = {nd.x, nd.y, nd.z, nd.w};
        const double4 a = x4[cn[flnodes[3 * f + 0]]];
        const double4 b = x4[cn[flnodes[3 * f + 1]]];
        const double4 c = x4[cn[flnodes[3 * f + 2]]];
        const double J00 = a.x - c.x, J01 = a.y - c.y, J02 = a.z - c.z;
        const double J10 = b.x - c.x, J11 = b.y - c.y, J12 = b.z - c.z;
        const double j23 = J01 * J12 - J02 * J11;
        const double j31 = J02 * J10 - J00 * J12;
        const double j12 = J00 * J11 - J01 * J10;
        sum += (scale * 0.5) * sqrt(j23 * j23 + j31 * j31 + j12 * j12);
    }
    const double basis = 1.0 / 3.0;
    R[(long)node * DPN + 0] += basis * v0 * sum;
    if constexpr (DPN == 3) {
        R[(long)node * DPN + 1] += basis * v1 * sum;
        R[(long)node * DPN + 2] += basis * v2 * sum;
    }
}

// ---------------------------------------------------------------------------
// applyBlockConstraints (src/ApplyConstraints.hpp:40-112).  The reference
// subtracts A[r][c] * g_r from rhs[c] for every constrained dof r and every
// other dof c of its row with one atomic add each, so the order of the adds
// into a dof next to several constrained dofs changes from run to run.  Here
// the same sums run in a fixed order: the constrained values are scattered
// into a dense vector, each node next to a constrained node gathers its
// terms from the constrained rows (the graph is symmetric: row n holds every
// node whose row holds n), and only then are the rows and columns cleared.
// ---------------------------------------------------------------------------
// g[dof] = scale * value of each constrained dof; touched[n] = 1 for the
// nodes of the constrained rows.  g and touched start zeroed.
__global__ void k_constraints_mark(int dpn, const int* __restrict__ rowmap, const int* __restrict__ colidx, int nbc,
                                   const int* __restrict__ dofs, const double* __restrict__ bcv, double scale,
                                   double* __restrict__ g, unsigned char* __restrict__ touched) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbc) return;
    const int rdof = dofs[b];
    g[rdof] = scale * bcv[b];
    const int rnode = rdof / dpn;
    for (int off = rowmap[rnode]; off < rowmap[rnode + 1]; ++off) touched[colidx[off]] = 1;
}

// rhs[c] -= A[r][c] * g[r] over the constrained dofs r != c whose rows hold c,
// in the order of node c's row and then of r's local dof.  One thread per node.
template <class VT>
__global__ void k_constraints_rhs(int N, int dpn, const int* __restrict__ rowmap, const int* __restrict__ colidx,
                                  const VT* __restrict__ vals, const double* __restrict__ g,
                                  const unsigned char* __restrict__ touched, double* __restrict__ rhs) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N || !touched[n]) return;
    double acc[4];
    for (int lc = 0; lc < dpn; ++lc) acc[lc] = rhs[(long)n * dpn + lc];
    for (int e = rowmap[n]; e < rowmap[n + 1]; ++e) {
        const int r = colidx[e];
        bool any = false;
        for (int lr = 0; lr < dpn; ++lr) any |= g[(long)r * dpn + lr] != 0.0;
        if (!any) continue;
        int o = rowmap[r];
        const int o1 = rowmap[r + 1];
        while (o < o1 && colidx[o] != n) ++o;
        if (o == o1) continue;
        for (int lr = 0; lr < dpn; ++lr) {
            const double v = g[(long)r * dpn + lr];
            if (v == 0.0) continue;
            for (int lc = 0; lc < dpn; ++lc)
                if (r != n || lr != lc) acc[lc] -= (double)vals[(long)dpn * dpn * o + dpn * lr + lc] * v;
        }
    }
    for (int lc = 0; lc < dpn; ++lc) rhs[(long)n * dpn + lc] = acc[lc];
}

// Constrained rows become identity rows and their columns are cleared (the
// reference's first loop without its right-hand-side update); every write is
// a constant, so the threads' order does not matter.
template <class VT>
__global__ void k_constraints_1(int dpn, const int* __restrict__ rowmap, const int* __restrict__ colidx,
                                VT* __restrict__ vals, int nbc, const int* __restrict__ dofs) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbc) return;
    const int rdof = dofs[b];
    const int rnode = rdof / dpn, lr = rdof % dpn;
    for (int off = rowmap[rnode]; off < rowmap[rnode + 1]; ++off) {
        const int cnode = colidx[off];
        for (int lc = 0; lc < dpn; ++lc) {
            const long ent = (long)dpn * dpn * off + dpn * lr + lc;
            const int cdof = dpn * cnode + lc;
            if (cdof == rdof) {
                vals[ent] = (VT)1.0;
            } else {
                vals[ent] = (VT)0.0;
                for (int o2 = rowmap[cnode]; o2 < rowmap[cnode + 1]; ++o2)
                    if (dpn * colidx[o2] + lr == rdof) vals[(long)dpn * dpn * o2 + dpn * lc + lr] = (VT)0.0;
            }
        }
    }
}

__global__ void k_constraints_2(int nbc, const int* __restrict__ dofs, const double* __restrict__ bcv, double scale,
                                double* __restrict__ rhs) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nbc) rhs[dofs[b]] = scale * bcv[b];
}

// y = A x + beta y on the block graph, one group of 4 lanes per block row.
template <int BI, int BJ, class VT>
__global__ __launch_bounds__(kBlock) void k_block_spmv(int N, const int* __restrict__ rowmap,
                                                       const int* __restrict__ colidx, const VT* __restrict__ vals,
                                                       const double* __restrict__ x, double beta,
                                                       double* __restrict__ y) {
    constexpr int G = 4;
    const int grp = (blockIdx.x * kBlock + threadIdx.x) / G, lane = threadIdx.x % G;
    const int row = grp;
    if (row >= N) return;
    double acc[BI];
#pragma unroll
    for (int i = 0; i < BI; ++i) acc[i] = 0.0;
    for (int e = rowmap[row] + lane; e < rowmap[row + 1]; e += G) {
        const int c = colidx[e];
        double xv[BJ];
#pragma unroll
        for (int j = 0; j < BJ; ++j) xv[j] = x[(long)c * BJ + j];
        const VT* v = vals + (long)e * BI * BJ;
#pragma unroll
        for (int i = 0; i < BI; ++i)
#pragma unroll
            for (int j = 0; j < BJ; ++j) acc[i] += (double)v[i * BJ + j] * xv[j];
    }
#pragma unroll
    for (int i = 0; i < BI; ++i) acc[i] = group_sum<G>(acc[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < BI; ++i) {
            const long o = (long)row * BI + i;
            y[o] = acc[i] + (beta == 0.0 ? 0.0 : beta * y[o]);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_copy16(long n, const int4* __restrict__ src, int4* __restrict__ dst) {
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) dst[i] = src[i];
}

// ------------------------------------------------------------------ helpers
Material material_of(int physics, const double* mat, const double* pen) {
    Material m{};
    if (physics == PA_MECHANICS) {
        const double E = mat[0], nu = mat[1];
        const double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
        m.lambda = c * nu;
        m.mu = 1.0 / 2.0 * c * (1.0 - 2.0 * nu);
    } else {
        m.k = mat[0];
    }
    m.p = pen[0];
    m.emin = pen[1];
    return m;
}

RowMesh row_mesh(const pa_mesh* m) {
    RowMesh r;
    r.N = m->N;
    r.x4 = reinterpret_cast<const double4*>(m->d_coords4.p);
    r.conn = reinterpret_cast<const int4*>(m->d_conn.p);
    r.rowmap = m->d_rowmap.p;
    r.n2c_off = m->d_n2c_off.p;
    r.pairs = m->d_pairs.p;
    r.pcell = m->d_pcell.p;
    r.row_pat = m->d_row_pat.p;
    r.pat_info = m->d_pat_info.p;
    r.pat_a = m->d_pat_a.p;
    r.pat_ent = m->d_pat_ent.p;
    r.pat_item = m->d_pat_item.p;
    r.pat_delta = m->d_pat_delta.p;
    r.pat_item2 = m->d_pat_item2.p;
    return r;
}

TileMesh tile_mesh(const pa_mesh* m) {
    TileMesh t;
    t.tile = m->d_tile.p;
    t.rows = m->d_tp_row.p;
    t.runs = m->d_tp_run.p;
    t.cells = m->d_tp_cell.p;
    t.cslots = m->d_tp_cslot.p;
    t.hslots = m->d_tp_hslot.p;
    t.units = m->d_tp_unit.p;
    t.ents = reinterpret_cast<const uint4*>(m->d_tp_ent.p);
    t.rowmap = m->d_rowmap.p;
    t.d0 = m->grid_dims[0];
    t.d1 = m->grid_dims[1];
    t.d2 = m->grid_dims[2];
    return t;
}

// Dynamic LDS of a tile kernel, or 0 when the device cannot give it (the
// caller then takes the row kernels).  Above 64 KiB the kernel is opted in.
template <class K>
size_t tile_lds(K kernel, size_t bytes) {
    static int limit = -1;
    if (limit < 0) {
        int dev = 0;
        PA_HIP(hipGetDevice(&dev));
        PA_HIP(hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    }
    if (bytes + 64 * sizeof(int) > (size_t)limit) return 0;
    if (bytes > 65536) PA_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return bytes;
}

constexpr int kG = 8;  // lanes per row

int row_blocks(int N, int G) { return (N + kBlock / G - 1) / (kBlock / G); }

void check_launch() { PA_HIP(hipGetLastError()); }

void natural_bcs(const pa_vector_function* vf, double* R, hipStream_t st) {
    const pa_mesh* m = vf->mesh;
    const double scale = vf->physics == PA_MECHANICS ? -1.0 : 1.0;  // evaluate_boundary scales
    for (const auto& bc : vf->natural_bcs) {
        auto it = m->side_sets.find(bc.side_set);
        PA_REQUIRE(it != m->side_sets.end(), PA_ERR_ARG, "natural BC: unknown side set " + bc.side_set);
        const SideSet& s = it->second;
        if (s.num_bnodes == 0) continue;
        const int nb = s.num_bnodes;
        const dim3 grid((nb + 255) / 256), blk(256);
        if (vf->dpn == 3)
            k_natural_bc<3><<<grid, blk, 0, st>>>(nb, s.d_bnodes.p, s.d_bface_off.p, s.d_bface.p, s.d_cells.p,
                                                  s.d_lnodes.p, reinterpret_cast<const int4*>(m->d_conn.p),
                                                  reinterpret_cast<const double4*>(m->d_coords4.p), bc.values[0],
                                                  bc.values[1], bc.values[2], scale, R);
        else
            k_natural_bc<1><<<grid, blk, 0, st>>>(nb, s.d_bnodes.p, s.d_bface_off.p, s.d_bface.p, s.d_cells.p,
                                                  s.d_lnodes.p, reinterpret_cast<const int4*>(m->d_conn.p),
                                                  reinterpret_cast<const double4*>(m->d_coords4.p), bc.values[0],
                                                  0.0, 0.0, scale, R);
        check_launch();
    }
}

template <int DPN, int KMAX, class VT, bool RES, bool TR>
void launch_jac(const pa_vector_function* vf, const double* u, const double* z, VT* vals, double* R, hipStream_t st) {
    const pa_mesh* m = vf->mesh;
    const Material mt = material_of(vf->physics, vf->material, vf->penalty);
    const double body = vf->has_body ? vf->body_value : 0.0;
    // wider groups (fewer rows per block) keep the LDS rows of high-degree meshes small
    constexpr int G = KMAX <= 16 ? kG : 32;
    k_jacobian_rows<DPN, G, KMAX, VT, RES, TR>
        <<<row_blocks(m->N, G), kBlock, 0, st>>>(row_mesh(m), mt, u, z, body, vals, R);
    check_launch();
}

template <int DPN, class VT, bool RES, bool TR>
void dispatch_jac(const pa_vector_function* vf, const double* u, const double* z, VT* vals, double* R,
                  hipStream_t st) {
    const int K = vf->mesh->max_row_len;
    if (K <= 16) launch_jac<DPN, 16, VT, RES, TR>(vf, u, z, vals, R, st);
    else if (K <= 32) launch_jac<DPN, 32, VT, RES, TR>(vf, u, z, vals, R, st);
    else if (K <= 64 || DPN == 1) launch_jac<DPN, 64, VT, RES, TR>(vf, u, z, vals, R, st);
    else throw Error(PA_ERR_LIMIT, "gradient_u: node degree above 63 is not supported");
}


// Residual-type row sums: through row patterns and packed node records when
// the mesh has them; else one incident cell per lane (32 lanes per row) when
// the node degree allows, else 8 lanes per row looping over the cells.
// Residual rows by the tile kernel, plus a nodal load fext when given;
// false when the mesh has no usable tiles.
bool tile_residual(const pa_mesh* m, int dpn, const Material& mt, const double* u, const double* z, double body,
                   double scale, const double* fext, double* R, hipStream_t st) {
    if (!m->has_tiles || m->tile_rows_max * 8 > kResThreads) return false;
    const int hoff = m->tile_rows_max * 32 * dpn;  // doubles; a multiple of 32 keeps the records aligned
    const size_t lds = (size_t)(hoff + (4 + dpn) * kTileHalo) * sizeof(double);
    const size_t ok = dpn == 3 ? tile_lds(k_tile_residual<3>, lds) : tile_lds(k_tile_residual<1>, lds);
    if (!ok) return false;
    const double4* x4 = reinterpret_cast<const double4*>(m->d_coords4.p);
    if (dpn == 3)
        k_tile_residual<3><<<m->num_tiles, kResThreads, lds, st>>>(tile_mesh(m), x4, z, mt, u, body, scale, hoff, fext,
                                                                   R);
    else
        k_tile_residual<1><<<m->num_tiles, kResThreads, lds, st>>>(tile_mesh(m), x4, z, mt, u, body, scale, hoff, fext,
                                                                   R);
    check_launch();
    return true;
}

void residual_rows(const pa_mesh* m, int dpn, const Material& mt, const double* u, const double* z, double body,
                   double scale, double* R, hipStream_t st) {
    if (grid_residual(m, dpn, mt, u, z, body, scale, nullptr, R, st)) return;
    if (tile_residual(m, dpn, mt, u, z, body, scale, nullptr, R, st)) return;
    if (m->has_patterns) {
        if (dpn == 3)
            k_residual_pattern<3><<<row_blocks(m->N, 32), kBlock, 0, st>>>(row_mesh(m), z, mt, u, body, scale, R);
        else
            k_residual_pattern<1><<<row_blocks(m->N, 32), kBlock, 0, st>>>(row_mesh(m), z, mt, u, body, scale, R);
    } else if (m->max_cells_per_node <= 32) {
        if (dpn == 3)
            k_residual_rows<3, 32><<<row_blocks(m->N, 32), kBlock, 0, st>>>(row_mesh(m), mt, u, z, body, scale, R);
        else
            k_residual_rows<1, 32><<<row_blocks(m->N, 32), kBlock, 0, st>>>(row_mesh(m), mt, u, z, body, scale, R);
    } else {
        if (dpn == 3)
            k_residual_rows<3, kG><<<row_blocks(m->N, kG), kBlock, 0, st>>>(row_mesh(m), mt, u, z, body, scale, R);
        else
            k_residual_rows<1, kG><<<row_blocks(m->N, kG), kBlock, 0, st>>>(row_mesh(m), mt, u, z, body, scale, R);
    }
    check_launch();
}

// ---------------------------------------------------------------------------
// Body loads given per cell (BodyLoad::get, src/BodyLoads.hpp:54-122, with
// the function's value f_c at the centroid): a cell adds
// b_c = -(1/6) |det J| f_c (1/4) rho_c to the dof of each of its nodes.  One
// thread per row, over the row's incident cells in order: no atomics.
// ---------------------------------------------------------------------------
// R[row, dof] += sum over the row's cells of b_c
template <int DPN>
__global__ __launch_bounds__(256) void k_body_cells_residual(RowMesh m, const double* __restrict__ z,
                                                             const double* __restrict__ f, int dof,
                                                             double* __restrict__ R) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= m.N) return;
    double r = 0.0;
    for (int p = m.n2c_off[row]; p < m.n2c_off[row + 1]; ++p) {
        CellEval c;
        load_cell(m, z, p, c);
        r += -1.0 * (c.absdet / 6.0) * f[m.pairs[p].x >> 2] * 0.25 * c.rho;
    }
    R[(long)row * DPN + dof] += r;
}

// gradient_z blocks (row = control node b, column = state node a, 1 x DPN):
// d b_c / d z_b = -(1/6) |det J| f_c (1/4) (1/4) for each node a of the cell
template <int DPN>
__global__ __launch_bounds__(256) void k_body_cells_gradz(RowMesh m, const double* __restrict__ z,
                                                          const double* __restrict__ f, int dof,
                                                          double* __restrict__ vals) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= m.N) return;
    const long rs = m.rowmap[row];
    for (int p = m.n2c_off[row]; p < m.n2c_off[row + 1]; ++p) {
        CellEval c;
        load_cell(m, z, p, c);
        const double d = -1.0 * (c.absdet / 6.0) * f[m.pairs[p].x >> 2] * 0.25 * 0.25;
#pragma unroll
        for (int b = 0; b < 4; ++b) vals[(rs + ((c.slots >> (8 * b)) & 0xff)) * DPN + dof] += d;
    }
}

// gradient_x blocks (row = coordinate node a, column = state node, 3 x DPN):
// d b_c / d x_{a,i} = b_c (grad N_a)_i, f_c held constant per cell as in the
// reference (its cubature point is mapped in plain doubles)
template <int DPN>
__global__ __launch_bounds__(256) void k_body_cells_gradx(RowMesh m, const double* __restrict__ z,
                                                          const double* __restrict__ f, int dof,
                                                          double* __restrict__ vals) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= m.N) return;
    const long rs = m.rowmap[row];
    for (int p = m.n2c_off[row]; p < m.n2c_off[row + 1]; ++p) {
        CellEval c;
        load_cell(m, z, p, c);
        const double bl = -1.0 * (c.absdet / 6.0) * f[m.pairs[p].x >> 2] * 0.25 * c.rho;
        double ga[3];
        pick_row(c.g, c.a, ga);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const long e = (rs + ((c.slots >> (8 * b)) & 0xff)) * 3 * DPN;
#pragma unroll
            for (int i = 0; i < 3; ++i) vals[e + i * DPN + dof] += bl * ga[i];
        }
    }
}

// out[b] += sum over the row's cells and their nodes a of d b_c / d z_b * lam_a,
// lam = lscale * u (the fused self-adjoint criterion gradient's (dR/dz)^T lam)
template <int DPN>
__global__ __launch_bounds__(256) void k_body_cells_adjoint(RowMesh m, const double* __restrict__ z,
                                                            const double* __restrict__ f, int dof,
                                                            const double* __restrict__ u, double lscale,
                                                            double* __restrict__ out) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= m.N) return;
    double r = 0.0;
    for (int p = m.n2c_off[row]; p < m.n2c_off[row + 1]; ++p) {
        CellEval c;
        load_cell(m, z, p, c);
        const double d = -1.0 * (c.absdet / 6.0) * f[m.pairs[p].x >> 2] * 0.25 * 0.25;
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b) s += lscale * u[(long)c.nodes[b] * DPN + dof];
        r += d * s;
    }
    out[row] += r;
}

enum CellLoadTerm { kLoadResidual, kLoadGradZ, kLoadGradX, kLoadAdjoint };

// Adds the per-cell body loads of vf to the residual, gradient_z, gradient_x
// or the self-adjoint criterion gradient.
void cell_loads(const pa_vector_function* vf, CellLoadTerm term, const double* u, const double* z, double* out,
                hipStream_t st) {
    const pa_mesh* m = vf->mesh;
    const int nb = (m->N + 255) / 256;
    for (const auto& L : vf->cell_loads) {
        const double* f = L->f.p;
        auto go = [&](auto dpn) {
            constexpr int D = decltype(dpn)::value;
            if (term == kLoadResidual) k_body_cells_residual<D><<<nb, 256, 0, st>>>(row_mesh(m), z, f, L->dof, out);
            else if (term == kLoadGradZ) k_body_cells_gradz<D><<<nb, 256, 0, st>>>(row_mesh(m), z, f, L->dof, out);
            else if (term == kLoadGradX) k_body_cells_gradx<D><<<nb, 256, 0, st>>>(row_mesh(m), z, f, L->dof, out);
            else k_body_cells_adjoint<D><<<nb, 256, 0, st>>>(row_mesh(m), z, f, L->dof, u, -1.0, out);
        };
        if (vf->dpn == 3) go(std::integral_constant<int, 3>{});
        else go(std::integral_constant<int, 1>{});
        check_launch();
    }
}

// Tile-kernel Jacobian; false when the mesh has no tiles or the device
// cannot give the kernel its LDS.  Self-adjoint operators: the transposed
// assembly is the direct one.
template <class VT>
bool tile_jacobian(const pa_vector_function* vf, const double* z, VT* vals, hipStream_t st) {
    const pa_mesh* m = vf->mesh;
    if (!m->has_tiles) return false;
    const Material mt = material_of(vf->physics, vf->material, vf->penalty);
    const int cmax = m->tile_cells_max + 1;  // + the zero cell
    // 3 x 3 blocks: cell data; 1 x 1: cell products, then the node halo
    const size_t cells = vf->dpn == 3 ? (size_t)cmax * kTileStride : ((size_t)cmax * kThermalStride + 1) / 2 * 2 + 4 * kTileHalo;
    const size_t lds = std::max(cells * sizeof(double), (size_t)m->tile_blocks_max * vf->dpn * vf->dpn * sizeof(VT));
    auto run = [&](auto kernel) {
        if (!tile_lds(kernel, lds)) return false;
        const double4* x4 = reinterpret_cast<const double4*>(m->d_coords4.p);
        kernel<<<m->num_tiles, kTileThreads, lds, st>>>(tile_mesh(m), x4, z, mt, cmax, vals);
        check_launch();
        return true;
    };
    const bool six = m->tile_unit_entries <= 6;
    if (vf->dpn == 3) return six ? run(k_tile_jacobian<3, VT, 6>) : run(k_tile_jacobian<3, VT, 8>);
    return six ? run(k_tile_jacobian<1, VT, 6>) : run(k_tile_jacobian<1, VT, 8>);
}

template <class VT, bool RES, bool TR>
void jacobian(const pa_vector_function* vf, const double* u, const double* z, VT* vals, double* R, hipStream_t st) {
    const pa_mesh* m = vf->mesh;
    if constexpr (!RES) {
        // Kuhn grids: the grid kernels (symmetric operators: gradient_u_T is
        // the same matrix, bit for bit); else the tiles
        if (vf->dpn == 1 && m->has_grid &&
            grid_jacobian1(m, material_of(vf->physics, vf->material, vf->penalty), z, vals, st))
            return;
        if (vf->dpn == 3 && m->has_grid &&
            grid_jacobian3(m, material_of(vf->physics, vf->material, vf->penalty), z, vals, st))
            return;
        if (tile_jacobian(vf, z, vals, st)) return;
    }
    if (m->has_patterns) {
        // The operators are self-adjoint: the transposed assembly equals the
        // direct one (block (i,j) = K_ji^T = K_ij) up to rounding.
        const Material mt = material_of(vf->physics, vf->material, vf->penalty);
        const double body = vf->has_body ? vf->body_value : 0.0;
        const int nb = row_blocks(m->N, 32);
        if (m->max_cells_per_node <= 24) {
            if (vf->dpn == 3)
                k_jacobian_pattern<3, VT, RES, 24><<<nb, kBlock, 0, st>>>(row_mesh(m), z, mt, u, 0.0, vals, R);
            else
                k_jacobian_pattern<1, VT, RES, 24><<<nb, kBlock, 0, st>>>(row_mesh(m), z, mt, u, body, vals, R);
        } else {
            if (vf->dpn == 3)
                k_jacobian_pattern<3, VT, RES, 32><<<nb, kBlock, 0, st>>>(row_mesh(m), z, mt, u, 0.0, vals, R);
            else
                k_jacobian_pattern<1, VT, RES, 32><<<nb, kBlock, 0, st>>>(row_mesh(m), z, mt, u, body, vals, R);
        }
        check_launch();
    } else {
        PA_REQUIRE(m->max_row_len <= 64, PA_ERR_LIMIT, "gradient_u: node degree above 63 is not supported");
        if (vf->dpn == 3) dispatch_jac<3, VT, RES, TR>(vf, u, z, vals, R, st);
        else dispatch_jac<1, VT, RES, TR>(vf, u, z, vals, R, st);
    }
    if (RES) {
        natural_bcs(vf, R, st);
        cell_loads(vf, kLoadResidual, u, z, R, st);
    }
}

__global__ void k_zero(long n, double* __restrict__ v) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) v[i] = 0.0;
}

// The natural BCs' nodal load of vf (null without natural BCs): assembled by
// the node-ordered load kernel into a zeroed vector when a load is added, and
// again on the first call after the mesh's side sets change, each time
// synchronously, so every later call on any stream reads a finished vector.
void build_natural_load(const pa_vector_function* vf) {
    const pa_mesh* m = vf->mesh;
    const long n = (long)m->N * vf->dpn;
    if (vf->fext.n != (size_t)n) vf->fext.alloc(n);
    k_zero<<<(int)std::min<long>(4096, (n + kBlock - 1) / kBlock), kBlock, 0, nullptr>>>(n, vf->fext.p);
    check_launch();
    natural_bcs(vf, vf->fext.p, nullptr);
    PA_HIP(hipStreamSynchronize(nullptr));
    vf->fext_nbc = vf->natural_bcs.size();
    vf->fext_version = m->side_set_version;
}

const double* natural_load(const pa_vector_function* vf) {
    if (vf->natural_bcs.empty()) return nullptr;
    if (vf->fext_nbc != vf->natural_bcs.size() || vf->fext_version != vf->mesh->side_set_version)
        build_natural_load(vf);
    return vf->fext.p;
}

void residual(const pa_vector_function* vf, const double* u, const double* z, double* R, hipStream_t st) {
    const pa_mesh* m = vf->mesh;
    const Material mt = material_of(vf->physics, vf->material, vf->penalty);
    const double body = vf->dpn == 3 ? 0.0 : (vf->has_body ? vf->body_value : 0.0);
    // on grid and tile meshes the natural BCs' load is added as the rows are written
    if (m->has_grid && grid_residual(m, vf->dpn, mt, u, z, body, 1.0, natural_load(vf), R, st)) {
    } else if (m->has_tiles && tile_residual(m, vf->dpn, mt, u, z, body, 1.0, natural_load(vf), R, st)) {
    } else {
        residual_rows(m, vf->dpn, mt, u, z, body, 1.0, R, st);
        natural_bcs(vf, R, st);
    }
    cell_loads(vf, kLoadResidual, u, z, R, st);
}

void check_vf(const pa_vector_function* vf, const void* a, const void* b, const void* c) {
    PA_REQUIRE(vf && vf->mesh, PA_ERR_ARG, "vector function: null handle");
    PA_REQUIRE(a && b && c, PA_ERR_ARG, "vector function: null device pointer");
}

}  // namespace

void criterion_gradient_x(const pa_scalar_function* sf, const double* u, const double* z, double scale, double* g,
                          hipStream_t st) {
    const pa_mesh* m = sf->mesh;
    const int phys = sf->dpn == 3 ? PA_MECHANICS : PA_THERMAL;
    const Material mt = material_of(phys, sf->material, sf->penalty);
    const int nb = row_blocks(m->N, kG);
    if (sf->dpn == 3) k_energy_gradx_rows<3, kG><<<nb, kBlock, 0, st>>>(row_mesh(m), mt, u, z, scale, g);
    else k_energy_gradx_rows<1, kG><<<nb, kBlock, 0, st>>>(row_mesh(m), mt, u, z, scale, g);
    check_launch();
}

void self_adjoint_gradient(const pa_vector_function* vf, const pa_scalar_function* sf, const double* u,
                           const double* z, double* out, hipStream_t st) {
    const pa_mesh* m = vf->mesh;
    const Material mp = material_of(vf->physics, vf->material, vf->penalty);
    const Material mc = material_of(vf->physics, sf->material, sf->penalty);
    const double body = vf->has_body ? vf->body_value : 0.0;
    if (m->max_cells_per_node <= 32) {
        if (vf->dpn == 3)
            k_adjoint_gradz_rows<3, 32><<<row_blocks(m->N, 32), kBlock, 0, st>>>(row_mesh(m), mc, mp, u, -1.0, z,
                                                                                  0.0, out);
        else
            k_adjoint_gradz_rows<1, 32><<<row_blocks(m->N, 32), kBlock, 0, st>>>(row_mesh(m), mc, mp, u, -1.0, z,
                                                                                  body, out);
    } else {
        if (vf->dpn == 3)
            k_adjoint_gradz_rows<3, kG><<<row_blocks(m->N, kG), kBlock, 0, st>>>(row_mesh(m), mc, mp, u, -1.0, z,
                                                                                  0.0, out);
        else
            k_adjoint_gradz_rows<1, kG><<<row_blocks(m->N, kG), kBlock, 0, st>>>(row_mesh(m), mc, mp, u, -1.0, z,
                                                                                  body, out);
    }
    check_launch();
    cell_loads(vf, kLoadAdjoint, u, z, out, st);
}

void apply_block_constraints(const pa_mesh* m, int dpn, void* vals, int precision, double* rhs, int nbc,
                             const int* dofs, const double* bcv, double scale, double* dense, unsigned char* touched,
                             hipStream_t st) {
    if (nbc == 0) return;
    PA_REQUIRE(precision == PA_FP64 || precision == PA_FP32, PA_ERR_ARG, "apply_block_constraints: unknown precision");
    const long nd = (long)m->N * dpn;
    PA_HIP(hipMemsetAsync(dense, 0, nd * sizeof(double), st));
    PA_HIP(hipMemsetAsync(touched, 0, (size_t)m->N, st));
    const dim3 grid((nbc + 255) / 256), blk(256), ngrid((m->N + 255) / 256);
    k_constraints_mark<<<grid, blk, 0, st>>>(dpn, m->d_rowmap.p, m->d_colidx.p, nbc, dofs, bcv, scale, dense, touched);
    check_launch();
    if (precision == PA_FP64) {
        k_constraints_rhs<double><<<ngrid, blk, 0, st>>>(m->N, dpn, m->d_rowmap.p, m->d_colidx.p,
                                                         (const double*)vals, dense, touched, rhs);
        check_launch();
        k_constraints_1<double><<<grid, blk, 0, st>>>(dpn, m->d_rowmap.p, m->d_colidx.p, (double*)vals, nbc, dofs);
    } else {
        k_constraints_rhs<float><<<ngrid, blk, 0, st>>>(m->N, dpn, m->d_rowmap.p, m->d_colidx.p, (const float*)vals,
                                                        dense, touched, rhs);
        check_launch();
        k_constraints_1<float><<<grid, blk, 0, st>>>(dpn, m->d_rowmap.p, m->d_colidx.p, (float*)vals, nbc, dofs);
    }
    check_launch();
    k_constraints_2<<<grid, blk, 0, st>>>(nbc, dofs, bcv, scale, rhs);
    check_launch();
}

}  // namespace pa

using pa::guarded;

extern "C" {

int pa_copy_kernel(void* dst, const void* src, size_t bytes, void* stream) {
    return guarded([&] {
        PA_REQUIRE(dst && src && bytes % 16 == 0, PA_ERR_ARG, "pa_copy_kernel: bytes must be a multiple of 16");
        const long n = (long)(bytes / 16);
        if (n == 0) return;
        const int nb = (int)std::min<long>((n + pa::kBlock - 1) / pa::kBlock, 8192);
        pa::k_copy16<<<nb, pa::kBlock, 0, (hipStream_t)stream>>>(n, (const int4*)src, (int4*)dst);
        pa::check_launch();
    });
}

int pa_vector_function_create(pa_mesh* mesh, int physics, const double* material, const double* penalty,
                              pa_vector_function** out) {
    return guarded([&] {
        PA_REQUIRE(mesh && material && penalty && out, PA_ERR_ARG, "pa_vector_function_create: null argument");
        PA_REQUIRE(physics == PA_MECHANICS || physics == PA_THERMAL, PA_ERR_ARG,
                   "pa_vector_function_create: unknown physics");
        if (physics == PA_MECHANICS)
            PA_REQUIRE(material[1] > -1.0 && material[1] < 0.5, PA_ERR_ARG, "Poissons Ratio out of (-1, 0.5)");
        auto* vf = new pa_vector_function();
        vf->mesh = mesh;
        vf->physics = physics;
        vf->dpn = physics == PA_MECHANICS ? 3 : 1;
        vf->material[0] = material[0];
        vf->material[1] = physics == PA_MECHANICS ? material[1] : 0.0;
        vf->penalty[0] = penalty[0];
        vf->penalty[1] = penalty[1];
        *out = vf;
    });
}

int pa_vector_function_destroy(pa_vector_function* vf) {
    return guarded([&] { delete vf; });
}

int pa_vector_function_dofs_per_node(const pa_vector_function* vf, int* dpn) {
    return guarded([&] {
        PA_REQUIRE(vf && dpn, PA_ERR_ARG, "null argument");
        *dpn = vf->dpn;
    });
}

int pa_vector_function_add_uniform_natural_bc(pa_vector_function* vf, const char* side_set, const double* values) {
    return guarded([&] {
        PA_REQUIRE(vf && side_set && values, PA_ERR_ARG, "pa_vector_function_add_uniform_natural_bc: null argument");
        PA_REQUIRE(vf->mesh->side_sets.count(side_set), PA_ERR_ARG,
                   std::string("pa_vector_function_add_uniform_natural_bc: unknown side set ") + side_set);
        NaturalBC bc;
        bc.side_set = side_set;
        for (int i = 0; i < 3; ++i) bc.values[i] = i < vf->dpn ? values[i] : 0.0;
        vf->natural_bcs.push_back(bc);
        pa::build_natural_load(vf);
    });
}

int pa_vector_function_add_body_load(pa_vector_function* vf, int dof, double value) {
    return guarded([&] {
        PA_REQUIRE(vf, PA_ERR_ARG, "pa_vector_function_add_body_load: null handle");
        PA_REQUIRE(vf->physics == PA_THERMAL && dof == 0, PA_ERR_ARG,
                   "pa_vector_function_add_body_load: only the thermal source (dof 0) is supported");
        vf->has_body = true;
        vf->body_dof = dof;
        vf->body_value += value;
    });
}

int pa_vector_function_add_body_load_function(pa_vector_function* vf, int dof, const char* function) {
    return guarded([&] {
        PA_REQUIRE(vf && function, PA_ERR_ARG, "pa_vector_function_add_body_load_function: null argument");
        PA_REQUIRE(dof >= 0 && dof < vf->dpn, PA_ERR_ARG, "pa_vector_function_add_body_load_function: dof out of range");
        const pa::Expr f(function);
        const pa_mesh* m = vf->mesh;
        // the function at each cell's centroid, the one-point cubature of Tet4
        std::vector<double> xyz((size_t)m->N * 3), fc((size_t)m->C);
        std::vector<int> conn((size_t)m->C * 4);
        PA_HIP(hipDeviceSynchronize());
        PA_HIP(hipMemcpy(xyz.data(), m->d_coords.p, xyz.size() * sizeof(double), hipMemcpyDeviceToHost));
        PA_HIP(hipMemcpy(conn.data(), m->d_conn.p, conn.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (int c = 0; c < m->C; ++c) {
            double x[3] = {0.0, 0.0, 0.0};
            for (int d = 0; d < 3; ++d) {
                for (int n = 0; n < 4; ++n) x[d] += xyz[(size_t)conn[(size_t)c * 4 + n] * 3 + d];
                x[d] *= 0.25;
            }
            fc[c] = f.eval(x[0], x[1], x[2], 0.0);
        }
        auto L = std::make_unique<pa_vector_function::CellLoad>();
        L->dof = dof;
        L->function = function;
        L->f.upload(fc.data(), fc.size());
        vf->cell_loads.push_back(std::move(L));
    });
}

int pa_vector_function_value(const pa_vector_function* vf, const double* u, const double* z, double* R,
                             void* stream) {
    return guarded([&] {
        pa::check_vf(vf, u, z, R);
        pa::residual(vf, u, z, R, (hipStream_t)stream);
    });
}

int pa_vector_function_gradient_u(const pa_vector_function* vf, const double* u, const double* z, void* vals,
                                  int precision, void* stream) {
    return guarded([&] {
        pa::check_vf(vf, u, z, vals);
        if (precision == PA_FP64)
            pa::jacobian<double, false, false>(vf, u, z, (double*)vals, nullptr, (hipStream_t)stream);
        else if (precision == PA_FP32)
            pa::jacobian<float, false, false>(vf, u, z, (float*)vals, nullptr, (hipStream_t)stream);
        else
            throw pa::Error(PA_ERR_ARG, "gradient_u: unknown precision");
    });
}

int pa_vector_function_gradient_u_T(const pa_vector_function* vf, const double* u, const double* z, void* vals,
                                    int precision, void* stream) {
    return guarded([&] {
        pa::check_vf(vf, u, z, vals);
        if (precision == PA_FP64)
            pa::jacobian<double, false, true>(vf, u, z, (double*)vals, nullptr, (hipStream_t)stream);
        else if (precision == PA_FP32)
            pa::jacobian<float, false, true>(vf, u, z, (float*)vals, nullptr, (hipStream_t)stream);
        else
            throw pa::Error(PA_ERR_ARG, "gradient_u_T: unknown precision");
    });
}

int pa_vector_function_value_gradient_u(const pa_vector_function* vf, const double* u, const double* z, double* R,
                                        void* vals, int precision, void* stream) {
    return guarded([&] {
        pa::check_vf(vf, u, z, vals);
        PA_REQUIRE(R, PA_ERR_ARG, "value_gradient_u: null residual");
        PA_REQUIRE(precision == PA_FP64 || precision == PA_FP32, PA_ERR_ARG, "value_gradient_u: unknown precision");
        const pa_mesh* m = vf->mesh;
        if (vf->dpn == 3 && m->has_grid) {
            // elastic Kuhn grids: the residual from the blocks in the Jacobian
            // kernel (R = K u + f), then the body-load functions
            const pa::Material mt = pa::material_of(vf->physics, vf->material, vf->penalty);
            const double* f = pa::natural_load(vf);
            const bool done = precision == PA_FP64
                                  ? pa::grid_value_jacobian3(m, mt, u, z, f, R, (double*)vals, (hipStream_t)stream)
                                  : pa::grid_value_jacobian3(m, mt, u, z, f, R, (float*)vals, (hipStream_t)stream);
            if (done) {
                pa::cell_loads(vf, pa::kLoadResidual, u, z, R, (hipStream_t)stream);
                return;
            }
        }
        if (m->has_tiles || m->has_grid) {
            // on grid and tile meshes the residual kernel and the Jacobian
            // kernel one after the other beat one fused pass (whose cell data
            // halves the workgroups per CU)
            pa::residual(vf, u, z, R, (hipStream_t)stream);
            if (precision == PA_FP64)
                pa::jacobian<double, false, false>(vf, u, z, (double*)vals, nullptr, (hipStream_t)stream);
            else
                pa::jacobian<float, false, false>(vf, u, z, (float*)vals, nullptr, (hipStream_t)stream);
            return;
        }
        if (precision == PA_FP64)
            pa::jacobian<double, true, false>(vf, u, z, (double*)vals, R, (hipStream_t)stream);
        else if (precision == PA_FP32)
            pa::jacobian<float, true, false>(vf, u, z, (float*)vals, R, (hipStream_t)stream);
        else
            throw pa::Error(PA_ERR_ARG, "value_gradient_u: unknown precision");
    });
}

int pa_vector_function_gradient_z(const pa_vector_function* vf, const double* u, const double* z, double* vals,
                                  void* stream) {
    return guarded([&] {
        pa::check_vf(vf, u, z, vals);
        const pa_mesh* m = vf->mesh;
        PA_REQUIRE(m->max_row_len <= 64, PA_ERR_LIMIT, "gradient_z: node degree above 63 is not supported");
        const pa::Material mt = pa::material_of(vf->physics, vf->material, vf->penalty);
        const double body = vf->has_body ? vf->body_value : 0.0;
        const hipStream_t st = (hipStream_t)stream;
        const int nb8 = pa::row_blocks(m->N, pa::kG), nb32 = pa::row_blocks(m->N, 32);
        const double bd = vf->dpn == 3 ? 0.0 : body;
        if (vf->dpn == 3) {
            if (m->max_row_len <= 16)
                pa::k_gradz_rows<3, pa::kG, 16><<<nb8, pa::kBlock, 0, st>>>(pa::row_mesh(m), mt, u, z, bd, vals);
            else
                pa::k_gradz_rows<3, 32, 64><<<nb32, pa::kBlock, 0, st>>>(pa::row_mesh(m), mt, u, z, bd, vals);
        } else {
            if (m->max_row_len <= 16)
                pa::k_gradz_rows<1, pa::kG, 16><<<nb8, pa::kBlock, 0, st>>>(pa::row_mesh(m), mt, u, z, bd, vals);
            else
                pa::k_gradz_rows<1, 32, 64><<<nb32, pa::kBlock, 0, st>>>(pa::row_mesh(m), mt, u, z, bd, vals);
        }
        pa::check_launch();
        pa::cell_loads(vf, pa::kLoadGradZ, u, z, vals, st);
    });
}

int pa_vector_function_gradient_x(const pa_vector_function* vf, const double* u, const double* z, double* vals,
                                  void* stream) {
    return guarded([&] {
        pa::check_vf(vf, u, z, vals);
        const pa_mesh* m = vf->mesh;
        PA_REQUIRE(m->max_row_len <= 64, PA_ERR_LIMIT, "gradient_x: node degree above 63 is not supported");
        const pa::Material mt = pa::material_of(vf->physics, vf->material, vf->penalty);
        const double body = vf->has_body ? vf->body_value : 0.0;
        const hipStream_t st = (hipStream_t)stream;
        const int nb8 = pa::row_blocks(m->N, pa::kG), nb32 = pa::row_blocks(m->N, 32);
        if (vf->dpn == 3) {
            if (m->max_row_len <= 16)
                pa::k_gradx_rows<3, pa::kG, 16><<<nb8, pa::kBlock, 0, st>>>(pa::row_mesh(m), mt, u, z, 0.0, vals);
            else
                pa::k_gradx_rows<3, 32, 64><<<nb32, pa::kBlock, 0, st>>>(pa::row_mesh(m), mt, u, z, 0.0, vals);
        } else {
            if (m->max_row_len <= 16)
                pa::k_gradx_rows<1, pa::kG, 16><<<nb8, pa::kBlock, 0, st>>>(pa::row_mesh(m), mt, u, z, body, vals);
            else
                pa::k_gradx_rows<1, 32, 64><<<nb32, pa::kBlock, 0, st>>>(pa::row_mesh(m), mt, u, z, body, vals);
        }
        pa::check_launch();
        // evaluate_boundary: the same scale as the residual's natural BCs
        const double scale = vf->physics == PA_MECHANICS ? -1.0 : 1.0;
        for (const auto& bc : vf->natural_bcs) {
            auto it = m->side_sets.find(bc.side_set);
            PA_REQUIRE(it != m->side_sets.end(), PA_ERR_ARG, "natural BC: unknown side set " + bc.side_set);
            const pa::SideSet& ss = it->second;
            if (ss.num_bnodes == 0) continue;
            const int nb = ss.num_bnodes;
            const auto* c4 = reinterpret_cast<const int4*>(m->d_conn.p);
            const auto* x4 = reinterpret_cast<const double4*>(m->d_coords4.p);
            if (vf->dpn == 3)
                pa::k_gradx_natural_bc<3><<<(nb + 255) / 256, 256, 0, st>>>(
                    nb, ss.d_bnodes.p, ss.d_bface_off.p, ss.d_bface.p, ss.d_cells.p, ss.d_lnodes.p, c4, x4,
                    m->d_rowmap.p, m->d_colidx.p, bc.values[0], bc.values[1], bc.values[2], scale, vals);
            else
                pa::k_gradx_natural_bc<1><<<(nb + 255) / 256, 256, 0, st>>>(
                    nb, ss.d_bnodes.p, ss.d_bface_off.p, ss.d_bface.p, ss.d_cells.p, ss.d_lnodes.p, c4, x4,
                    m->d_rowmap.p, m->d_colidx.p, bc.values[0], 0.0, 0.0, scale, vals);
            pa::check_launch();
        }
        pa::cell_loads(vf, pa::kLoadGradX, u, z, vals, st);
    });
}

int pa_scalar_function_create(pa_mesh* mesh, int type, const double* material, const double* penalty,
                              pa_scalar_function** out) {
    return guarded([&] {
        PA_REQUIRE(mesh && material && penalty && out, PA_ERR_ARG, "pa_scalar_function_create: null argument");
        PA_REQUIRE(type == PA_INTERNAL_ELASTIC_ENERGY || type == PA_INTERNAL_THERMAL_ENERGY, PA_ERR_ARG,
                   "pa_scalar_function_create: unknown type");
        auto* sf = new pa_scalar_function();
        sf->mesh = mesh;
        sf->type = type;
        sf->dpn = type == PA_INTERNAL_ELASTIC_ENERGY ? 3 : 1;
        sf->material[0] = material[0];
        sf->material[1] = type == PA_INTERNAL_ELASTIC_ENERGY ? material[1] : 0.0;
        sf->penalty[0] = penalty[0];
        sf->penalty[1] = penalty[1];
        try {
            sf->d_partials.alloc(1024);
        } catch (...) {
            delete sf;
            throw;
        }
        *out = sf;
    });
}

int pa_scalar_function_destroy(pa_scalar_function* sf) {
    return guarded([&] { delete sf; });
}

int pa_scalar_function_value(const pa_scalar_function* sf, const double* u, const double* z, double* value,
                             void* stream) {
    return guarded([&] {
        PA_REQUIRE(sf && u && z && value, PA_ERR_ARG, "pa_scalar_function_value: null argument");
        const pa_mesh* m = sf->mesh;
        const int phys = sf->dpn == 3 ? PA_MECHANICS : PA_THERMAL;
        const pa::Material mt = pa::material_of(phys, sf->material, sf->penalty);
        const hipStream_t st = (hipStream_t)stream;
        const int nb = (int)std::min<long>(1024, (m->C + pa::kBlock - 1) / pa::kBlock);
        const auto* x4 = reinterpret_cast<const double4*>(m->d_coords4.p);
        const auto* c4 = reinterpret_cast<const int4*>(m->d_conn.p);
        if (sf->dpn == 3) pa::k_energy_cells<3><<<nb, pa::kBlock, 0, st>>>(m->C, x4, c4, mt, u, z, sf->d_partials.p);
        else pa::k_energy_cells<1><<<nb, pa::kBlock, 0, st>>>(m->C, x4, c4, mt, u, z, sf->d_partials.p);
        pa::check_launch();
        pa::k_sum_partials<<<1, 64, 0, st>>>(nb, sf->d_partials.p, value);
        pa::check_launch();
    });
}

// Device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array).
static double* scratch_on_device() {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double* ptr[kMaxDevices] = {};
    int dev = 0;
    PA_HIP(hipGetDevice(&dev));
    PA_REQUIRE(dev >= 0 && dev < kMaxDevices, PA_ERR_LIMIT, "scratch: device ordinal out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        PA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptr[dev] = static_cast<double*>(p);
    }
    return ptr[dev];
}

int pa_scalar_function_value_noalloc(const pa_scalar_function* sf, const double* u, const double* z, double* value,
                                     void* stream) {
    return guarded([&] {
        PA_REQUIRE(sf && u && z && value, PA_ERR_ARG, "pa_scalar_function_value_noalloc: null argument");
        const pa_mesh* m = sf->mesh;
        const int phys = sf->dpn == 3 ? PA_MECHANICS : PA_THERMAL;
        const pa::Material mt = pa::material_of(phys, sf->material, sf->penalty);
        const hipStream_t st = (hipStream_t)stream;
        // one partial per workgroup of the grid-stride cell kernel
        const long nb = std::min<long>(kNoAllocPartials, (m->C + pa::kBlock - 1) / pa::kBlock);
        PA_REQUIRE(nb >= 1 && nb <= kNoAllocPartials, PA_ERR_LIMIT,
                   "pa_scalar_function_value_noalloc: more partial sums than the scratch holds");
        double* partials = scratch_on_device();
        const auto* x4 = reinterpret_cast<const double4*>(m->d_coords4.p);
        const auto* c4 = reinterpret_cast<const int4*>(m->d_conn.p);
        // every partial the sum reads is written by this call first: nothing to clear
        if (sf->dpn == 3) pa::k_energy_cells<3><<<(int)nb, pa::kBlock, 0, st>>>(m->C, x4, c4, mt, u, z, partials);
        else pa::k_energy_cells<1><<<(int)nb, pa::kBlock, 0, st>>>(m->C, x4, c4, mt, u, z, partials);
        pa::check_launch();
        pa::k_sum_partials<<<1, 64, 0, st>>>((int)nb, partials, value);
        pa::check_launch();
    });
}

int pa_scalar_function_gradient_u(const pa_scalar_function* sf, const double* u, const double* z, double* g,
                                  void* stream) {
    return guarded([&] {
        PA_REQUIRE(sf && u && z && g, PA_ERR_ARG, "pa_scalar_function_gradient_u: null argument");
        const pa_mesh* m = sf->mesh;
        const int phys = sf->dpn == 3 ? PA_MECHANICS : PA_THERMAL;
        const pa::Material mt = pa::material_of(phys, sf->material, sf->penalty);
        const hipStream_t st = (hipStream_t)stream;
        // d/du of 1/2 u.K u is K u; of the thermal energy u.K u it is 2 K u.
        pa::residual_rows(m, sf->dpn, mt, u, z, 0.0, sf->dpn == 3 ? 1.0 : 2.0, g, st);
    });
}

int pa_scalar_function_gradient_z(const pa_scalar_function* sf, const double* u, const double* z, double* g,
                                  void* stream) {
    return guarded([&] {
        PA_REQUIRE(sf && u && z && g, PA_ERR_ARG, "pa_scalar_function_gradient_z: null argument");
        const pa_mesh* m = sf->mesh;
        const int phys = sf->dpn == 3 ? PA_MECHANICS : PA_THERMAL;
        const pa::Material mt = pa::material_of(phys, sf->material, sf->penalty);
        const hipStream_t st = (hipStream_t)stream;
        const int nb = pa::row_blocks(m->N, pa::kG);
        if (sf->dpn == 3) pa::k_energy_gradz_rows<3, pa::kG><<<nb, pa::kBlock, 0, st>>>(pa::row_mesh(m), mt, u, z, g);
        else pa::k_energy_gradz_rows<1, pa::kG><<<nb, pa::kBlock, 0, st>>>(pa::row_mesh(m), mt, u, z, g);
        pa::check_launch();
    });
}

int pa_scalar_function_gradient_x(const pa_scalar_function* sf, const double* u, const double* z, double* g,
                                  void* stream) {
    return guarded([&] {
        PA_REQUIRE(sf && u && z && g, PA_ERR_ARG, "pa_scalar_function_gradient_x: null argument");
        pa::criterion_gradient_x(sf, u, z, 1.0, g, (hipStream_t)stream);
    });
}

int pa_apply_block_constraints(const pa_mesh* m, int dpn, void* vals, int precision, double* rhs, int nbc,
                               const int* dofs, const double* bcv, double scale, void* stream) {
    return guarded([&] {
        PA_REQUIRE(m && vals && rhs, PA_ERR_ARG, "pa_apply_block_constraints: null argument");
        PA_REQUIRE(dpn >= 1 && dpn <= 4, PA_ERR_ARG, "pa_apply_block_constraints: dpn out of range");
        if (nbc == 0) return;
        PA_REQUIRE(nbc > 0 && dofs && bcv, PA_ERR_ARG, "pa_apply_block_constraints: bad constraint list");
        PA_REQUIRE(precision == PA_FP64 || precision == PA_FP32, PA_ERR_ARG,
                   "pa_apply_block_constraints: unknown precision");
        // work vectors for this call, freed once the stream has run it
        const hipStream_t st = (hipStream_t)stream;
        pa::DevBuf<double> dense;
        pa::DevBuf<unsigned char> touched;
        dense.alloc((size_t)m->N * dpn);
        touched.alloc((size_t)m->N);
        pa::apply_block_constraints(m, dpn, vals, precision, rhs, nbc, dofs, bcv, scale, dense.p, touched.p, st);
        PA_HIP(hipStreamSynchronize(st));
    });
}

int pa_block_spmv(const pa_mesh* m, int bi, int bj, const void* vals, int precision, const double* x, double beta,
                  double* y, void* stream) {
    return guarded([&] {
        PA_REQUIRE(m && vals && x && y, PA_ERR_ARG, "pa_block_spmv: null argument");
        const hipStream_t st = (hipStream_t)stream;
        const int nb = (int)(((long)m->N * 4 + pa::kBlock - 1) / pa::kBlock);
        const int* rm = m->d_rowmap.p;
        const int* ci = m->d_colidx.p;
#define PA_SPMV(BI, BJ)                                                                                          \
    if (bi == BI && bj == BJ) {                                                                                  \
        if (precision == PA_FP64)                                                                                \
            pa::k_block_spmv<BI, BJ, double><<<nb, pa::kBlock, 0, st>>>(m->N, rm, ci, (const double*)vals, x,    \
                                                                         beta, y);                               \
        else                                                                                                     \
            pa::k_block_spmv<BI, BJ, float><<<nb, pa::kBlock, 0, st>>>(m->N, rm, ci, (const float*)vals, x,      \
                                                                        beta, y);                                \
        pa::check_launch();                                                                                      \
        return;                                                                                                  \
    }
        PA_REQUIRE(precision == PA_FP64 || precision == PA_FP32, PA_ERR_ARG, "pa_block_spmv: unknown precision");
        PA_SPMV(3, 3)
        PA_SPMV(1, 1)
        PA_SPMV(1, 3)
        PA_SPMV(3, 1)
        PA_SPMV(4, 4)
#undef PA_SPMV
        throw pa::Error(PA_ERR_ARG, "pa_block_spmv: unsupported block shape");
    });
}

}  // extern "C"
