// Geometric multigrid preconditioner for the structured Kuhn box meshes.
//
// Levels coarsen the hex grid by two in every direction while all cell counts
// stay even.  Transfer is trilinear interpolation on the grid vertices
// (fine vertex i takes weight 1 from the coarse vertex at i/2 when i is even,
// 1/2 from each neighbour when i is odd, per direction).  Coarse operators
// are Galerkin products A_{l+1} = P^T A_l P, computed on the device and
// stored as 27-point block stencils; the finest level is an FP32 copy of the
// assembled block-CRS matrix laid out as a stencil over the offsets the mesh
// graph holds (15 on Kuhn grids; symmetric operators store the diagonal and
// the 7 blocks towards larger offsets only), its Dirichlet dofs masked out of
// P.  Smoother:
// damped block Jacobi, symmetric pre/post sweeps, so the V-cycle is an SPD
// preconditioner for CG.  Coarsest level: dense inverse (FP64 Gauss-Jordan in
// one workgroup).  Level matrices and vectors are FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pa_cell.h"
#include "pa_internal.h"
#include "pa_stencil.h"

namespace pa {
namespace {

constexpr int kT = 256;
constexpr int kChebMax = 8;  // most sweeps per smoothing with Chebyshev dampings

__device__ inline int sidx(int dx, int dy, int dz) { return (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1); }

// Block inverse in double, stored as float: cofactors for 3x3, Gauss-Jordan
// with partial pivoting for 4x4 (the nodal blocks of the mixed u-p tangent
// are indefinite), reciprocal for 1x1.
template <int D>
__device__ inline void block_inverse(const double* a, float* out) {
    if constexpr (D == 4) {
        double A[4][8];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                A[i][j] = a[i * 4 + j];
                A[i][4 + j] = (i == j) ? 1.0 : 0.0;
            }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int piv = k;
            double best = fabs(A[k][k]);
#pragma unroll
            for (int i = k + 1; i < 4; ++i)
                if (fabs(A[i][k]) > best) {
                    best = fabs(A[i][k]);
                    piv = i;
                }
#pragma unroll
            for (int i = k + 1; i < 4; ++i)
                if (i == piv)
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const double tmp = A[k][j];
                        A[k][j] = A[i][j];
                        A[i][j] = tmp;
                    }
            const double inv = A[k][k] != 0.0 ? 1.0 / A[k][k] : 0.0;
#pragma unroll
            for (int j = 0; j < 8; ++j) A[k][j] *= inv;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (i != k) {
                    const double f = A[i][k];
#pragma unroll
                    for (int j = 0; j < 8; ++j) A[i][j] -= f * A[k][j];
                }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) out[i * 4 + j] = (float)A[i][4 + j];
    } else if constexpr (D == 3) {
        const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[2] * a[7] - a[1] * a[8], c02 = a[1] * a[5] - a[2] * a[4];
        const double c10 = a[5] * a[6] - a[3] * a[8], c11 = a[0] * a[8] - a[2] * a[6], c12 = a[2] * a[3] - a[0] * a[5];
        const double c20 = a[3] * a[7] - a[4] * a[6], c21 = a[1] * a[6] - a[0] * a[7], c22 = a[0] * a[4] - a[1] * a[3];
        const double det = a[0] * c00 + a[1] * c10 + a[2] * c20;
        const double id = det != 0.0 ? 1.0 / det : 0.0;
        out[0] = (float)(c00 * id); out[1] = (float)(c01 * id); out[2] = (float)(c02 * id);
        out[3] = (float)(c10 * id); out[4] = (float)(c11 * id); out[5] = (float)(c12 * id);
        out[6] = (float)(c20 * id); out[7] = (float)(c21 * id); out[8] = (float)(c22 * id);
    } else {
        out[0] = a[0] != 0.0 ? (float)(1.0 / a[0]) : 0.0f;
    }
}

// ------------------------------------------------------- fine level (stencil)
template <int D>
__global__ void k_diag_inv_fst(int N, BlockStencil fs, const float* __restrict__ A, float* __restrict__ dinv) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double a[D * D];
#pragma unroll
    for (int q = 0; q < D * D; ++q) a[q] = (double)A[((long)n * fs.SS + fs.diag) * D * D + q];
    block_inverse<D>(a, dinv + (long)n * D * D);
}


template <int D>
__global__ void k_diag_inv_st(int N, const float* __restrict__ A, float* __restrict__ dinv) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double a[D * D];
#pragma unroll
    for (int q = 0; q < D * D; ++q) a[q] = (double)A[((long)n * 27 + 13) * D * D + q];
    block_inverse<D>(a, dinv + (long)n * D * D);
}

// y = b - A x on a level's stencil, G lanes per node, kFsSlots offsets per
// lane; masked dofs get 0 (mask null: none).  zero_x: x is zero (y = b).
// With omega != 0 the result is the Jacobi update x_out = x + omega Dinv
// (b - A x) instead (omega_dev non-null: the damping read from there).  Branch-free: an offset that is absent or points out of
// the grid reads the node's own block and x and adds zero, so every load of
// the lane issues at once.  SYM: some offsets are read transposed.  Lane
// l < D finishes dof l.
template <int D, int G, bool SYM>
__global__ __launch_bounds__(kT) void k_fst_resid_jac(Grid g, BlockStencil fs, const float* __restrict__ A,
                                                      const float* __restrict__ dinv,
                                                      const unsigned char* __restrict__ mask,
                                                      const float* __restrict__ b, const float* __restrict__ x,
                                                      float* __restrict__ out, float omega,
                                                      const float* __restrict__ omega_dev, int zero_x,
                                                      const double* __restrict__ b64, float* __restrict__ b32,
                                                      double* __restrict__ out64) {
    constexpr int ROWS = kT / G, BS = D * D, NS = kFsSlots;
    if (omega_dev) omega = omega_dev[0];  // a Chebyshev step's damping, computed on the device
    const int l = threadIdx.x & (G - 1);
    const int n = xcd_block() * ROWS + threadIdx.x / G;
    if (n >= g.nodes()) return;
    int i, j, k;
    stencil_ijk(g, fs, n, i, j, k);
    // the node's dof min(l, D - 1) for the update, in flight with the stencil
    const int ld = l < D ? l : D - 1;
    // b64: the right-hand side in FP64 (the V-cycle's first sweep), kept as
    // b32 for the level's later sweeps
    const float bl = b64 ? (float)b64[(long)n * D + ld] : b[(long)n * D + ld], xl = x[(long)n * D + ld];
    const unsigned ml = mask ? mask[(long)n * D + ld] : 1u;
    float dl[D];
#pragma unroll
    for (int q = 0; q < D; ++q) dl[q] = dinv[((long)n * D + ld) * D + q];
    float acc[D];
#pragma unroll
    for (int p = 0; p < D; ++p) acc[p] = 0.0f;
    if (!zero_x) {
        float blk[NS][BS], xc[NS][D];
        bool ok[NS], tr[NS];
#pragma unroll
        for (int m = 0; m < NS; ++m) {
            long c, base;
            stencil_slot<G, SYM>(g, fs, m, l, n, i, j, k, BS, ok[m], tr[m], c, base);
            const float* v = A + base;
#pragma unroll
            for (int q = 0; q < BS; ++q) blk[m][q] = v[q];
#pragma unroll
            for (int q = 0; q < D; ++q) xc[m][q] = x[c * D + q];
        }
#pragma unroll
        for (int m = 0; m < NS; ++m) {
#pragma unroll
            for (int p = 0; p < D; ++p) {
                float yn = 0.0f, yt = 0.0f;
#pragma unroll
                for (int q = 0; q < D; ++q) {
                    yn = fmaf(blk[m][p * D + q], xc[m][q], yn);
                    if constexpr (SYM) yt = fmaf(blk[m][q * D + p], xc[m][q], yt);
                }
                const float y = (SYM && tr[m]) ? yt : yn;
                acc[p] += ok[m] ? y : 0.0f;
            }
        }
#pragma unroll
        for (int p = 0; p < D; ++p) acc[p] = node_lanes_sum<G>(acc[p]);
    }
    // r_ld on lane l, then r_q from lane q of the quad for the block row
    float al = acc[0];
#pragma unroll
    for (int p = 1; p < D; ++p) al = ld == p ? acc[p] : al;
    const float rl = bl - al;
    float t = 0.0f;
    if (omega != 0.0f) {
        float rq[D];
        rq[0] = dpp_f32<0x00>(rl);  // quad_perm [0, 0, 0, 0]
        if constexpr (D > 1) rq[1] = dpp_f32<0x55>(rl);
        if constexpr (D > 2) rq[2] = dpp_f32<0xaa>(rl);
        if constexpr (D > 3) rq[3] = dpp_f32<0xff>(rl);
#pragma unroll
        for (int q = 0; q < D; ++q) t += dl[q] * rq[q];
    }
    if (l < D) {
        const float v = ml ? (omega != 0.0f ? (zero_x ? 0.0f : xl) + omega * t : rl) : 0.0f;
        if (out64) out64[(long)n * D + l] = (double)v;  // the V-cycle's last sweep writes z in FP64
        else out[(long)n * D + l] = v;
        if (b64) b32[(long)n * D + l] = bl;
    }
}

// ------------------------------------------------------------------ transfer
// b_c[J] = sum_i w_iJ mask_i r_f[i]
// A quad of lanes per coarse node: lane l takes the fine vertices 2J +
// offset(t), t = l, l + 4, ... < 27, branch-free (a vertex outside the fine
// grid reads the centre vertex with weight 0), then the quad sums by DPP and
// lane p < D writes component p.  MASK: bit p of mbits[i] is dof p of fine
// vertex i being free.
template <int D, bool MASK>
__global__ __launch_bounds__(kT) void k_restrict(Grid gf, Grid gc, const float* __restrict__ rf,
                                                 const unsigned char* __restrict__ mbits, float* __restrict__ bc) {
    constexpr int NODES = kT / 4, NV = 7;
    const int l = threadIdx.x & 3;
    const int n = blockIdx.x * NODES + (threadIdx.x >> 2);
    if (n >= gc.nodes()) return;
    int I, J, K;
    gc.ijk(n, I, J, K);
    const long f0 = gf.id(2 * I, 2 * J, 2 * K);
    float v[NV][D], w[NV];
    unsigned b[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) {
        const int t = l + 4 * m;
        const int a = t / 9 - 1, c = (t / 3) % 3 - 1, e = t % 3 - 1;
        const int i = 2 * I + a, j = 2 * J + c, k = 2 * K + e;
        const bool ok = t < 27 && gf.in(i, j, k);
        w[m] = ok ? (a ? 0.5f : 1.0f) * (c ? 0.5f : 1.0f) * (e ? 0.5f : 1.0f) : 0.0f;
        const long f = ok ? gf.id(i, j, k) : f0;
#pragma unroll
        for (int p = 0; p < D; ++p) v[m][p] = rf[f * D + p];
        b[m] = MASK ? mbits[f] : 0xffu;
    }
    float acc[D];
#pragma unroll
    for (int p = 0; p < D; ++p) {
        acc[p] = 0.0f;
#pragma unroll
        for (int m = 0; m < NV; ++m) acc[p] += ((b[m] >> p) & 1u) ? w[m] * v[m][p] : 0.0f;
        acc[p] = node_lanes_sum<4>(acc[p]);
    }
    float al = acc[0];
#pragma unroll
    for (int p = 1; p < D; ++p) al = l == p ? acc[p] : al;
    if (l < D) bc[(long)n * D + l] = al;
}

// x_f[i] += mask_i sum_J w_iJ x_c[J]: thread per fine node, its eight
// candidate parents per trilinear cell read branch-free (a parent repeated
// for an even coordinate has weight 0).
template <int D, bool MASK>
__global__ void k_prolong_add(Grid gf, Grid gc, const float* __restrict__ xc, const unsigned char* __restrict__ mbits,
                              float* __restrict__ xf) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= gf.nodes()) return;
    int i, j, k;
    gf.ijk(n, i, j, k);
    const int I[2] = {i >> 1, (i + 1) >> 1}, J[2] = {j >> 1, (j + 1) >> 1}, K[2] = {k >> 1, (k + 1) >> 1};
    const float wi[2] = {(i & 1) ? 0.5f : 1.0f, (i & 1) ? 0.5f : 0.0f};
    const float wj[2] = {(j & 1) ? 0.5f : 1.0f, (j & 1) ? 0.5f : 0.0f};
    const float wk[2] = {(k & 1) ? 0.5f : 1.0f, (k & 1) ? 0.5f : 0.0f};
    float v[8][D];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const long c = gc.id(I[q >> 2], J[(q >> 1) & 1], K[q & 1]);
#pragma unroll
        for (int p = 0; p < D; ++p) v[q][p] = xc[c * D + p];
    }
    const unsigned b = MASK ? mbits[n] : 0xffu;
#pragma unroll
    for (int p = 0; p < D; ++p) {
        float acc = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q) acc += (wi[q >> 2] * wj[(q >> 1) & 1] * wk[q & 1]) * v[q][p];
        if ((b >> p) & 1u) xf[(long)n * D + p] += acc;
    }
}

// --------------------------------------------------------- Galerkin products
// Galerkin product by gather.  Block A_c(I, I + offset(S)) is a fixed-order
// sum over the pairs (a, d) of the list of S: fine node i = 2I + a (a in
// {-1, 0, 1}^3), its neighbour j = i + d (d an offset of the source stencil,
// with e = j - 2(I + offset(S)) in {-1, 0, 1}^3), weight w(i, I) w(j, J) =
// 2^-(nonzeros of a and e), and the source block A(i, j) (read transposed
// from node j when the source stencil keeps half); pairs whose nodes fall
// outside the fine grid add nothing.  One wave per coarse node: the lists
// are cut into at most 64 chunks of about equal length, one per lane, and
// the chunks of one list are added in lane order through LDS.  No atomics:
// deterministic.  MASK: the fine level's Dirichlet dofs (bit p of mbits[n])
// are masked out of P.  Entry bits: a (0-4), source slot (5-9), transposed
// (10), d (11-15), weight exponent (16-18).
struct RapLanes {
    int lane[64];  // S (bits 0-4), chunk start (5-18), chunk length (19-26), first lane of S (bit 27 set)
    int nlanes;
};

template <int D, bool MASK>
__global__ __launch_bounds__(256) void k_rap_gather(Grid gf, Grid gc, RapLanes rl, const unsigned* __restrict__ tab,
                                                    const float* __restrict__ Af, int SSf,
                                                    const unsigned char* __restrict__ mbits, float* __restrict__ Ac) {
    constexpr int BS = D * D;
    __shared__ float part[4][64][BS];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int cn = blockIdx.x * 4 + w;
    const bool active = cn < gc.nodes() && l < rl.nlanes;
    const int desc = active ? rl.lane[l] : 0;
    const int S = desc & 31;
    int I = 0, J = 0, K = 0;
    if (cn < gc.nodes()) gc.ijk(cn, I, J, K);
    const int CI = I + S / 9 - 1, CJ = J + (S / 3) % 3 - 1, CK = K + S % 3 - 1;
    float acc[BS];
#pragma unroll
    for (int q = 0; q < BS; ++q) acc[q] = 0.0f;
    if (active && gc.in(CI, CJ, CK)) {
        const int t0 = (desc >> 5) & 0x3fff, t1 = t0 + ((desc >> 19) & 0xff);
        // branch-free body, U pairs per trip so that their loads are in
        // flight together; a pair outside the grid or past the chunk reads the
        // coarse node's own fine node and adds zero
        constexpr int U = D == 4 ? 2 : 4;
        const long f0 = gf.id(2 * I, 2 * J, 2 * K);
        for (int t = t0; t < t1; t += U) {
            float v[U][BS], wgt[U];
            unsigned bi[U], bj[U], eu[U];
            bool tr[U];
#pragma unroll
            for (int u = 0; u < U; ++u) eu[u] = tab[min(t + u, t1 - 1)];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const unsigned e = eu[u];
                const int as = (int)(e & 31u), slot = (int)((e >> 5) & 31u), ds = (int)((e >> 11) & 31u);
                tr[u] = (e >> 10) & 1u;
                const int i = 2 * I + as / 9 - 1, j = 2 * J + (as / 3) % 3 - 1, k = 2 * K + as % 3 - 1;
                const int ii = i + ds / 9 - 1, jj = j + (ds / 3) % 3 - 1, kk = k + ds % 3 - 1;
                const bool ok = ((int)(t + u < t1) & (int)gf.in(i, j, k) & (int)gf.in(ii, jj, kk)) != 0;
                wgt[u] = ok ? __builtin_amdgcn_ldexpf(1.0f, -(int)((e >> 16) & 7u)) : 0.0f;
                const long fi = ok ? gf.id(i, j, k) : f0, fj = ok ? gf.id(ii, jj, kk) : f0;
                const float* blk = Af + ((tr[u] ? fj : fi) * SSf + slot) * BS;
#pragma unroll
                for (int q = 0; q < BS; ++q) v[u][q] = blk[q];
                bi[u] = MASK ? mbits[fi] : 0xffu;
                bj[u] = MASK ? mbits[fj] : 0xffu;
            }
            // the blocks are used from registers (not re-read through selected addresses)
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int q = 0; q < BS; ++q) asm volatile("" : "+v"(v[u][q]));
#pragma unroll
            for (int u = 0; u < U; ++u) {
                float mi[D], mj[D];
#pragma unroll
                for (int p = 0; p < D; ++p) {
                    mi[p] = ((bi[u] >> p) & 1u) ? wgt[u] : 0.0f;
                    mj[p] = ((bj[u] >> p) & 1u) ? 1.0f : 0.0f;
                }
#pragma unroll
                for (int p = 0; p < D; ++p)
#pragma unroll
                    for (int q = 0; q < D; ++q) {
                        const float a = tr[u] ? v[u][q * D + p] : v[u][p * D + q];
                        acc[p * D + q] = fmaf(mi[p] * mj[q], a, acc[p * D + q]);
                    }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < BS; ++q) part[w][l][q] = acc[q];
    __syncthreads();
    if (active && ((desc >> 27) & 1)) {
        // this lane is the first of its list: the list's chunks in lane order
        float sum[BS];
#pragma unroll
        for (int q = 0; q < BS; ++q) sum[q] = part[w][l][q];
        for (int r = l + 1; r < rl.nlanes && (rl.lane[r] & 31) == S; ++r)
#pragma unroll
            for (int q = 0; q < BS; ++q) sum[q] += part[w][r][q];
        float* dst = Ac + ((long)cn * 27 + S) * BS;
#pragma unroll
        for (int q = 0; q < BS; ++q) dst[q] = sum[q];
    }
}

// Bit p of bits[n]: dof p of node n is free.
__global__ void k_mask_bits(int N, int D, const unsigned char* __restrict__ mask, unsigned char* __restrict__ bits) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    unsigned b = 0;
    for (int p = 0; p < D; ++p) b |= mask[(long)n * D + p] ? 1u << p : 0u;
    bits[n] = (unsigned char)b;
}

// ------------------------------------------------------------ coarsest level
template <int D>
__global__ void k_dense_from_stencil(Grid g, const float* __restrict__ A, double* __restrict__ dense) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int nd = g.nodes() * D;
    if (n >= g.nodes()) return;
    int i, j, k;
    g.ijk(n, i, j, k);
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                if (!g.in(i + dx, j + dy, k + dz)) continue;
                const int c = g.id(i + dx, j + dy, k + dz);
                const float* v = A + ((long)n * 27 + sidx(dx, dy, dz)) * D * D;
                for (int p = 0; p < D; ++p)
                    for (int q = 0; q < D; ++q) dense[(long)(n * D + p) * nd + c * D + q] = v[p * D + q];
            }
}

// In-place Gauss-Jordan inverse with partial pivoting (rows swapped in both
// matrices; ties go to the lower row, so the result is deterministic), one
// workgroup of 1024 threads as 32 rows of 32 lanes: lane tx of row group ty
// updates the columns tx, tx + 32, ... of the rows ty, ty + 32, ...
__global__ __launch_bounds__(1024) void k_gauss_jordan(int n, double* __restrict__ a, double* __restrict__ inv) {
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < n; r += 32)
        for (int c = tx; c < n; c += 32) inv[(long)r * n + c] = r == c ? 1.0 : 0.0;
    __syncthreads();
    __shared__ double sval[1024];
    __shared__ int sidx_[1024];
    __shared__ double piv;
    for (int k = 0; k < n; ++k) {
        double best = -1.0;
        int bi = k;
        for (int r = k + (int)threadIdx.x; r < n; r += 1024) {
            const double v = fabs(a[(long)r * n + k]);
            if (v > best) {
                best = v;
                bi = r;
            }
        }
        sval[threadIdx.x] = best;
        sidx_[threadIdx.x] = bi;
        __syncthreads();
        for (int s = 512; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) {
                const double o = sval[threadIdx.x + s];
                const int oi = sidx_[threadIdx.x + s];
                if (o > sval[threadIdx.x] || (o == sval[threadIdx.x] && oi < sidx_[threadIdx.x])) {
                    sval[threadIdx.x] = o;
                    sidx_[threadIdx.x] = oi;
                }
            }
            __syncthreads();
        }
        const int pr = sidx_[0];
        if (pr != k)
            for (int c = threadIdx.x; c < n; c += 1024) {
                double t0 = a[(long)k * n + c];
                a[(long)k * n + c] = a[(long)pr * n + c];
                a[(long)pr * n + c] = t0;
                t0 = inv[(long)k * n + c];
                inv[(long)k * n + c] = inv[(long)pr * n + c];
                inv[(long)pr * n + c] = t0;
            }
        __syncthreads();
        if (threadIdx.x == 0) piv = a[(long)k * n + k];
        __syncthreads();
        const double ip = piv != 0.0 ? 1.0 / piv : 0.0;
        for (int c = threadIdx.x; c < n; c += 1024) {
            a[(long)k * n + c] *= ip;
            inv[(long)k * n + c] *= ip;
        }
        __syncthreads();
        for (int r = ty; r < n; r += 32) {
            if (r == k) continue;
            const double f = a[(long)r * n + k];
            for (int c = tx; c < n; c += 32) {
                if (c != k) a[(long)r * n + c] -= f * a[(long)k * n + c];
                inv[(long)r * n + c] -= f * inv[(long)k * n + c];
            }
        }
        __syncthreads();
        for (int r = threadIdx.x; r < n; r += 1024)
            if (r != k) a[(long)r * n + k] = 0.0;
        __syncthreads();
    }
}

// Gauss-Jordan inverse of a larger coarsest level, one launch per pivot step:
// M = [A | I] (n x 2n) ping-pongs between two buffers and rows are not
// swapped.  Every block finds the same pivot (largest magnitude in column k
// among the rows not yet pivots, ties to the lower row; used[r] = k + 1
// marks the pivot row of step k, so a block of step k never counts the row
// marked in this step), scales the pivot row and eliminates column k from
// the other rows.  Row k of the inverse is then the right half of the pivot
// row of step k.
__global__ void k_gj_init(int n, const double* __restrict__ a, double* __restrict__ m, int* __restrict__ used) {
    const long w = 2L * n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < (long)n * w; e += (long)gridDim.x * blockDim.x) {
        const int r = (int)(e / w), c = (int)(e % w);
        m[e] = c < n ? a[(long)r * n + c] : (c - n == r ? 1.0 : 0.0);
        if (c == 0) used[r] = 0;
    }
}

__global__ __launch_bounds__(256) void k_gj_step(int n, int k, const double* __restrict__ mi, double* __restrict__ mo,
                                                 int* __restrict__ used, int* __restrict__ pivrow) {
    __shared__ double sv[256];
    __shared__ int si[256];
    const long w = 2L * n;
    double best = -1.0;
    int bi = n;
    for (int r = threadIdx.x; r < n; r += 256) {
        const int u = used[r];
        if (u > 0 && u <= k) continue;
        const double v = fabs(mi[(long)r * w + k]);
        if (v > best) {
            best = v;
            bi = r;
        }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            const double o = sv[threadIdx.x + h];
            const int oi = si[threadIdx.x + h];
            if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                sv[threadIdx.x] = o;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    const int p = si[0];
    if (p >= n) return;  // no row left (cannot happen for k < n)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        used[p] = k + 1;
        pivrow[k] = p;
    }
    const double pv = mi[(long)p * w + k];
    const double ip = pv != 0.0 ? 1.0 / pv : 0.0;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < (long)n * w; e += (long)gridDim.x * 256) {
        const int r = (int)(e / w), c = (int)(e % w);
        const double prow = mi[(long)p * w + c] * ip;
        mo[e] = r == p ? prow : (c == k ? 0.0 : mi[e] - mi[(long)r * w + k] * prow);
    }
}

__global__ void k_gj_extract(int n, const double* __restrict__ m, const int* __restrict__ pivrow, double* __restrict__ inv) {
    const long w = 2L * n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < (long)n * n; e += (long)gridDim.x * blockDim.x) {
        const int k = (int)(e / n), j = (int)(e % n);
        inv[e] = m[(long)pivrow[k] * w + n + j];
    }
}

// x = inv b: one wave per row, lanes over the columns, FP64 sums in a fixed
// order.
__global__ __launch_bounds__(256) void k_dense_apply(int n, const double* __restrict__ inv, const float* __restrict__ b,
                                                     float* __restrict__ x) {
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (r >= n) return;
    double s = 0.0;
    for (int c = lane; c < n; c += 64) s = fma(inv[(long)r * n + c], (double)b[c], s);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) x[r] = (float)s;
}

// ------------------------------------------------- Chebyshev smoothing setup
// Power iteration for the largest eigenvalue of D^-1 A on a level: a fixed
// start vector (zero on masked dofs), then v <- D^-1 A v = v - S v, S v the
// sweep with b = 0 and omega = 1; the estimate is ||M^K v|| / ||M^(K-1) v||.
__global__ void k_pow_start(long n, const unsigned char* __restrict__ mask, float* __restrict__ v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned h = (unsigned)i * 2654435761u;
    v[i] = (mask && !mask[i]) ? 0.0f : 0.5f + (float)((h >> 8) & 0xffffu) * (1.0f / 65536.0f);
}

__global__ void k_fill_f(long n, float* __restrict__ v, float value) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = value;
}

__device__ inline double block_sum256(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    const double t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return t;
}

// v <- v - sv, and per-block partial sums of the new v squared
__global__ __launch_bounds__(256) void k_pow_diff(long n, float* __restrict__ v, const float* __restrict__ sv,
                                                  double* __restrict__ part) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const float y = v[i] - sv[i];
        v[i] = y;
        acc += (double)y * y;
    }
    const double t = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Chebyshev dampings on [b/12, b], b = 1.2 lambda (power iteration
// underestimates the largest eigenvalue; the wide interval measured best at
// 128x64x64 elastic: 83 PCG iterations against 92 on [b/4, 1.15 lambda] and
// 95 with damped Jacobi at 0.6): pre-smoothing step k uses
// 1 / (theta + delta cos(pi (2k + 1) / (2 m))), m = npre; post-smoothing runs
// the npost roots in reverse order (the adjoint), so the V-cycle stays
// symmetric.
__global__ __launch_bounds__(256) void k_cheb_dampings(const double* __restrict__ p0, const double* __restrict__ p1,
                                                       int np, int npre, int npost, float* __restrict__ om) {
    __shared__ double sh[4];
    double a0 = 0.0, a1 = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) {
        a0 += p0[i];
        a1 += p1[i];
    }
    const double s0 = block_sum256(a0, sh), s1 = block_sum256(a1, sh);
    if (threadIdx.x != 0) return;
    const double lam = s0 > 0.0 ? sqrt(s1 / s0) : 1.0;
    const double b = 1.2 * lam, a = b / 12.0, th = 0.5 * (a + b), de = 0.5 * (b - a);
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < npre; ++k) om[k] = (float)(1.0 / (th + de * cos(pi * (2 * k + 1) / (2.0 * npre))));
    for (int s = 0; s < npost; ++s) {
        const int k = npost - 1 - s;
        om[kChebMax + s] = (float)(1.0 / (th + de * cos(pi * (2 * k + 1) / (2.0 * npost))));
    }
}

// ----------------------------------------------------------------- helpers
__global__ void k_mask_init(long n, unsigned char* m) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) m[i] = 1;
}
__global__ void k_mask_bc(int nbc, const int* __restrict__ dofs, unsigned char* m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nbc) m[dofs[i]] = 0;
}

inline int nblk(long n, int t = kT) { return (int)((n + t - 1) / t); }
void chk() { PA_HIP(hipGetLastError()); }

}  // namespace
}  // namespace pa

struct pa_multigrid {
    const pa_mesh* mesh = nullptr;
    int dpn = 3;
    pa_multigrid_options opt{};
    std::vector<pa::Grid> grids;  // grids[0] is the finest
    pa::DevBuf<unsigned char> mask;
    std::vector<pa::DevBuf<float>*> A, dinv, b, x, t;  // per level (A[0] unused: the fine stencil)
    pa::DevBuf<double> dense, dense_inv;
    pa::DevBuf<double> gj[2];           // [A | I] ping-pong of the stepwise Gauss-Jordan inverse
    pa::DevBuf<int> gj_used, gj_piv;
    unsigned graph_offsets = 0;  // bit sidx: the mesh graph holds that offset
    bool symmetric = false;      // the fine operator is symmetric: half stencil
    std::vector<pa::BlockStencil> fsl;  // stencil of each level's smoother
    pa::RapLanes rap_fine{}, rap_coarse{};  // Galerkin pair tables: from the fine stencil, from a 27-point one
    pa::DevBuf<unsigned> d_rap_fine, d_rap_coarse;
    pa::DevBuf<unsigned char> mbits;  // free dofs of each fine node as bits
    int rap_half = -1;  // fine layout the tables were built for (-1: none)
    pa::DevBuf<float> fine;  // the fine matrix as an FP32 stencil
    bool chebyshev = false;       // Chebyshev smoothing (else damped block Jacobi)
    pa::DevBuf<float> omegas;     // per level: kChebMax pre-smoothing, then kChebMax post-smoothing dampings
    pa::DevBuf<double> powpart;   // power-iteration partial sums (two sets)
    bool ready = false;
    ~pa_multigrid() {
        for (auto* v : {&A, &dinv, &b, &x, &t})
            for (auto* p : *v) delete p;
    }
};

namespace pa {
namespace {

// One smoothing sweep (or, with om = 0, the residual) on level l.
template <int D>
void sweep_level(pa_multigrid* mg, int l, const float* b, const float* xin, float* xout, float om,
                 const float* om_dev, int zero, hipStream_t st, const double* b64 = nullptr, float* b32 = nullptr,
                 double* out64 = nullptr) {
    const Grid g = mg->grids[l];
    const BlockStencil& f = mg->fsl[l];
    auto run = [&](auto kernel) {
        kernel<<<nblk(g.nodes(), kT / f.G), kT, 0, st>>>(g, f, l == 0 ? mg->fine.p : mg->A[l]->p, mg->dinv[l]->p,
                                                         l == 0 ? mg->mask.p : nullptr, b, xin, xout, om, om_dev,
                                                         zero, b64, b32, out64);
    };
    if (f.G == 4) {
        if (f.half) run(k_fst_resid_jac<D, 4, true>);
        else run(k_fst_resid_jac<D, 4, false>);
    } else {
        if (f.half) run(k_fst_resid_jac<D, 8, true>);
        else run(k_fst_resid_jac<D, 8, false>);
    }
    chk();
}

// Smoothing sweeps of level l (pre or post): coarse_sweeps on the coarse
// levels of at most coarse_max_nodes nodes when set.
int level_sweeps(const pa_multigrid* mg, int l, bool pre) {
    const int dflt = std::max(1, pre ? mg->opt.pre_sweeps : mg->opt.post_sweeps);
    const int lim = mg->opt.coarse_max_nodes > 0 ? mg->opt.coarse_max_nodes : 16384;
    return (l > 0 && mg->opt.coarse_sweeps > 0 && mg->grids[l].nodes() <= lim) ? mg->opt.coarse_sweeps : dflt;
}

// r64/z64 (level 0 only): the FP64 right-hand side, read by the first sweep,
// and the FP64 correction, written by the last sweep.
template <int D>
void vcycle(pa_multigrid* mg, int l, hipStream_t st, const double* r64 = nullptr, double* z64 = nullptr) {
    const int L = (int)mg->grids.size();
    const Grid g = mg->grids[l];
    const long n = (long)g.nodes() * D;
    float* b = mg->b[l]->p;
    float* x = mg->x[l]->p;
    float* t = mg->t[l]->p;
    const float w = (float)mg->opt.omega;
    if (l == L - 1) {
        k_dense_apply<<<(int)((n + 3) / 4), 256, 0, st>>>((int)n, mg->dense_inv.p, b, x);
        chk();
        return;
    }
    // damping of pre-smoothing sweep s (post: s + kChebMax)
    auto dev = [&](int s) { return mg->chebyshev ? mg->omegas.p + (long)l * 2 * kChebMax + s : nullptr; };
    auto sweep = [&](const float* xin, float* xout, float om, const float* om_dev, int zero,
                     const double* in64 = nullptr, double* out64 = nullptr) {
        sweep_level<D>(mg, l, b, xin, xout, om, om_dev, zero, st, in64, in64 ? b : nullptr, out64);
    };
    // pre-smoothing (ping-pong x <-> t; the first sweep starts from zero)
    const int pre = level_sweeps(mg, l, true);
    sweep(x, x, w, dev(0), 1, r64);
    for (int s = 1; s < pre; ++s) {
        sweep(x, t, w, dev(s), 0);
        std::swap(mg->x[l], mg->t[l]);
        x = mg->x[l]->p;
        t = mg->t[l]->p;
    }
    // residual, restriction, coarse correction
    sweep(x, t, 0.0f, nullptr, 0);
    const Grid gc = mg->grids[l + 1];
    if (l == 0)
        k_restrict<D, true><<<nblk(gc.nodes(), kT / 4), kT, 0, st>>>(g, gc, t, mg->mbits.p, mg->b[l + 1]->p);
    else
        k_restrict<D, false><<<nblk(gc.nodes(), kT / 4), kT, 0, st>>>(g, gc, t, nullptr, mg->b[l + 1]->p);
    chk();
    vcycle<D>(mg, l + 1, st);
    if (l == 0)
        k_prolong_add<D, true><<<nblk(g.nodes()), kT, 0, st>>>(g, gc, mg->x[l + 1]->p, mg->mbits.p, x);
    else
        k_prolong_add<D, false><<<nblk(g.nodes()), kT, 0, st>>>(g, gc, mg->x[l + 1]->p, nullptr, x);
    chk();
    // post-smoothing
    for (int s = 0, post = level_sweeps(mg, l, false); s < post; ++s) {
        sweep(x, t, w, dev(kChebMax + s), 0, nullptr, s == post - 1 ? z64 : nullptr);
        std::swap(mg->x[l], mg->t[l]);
        x = mg->x[l]->p;
        t = mg->t[l]->p;
    }
}

// A coarse level's smoother stencil: all 27 offsets stored (the Galerkin
// stencil); with a symmetric operator the offsets of smaller stencil index
// are read transposed from the neighbour, so the smoother's operator is
// symmetric to the last bit.
BlockStencil make_coarse_stencil(const Grid& g, bool symmetric) {
    BlockStencil f{};
    f.S = f.SS = 27;
    f.G = 8;
    f.half = symmetric;
    f.diag = 13;
    for (int q = 0; q < 27; ++q) {
        const bool tr = symmetric && q < 13;
        const unsigned info = (unsigned)(q / 9) | (unsigned)((q / 3) % 3) << 2 | (unsigned)(q % 3) << 4 |
                              (tr ? 1u << 6 : 0u) | 1u << 7 | (unsigned)(tr ? 26 - q : q) << 8;
        const int m = q / 8, l = q % 8;
        f.lane[m][l / 2] |= info << ((l % 2) * 16);
        f.ssig[q / 6] |= (unsigned)q << ((q % 6) * 5);
    }
    f.rny = 1.0 / (g.ny + 1);
    f.rnz = 1.0 / (g.nz + 1);
    return f;
}

// The pair lists of the gather Galerkin product for a source stencil, and
// their cut into at most 64 lane chunks.
RapLanes make_rap_table(const BlockStencil& src, std::vector<unsigned>& tab) {
    struct Src {
        bool ok;
        int slot, tr;
    } of[27] = {};
    for (int s = 0; s < src.S; ++s) {
        const int m = s / src.G, l = s % src.G;
        const unsigned info = (src.lane[m][l / 2] >> ((l % 2) * 16)) & 0xffffu;
        const int sig = (int)(info & 3u) * 9 + (int)((info >> 2) & 3u) * 3 + (int)((info >> 4) & 3u);
        of[sig] = Src{true, (int)((info >> 8) & 31u), (int)((info >> 6) & 1u)};
    }
    tab.clear();
    int off[28];
    for (int S = 0; S < 27; ++S) {
        off[S] = (int)tab.size();
        for (int a = 0; a < 27; ++a)
            for (int d = 0; d < 27; ++d) {
                if (!of[d].ok) continue;
                bool in = true;
                int nz = 0;
                for (int c = 0; c < 3; ++c) {
                    const int av = (c == 0 ? a / 9 : (c == 1 ? (a / 3) % 3 : a % 3)) - 1;
                    const int dv = (c == 0 ? d / 9 : (c == 1 ? (d / 3) % 3 : d % 3)) - 1;
                    const int sv = (c == 0 ? S / 9 : (c == 1 ? (S / 3) % 3 : S % 3)) - 1;
                    const int e = av + dv - 2 * sv;
                    in = in && e >= -1 && e <= 1;
                    nz += (av != 0) + (e != 0);
                }
                if (in)
                    tab.push_back((unsigned)(a | of[d].slot << 5 | of[d].tr << 10 | d << 11 | nz << 16));
            }
    }
    off[27] = (int)tab.size();
    PA_REQUIRE(off[27] < (1 << 14), PA_ERR_LIMIT, "multigrid: Galerkin pair table too long");
    // chunks of at most T entries, T the smallest length that needs <= 64 lanes
    RapLanes rl{};
    for (int T = 1;; ++T) {
        PA_REQUIRE(T <= 255, PA_ERR_LIMIT, "multigrid: Galerkin pair lists too long for 64 lanes");
        int lanes = 0;
        for (int S = 0; S < 27; ++S) lanes += std::max(1, (off[S + 1] - off[S] + T - 1) / T);
        if (lanes > 64) continue;
        rl.nlanes = 0;
        for (int S = 0; S < 27; ++S) {
            const int c = off[S + 1] - off[S], n = std::max(1, (c + T - 1) / T);
            for (int r = 0; r < n; ++r) {
                const int t0 = off[S] + (c * r) / n, t1 = off[S] + (c * (r + 1)) / n;
                rl.lane[rl.nlanes++] = S | t0 << 5 | (t1 - t0) << 19 | (r == 0 ? 1 << 27 : 0);
            }
        }
        break;
    }
    return rl;
}

template <int D, class VT>
void setup_impl(pa_multigrid* mg, const VT* vals, hipStream_t st) {
    const pa_mesh* m = mg->mesh;
    const int L = (int)mg->grids.size();
    const Grid g0 = mg->grids[0];
    const long nq = (long)g0.nodes() * mg->fsl[0].SS;
    k_crs_to_stencil<D, VT, float><<<nblk(nq), kT, 0, st>>>(g0, mg->fsl[0], m->d_rowmap.p, m->d_colidx.p, vals, mg->fine.p);
    chk();
    k_diag_inv_fst<D><<<nblk(g0.nodes()), kT, 0, st>>>(g0.nodes(), mg->fsl[0], mg->fine.p, mg->dinv[0]->p);
    chk();
    k_mask_bits<<<nblk(g0.nodes()), kT, 0, st>>>(g0.nodes(), D, mg->mask.p, mg->mbits.p);
    chk();
    for (int l = 1; l < L; ++l) {
        const Grid gf = mg->grids[l - 1], gc = mg->grids[l];
        const int nw = (gc.nodes() + 3) / 4;  // a wave per coarse node
        if (l == 1)
            k_rap_gather<D, true><<<nw, 256, 0, st>>>(gf, gc, mg->rap_fine, mg->d_rap_fine.p, mg->fine.p,
                                                      mg->fsl[0].SS, mg->mbits.p, mg->A[1]->p);
        else
            k_rap_gather<D, false><<<nw, 256, 0, st>>>(gf, gc, mg->rap_coarse, mg->d_rap_coarse.p,
                                                       mg->A[l - 1]->p, 27, nullptr, mg->A[l]->p);
        chk();
        if (l < L - 1) {
            k_diag_inv_st<D><<<nblk(gc.nodes()), kT, 0, st>>>(gc.nodes(), mg->A[l]->p, mg->dinv[l]->p);
            chk();
        }
    }
    const Grid gl = mg->grids[L - 1];
    const long nd = (long)gl.nodes() * D;
    PA_HIP(hipMemsetAsync(mg->dense.p, 0, nd * nd * sizeof(double), st));
    k_dense_from_stencil<D><<<nblk(gl.nodes()), kT, 0, st>>>(gl, mg->A[L - 1]->p, mg->dense.p);
    chk();
    if (nd <= 256) {
        k_gauss_jordan<<<1, 1024, 0, st>>>((int)nd, mg->dense.p, mg->dense_inv.p);
        chk();
    } else {
        // one launch per pivot step over the whole GPU
        const int gb = (int)std::min<long>(1024, (2 * nd * nd + 255) / 256);
        k_gj_init<<<gb, 256, 0, st>>>((int)nd, mg->dense.p, mg->gj[0].p, mg->gj_used.p);
        chk();
        for (int k = 0; k < (int)nd; ++k) {
            k_gj_step<<<gb, 256, 0, st>>>((int)nd, k, mg->gj[k & 1].p, mg->gj[(k + 1) & 1].p, mg->gj_used.p,
                                          mg->gj_piv.p);
            chk();
        }
        k_gj_extract<<<gb, 256, 0, st>>>((int)nd, mg->gj[nd & 1].p, mg->gj_piv.p, mg->dense_inv.p);
        chk();
    }
    if (!mg->chebyshev) return;
    // Chebyshev dampings of each smoothed level from its largest eigenvalue of D^-1 A
    constexpr int K = 12;
    for (int l = 0; l < L - 1; ++l) {
        const long n = (long)mg->grids[l].nodes() * D;
        const int nb = (int)std::min<long>(1024, (n + 255) / 256);
        float* v = mg->x[l]->p;
        float* sv = mg->t[l]->p;
        k_fill_f<<<nblk(n), kT, 0, st>>>(n, mg->b[l]->p, 0.0f);
        k_pow_start<<<nblk(n), kT, 0, st>>>(n, l == 0 ? mg->mask.p : nullptr, v);
        chk();
        for (int it = 0; it < K; ++it) {
            sweep_level<D>(mg, l, mg->b[l]->p, v, sv, 1.0f, nullptr, 0, st);
            k_pow_diff<<<nb, 256, 0, st>>>(n, v, sv, mg->powpart.p + (it & 1) * 1024);
            chk();
        }
        k_cheb_dampings<<<1, 256, 0, st>>>(mg->powpart.p + ((K - 2) & 1) * 1024, mg->powpart.p + ((K - 1) & 1) * 1024,
                                           nb, level_sweeps(mg, l, true), level_sweeps(mg, l, false),
                                           mg->omegas.p + (long)l * 2 * kChebMax);
        chk();
    }
}

}  // namespace

// The operators handed to setup are symmetric positive definite (the
// elliptic problems): the fine level stores half of its stencil, and the
// default smoother is Chebyshev.
void multigrid_set_elliptic(pa_multigrid* mg) {
    mg->symmetric = true;
    // (a Chebyshev V-cycle is symmetric only when the post-smoother is the
    // pre-smoother's adjoint: the same number of sweeps)
    if (mg->opt.smoother == PA_SMOOTHER_DEFAULT && mg->opt.pre_sweeps == mg->opt.post_sweeps) mg->chebyshev = true;
}

// used by the PCG of pa_solver.hip
void multigrid_apply(pa_multigrid* mg, const double* r, double* z, hipStream_t st) {
    PA_REQUIRE(mg->ready, PA_ERR_STATE, "multigrid: setup has not been called");
    // at least two levels (pa_multigrid_create): the fine level's first and
    // last sweeps convert r and z
    if (mg->dpn == 3) vcycle<3>(mg, 0, st, r, z);
    else if (mg->dpn == 4) vcycle<4>(mg, 0, st, r, z);
    else vcycle<1>(mg, 0, st, r, z);
}

}  // namespace pa

using pa::guarded;

extern "C" {

int pa_multigrid_create(const pa_mesh* mesh, int nx, int ny, int nz, int dpn, const pa_multigrid_options* opt,
                        pa_multigrid** out) {
    return guarded([&] {
        PA_REQUIRE(mesh && opt && out, PA_ERR_ARG, "pa_multigrid_create: null argument");
        PA_REQUIRE(dpn == 1 || dpn == 3 || dpn == 4, PA_ERR_ARG, "pa_multigrid_create: dpn must be 1, 3 or 4");
        pa::Grid g0{nx, ny, nz};
        PA_REQUIRE(nx > 0 && ny > 0 && nz > 0 && g0.nodes() == mesh->N, PA_ERR_ARG,
                   "pa_multigrid_create: grid does not match the mesh node count");
        // the graph must be a vertex-neighbourhood graph of this grid
        unsigned offsets = 0;
        PA_REQUIRE(pa::grid_graph_offsets(mesh, g0, &offsets), PA_ERR_ARG,
                   "pa_multigrid_create: mesh graph is not a structured-grid neighbourhood graph");
        auto* mg = new pa_multigrid();
        try {
            mg->mesh = mesh;
            mg->graph_offsets = offsets;
            mg->symmetric = opt->symmetric != 0;
            PA_REQUIRE(opt->smoother >= PA_SMOOTHER_DEFAULT && opt->smoother <= PA_SMOOTHER_CHEBYSHEV, PA_ERR_ARG,
                       "pa_multigrid_create: unknown smoother");
            PA_REQUIRE(opt->coarse_sweeps >= 0 && opt->coarse_max_nodes >= 0, PA_ERR_ARG,
                       "pa_multigrid_create: negative coarse_sweeps or coarse_max_nodes");
            mg->chebyshev = opt->smoother == PA_SMOOTHER_CHEBYSHEV;
            PA_REQUIRE(!mg->chebyshev || opt->pre_sweeps == opt->post_sweeps, PA_ERR_ARG,
                       "pa_multigrid_create: Chebyshev smoothing needs as many post- as pre-smoothing sweeps");
            mg->dpn = dpn;
            mg->opt = *opt;
            if (mg->opt.omega <= 0.0) mg->opt.omega = 0.6;
            mg->grids.push_back(g0);
            const int maxc = opt->max_coarse_nodes > 0 ? opt->max_coarse_nodes : 2000;
            while (true) {
                const pa::Grid& g = mg->grids.back();
                if (g.nx % 2 || g.ny % 2 || g.nz % 2) break;
                if (opt->max_levels > 0 && (int)mg->grids.size() >= opt->max_levels) break;
                mg->grids.push_back(pa::Grid{g.nx / 2, g.ny / 2, g.nz / 2});
            }
            PA_REQUIRE(mg->grids.size() >= 2, PA_ERR_ARG, "pa_multigrid_create: the grid cannot be coarsened (odd cell counts)");
            PA_REQUIRE(mg->grids.back().nodes() <= maxc, PA_ERR_LIMIT,
                       "pa_multigrid_create: coarsest level too large for the dense solve");
            const int L = (int)mg->grids.size();
            const int bs = dpn * dpn;
            for (int l = 0; l < L; ++l) {
                const long nn = mg->grids[l].nodes();
                mg->A.push_back(new pa::DevBuf<float>());
                mg->dinv.push_back(new pa::DevBuf<float>());
                mg->b.push_back(new pa::DevBuf<float>());
                mg->x.push_back(new pa::DevBuf<float>());
                mg->t.push_back(new pa::DevBuf<float>());
                if (l > 0) mg->A[l]->alloc(nn * 27 * bs);
                mg->dinv[l]->alloc(nn * bs);
                mg->b[l]->alloc(nn * dpn);
                mg->x[l]->alloc(nn * dpn);
                mg->t[l]->alloc(nn * dpn);
            }
            const long nd = (long)mg->grids.back().nodes() * dpn;
            mg->dense.alloc(nd * nd);
            mg->dense_inv.alloc(nd * nd);
            if (nd > 256) {
                mg->gj[0].alloc(2 * nd * nd);
                mg->gj[1].alloc(2 * nd * nd);
                mg->gj_used.alloc(nd);
                mg->gj_piv.alloc(nd);
            }
            mg->mask.alloc((long)mesh->N * dpn);
            mg->omegas.alloc((long)L * 2 * pa::kChebMax);
            mg->powpart.alloc(2 * 1024);
            mg->mbits.alloc(mesh->N);
        } catch (...) {
            delete mg;
            throw;
        }
        *out = mg;
    });
}

int pa_multigrid_destroy(pa_multigrid* mg) {
    return guarded([&] { delete mg; });
}

int pa_multigrid_levels(const pa_multigrid* mg, int* levels, int* coarse_dofs) {
    return guarded([&] {
        PA_REQUIRE(mg, PA_ERR_ARG, "pa_multigrid_levels: null argument");
        if (levels) *levels = (int)mg->grids.size();
        if (coarse_dofs) *coarse_dofs = mg->grids.back().nodes() * mg->dpn;
    });
}

int pa_multigrid_setup(pa_multigrid* mg, const void* vals, int precision, int nbc, const int* bc_dofs, void* stream) {
    return guarded([&] {
        PA_REQUIRE(mg && vals, PA_ERR_ARG, "pa_multigrid_setup: null argument");
        PA_REQUIRE(precision == PA_FP64 || precision == PA_FP32, PA_ERR_ARG, "pa_multigrid_setup: unknown precision");
        const hipStream_t st = (hipStream_t)stream;
        const long n = (long)mg->mesh->N * mg->dpn;
        pa::k_mask_init<<<pa::nblk(n), pa::kT, 0, st>>>(n, mg->mask.p);
        pa::chk();
        if (nbc > 0) {
            PA_REQUIRE(bc_dofs, PA_ERR_ARG, "pa_multigrid_setup: null constraint list");
            pa::k_mask_bc<<<pa::nblk(nbc), pa::kT, 0, st>>>(nbc, bc_dofs, mg->mask.p);
            pa::chk();
        }
        PA_REQUIRE(!mg->chebyshev || (mg->opt.pre_sweeps <= pa::kChebMax && mg->opt.post_sweeps <= pa::kChebMax &&
                                      mg->opt.coarse_sweeps <= pa::kChebMax),
                   PA_ERR_ARG, "pa_multigrid_setup: at most 8 Chebyshev sweeps");
        // the smoother reads an FP32 stencil copy of the fine matrix
        mg->fsl.assign(1, pa::make_block_stencil(mg->grids[0], mg->graph_offsets, mg->symmetric));
        for (size_t l = 1; l < mg->grids.size(); ++l)
            mg->fsl.push_back(pa::make_coarse_stencil(mg->grids[l], mg->symmetric));
        if (mg->rap_half != mg->fsl[0].half) {
            // the pair tables follow the fine stencil's layout (built once)
            std::vector<unsigned> tab;
            mg->rap_fine = pa::make_rap_table(mg->fsl[0], tab);
            mg->d_rap_fine.upload(tab.data(), tab.size());
            mg->rap_coarse = pa::make_rap_table(pa::make_coarse_stencil(mg->grids[0], false), tab);
            mg->d_rap_coarse.upload(tab.data(), tab.size());
            mg->rap_half = mg->fsl[0].half;
        }
        const long nf = (long)mg->mesh->N * mg->fsl[0].SS * mg->dpn * mg->dpn;
        if (mg->fine.n != (size_t)nf) mg->fine.alloc(nf);
        auto run = [&](auto tag) {
            using VT = decltype(tag);
            const VT* v = (const VT*)vals;
            if (mg->dpn == 3) pa::setup_impl<3, VT>(mg, v, st);
            else if (mg->dpn == 4) pa::setup_impl<4, VT>(mg, v, st);
            else pa::setup_impl<1, VT>(mg, v, st);
        };
        if (precision == PA_FP64) run(0.0);
        else run(0.0f);
        mg->ready = true;
    });
}

int pa_multigrid_level_operator(const pa_multigrid* mg, int level, float* out, int* nx, int* ny, int* nz) {
    return guarded([&] {
        PA_REQUIRE(mg, PA_ERR_ARG, "pa_multigrid_level_operator: null argument");
        PA_REQUIRE(level >= 1 && level < (int)mg->grids.size(), PA_ERR_ARG, "pa_multigrid_level_operator: bad level");
        const pa::Grid g = mg->grids[level];
        if (nx) *nx = g.nx;
        if (ny) *ny = g.ny;
        if (nz) *nz = g.nz;
        if (out) {
            PA_REQUIRE(mg->ready, PA_ERR_STATE, "pa_multigrid_level_operator: setup has not been called");
            PA_HIP(hipDeviceSynchronize());
            PA_HIP(hipMemcpy(out, mg->A[level]->p, mg->A[level]->n * sizeof(float), hipMemcpyDeviceToHost));
        }
    });
}

int pa_multigrid_apply(pa_multigrid* mg, const double* r, double* z, void* stream) {
    return guarded([&] {
        PA_REQUIRE(mg && r && z, PA_ERR_ARG, "pa_multigrid_apply: null argument");
        pa::multigrid_apply(mg, r, z, (hipStream_t)stream);
    });
}

}  // extern "C"
