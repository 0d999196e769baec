// Preconditioned CG on the block-CRS matrices of the mesh graph, and the
// elliptic problem driver (solution / criterion value / criterion gradient).
//
// PCG keeps every scalar on the device: per iteration one SpMV fused with the
// p.q partial sums, one update kernel fused with the z = M r application and
// the r.z / r.r partial sums, one direction update, and two fixed-order
// reductions (deterministic results).  The host reads ||r||^2 only every
// `check_every` iterations.
#include <hip/hip_runtime.h>

#include <cmath>
#include <functional>
#include <vector>

#include "pa_cell.h"
#include "pa_internal.h"
#include "pa_stencil.h"

struct pa_solver {
    const pa_mesh* mesh = nullptr;
    int dpn = 3;
    int precond = PA_PRECOND_BLOCK_JACOBI;
    int nvb = 0;  // blocks of the vector kernels
    int nsb = 0;  // blocks of the SpMV kernel
    pa::DevBuf<int> d_diag;       // N: block index of the diagonal
    pa::DevBuf<double> d_minv;    // N*dpn*dpn (block Jacobi) or N*dpn (Jacobi)
    pa::DevBuf<double> r, zv, p, q;
    pa::DevBuf<double> part;      // partial sums: 4 x max(nvb, nsb)
    pa::DevBuf<double> scal;      // device scalars
    std::vector<double> h_scal;
    pa_multigrid* mg = nullptr;   // PA_PRECOND_MULTIGRID (not owned)
    // structured grid meshes: the matrix of a solve is copied once into a
    // block stencil that the SpMV reads (half of it for symmetric matrices)
    bool stencil = false;
    bool symmetric = false;
    pa::Grid sgrid{};
    unsigned soffsets = 0;
    pa::BlockStencil bs{};
    pa::DevBuf<double> st64;
    pa::DevBuf<float> st32;
    bool st_ready = false;  // the stencil copy of the current solve is built
    int nsp = 0;            // partial sums written by the last SpMV
};

struct pa_problem {
    pa_vector_function* vf = nullptr;
    pa_scalar_function* sf = nullptr;
    pa_problem_options opt{};
    pa_solver* solver = nullptr;
    int nbc = 0;
    pa::DevBuf<int> bc_dofs;
    pa::DevBuf<double> bc_vals;
    pa::DevBuf<double> bc_dense;        // constraint work vectors (N * dpn, N)
    pa::DevBuf<unsigned char> bc_touch;
    pa::DevBuf<double> K64;
    pa::DevBuf<float> K32;
    pa::DevBuf<double> rhs, res, dx, zero;
    pa_multigrid* mg = nullptr;  // owned
};

namespace pa {
namespace {

constexpr int kB = 256;

enum { S_RZ = 0, S_PQ = 1, S_RR = 2, S_BB = 3, S_RZOLD = 4, S_N = 8 };

template <int G>
__device__ inline double gsum(double v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}

// Block-wide sum of one value per thread; thread 0 gets the result.
__device__ inline double block_sum(double v, double* sh) {
    v = gsum<64>(v);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / 64); ++i) t += sh[i];
    __syncthreads();
    return t;
}

// y = A p (16 lanes per block row) and partial sums of p.y per thread block.
template <int D, class VT>
__global__ __launch_bounds__(kB) void k_spmv_dot(int N, const int* __restrict__ rowmap, const int* __restrict__ colidx,
                                                 const VT* __restrict__ vals, const double* __restrict__ x,
                                                 double* __restrict__ y, double* __restrict__ part) {
    constexpr int G = 16, ROWS = kB / G;
    __shared__ double sh[kB / 64];
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    const int row = blockIdx.x * ROWS + grp;
    double acc[D];
#pragma unroll
    for (int i = 0; i < D; ++i) acc[i] = 0.0;
    if (row < N) {
        for (int e = rowmap[row] + lane; e < rowmap[row + 1]; e += G) {
            const long c = colidx[e];
            double xv[D];
#pragma unroll
            for (int j = 0; j < D; ++j) xv[j] = x[c * D + j];
            const VT* v = vals + (long)e * D * D;
#pragma unroll
            for (int i = 0; i < D; ++i)
#pragma unroll
                for (int j = 0; j < D; ++j) acc[i] += (double)v[i * D + j] * xv[j];
        }
    }
    double d = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        acc[i] = gsum<G>(acc[i]);
        if (row < N && lane == 0) {
            y[(long)row * D + i] = acc[i];
            d += acc[i] * x[(long)row * D + i];
        }
    }
    const double t = block_sum(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// z = M r for one node (block Jacobi, Jacobi or identity).
template <int D>
__device__ inline void apply_prec(int prec, const double* __restrict__ minv, long node, const double* r, double* z) {
    if (prec == PA_PRECOND_BLOCK_JACOBI) {
        const double* m = minv + node * D * D;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) s += m[i * D + j] * r[j];
            z[i] = s;
        }
    } else if (prec == PA_PRECOND_JACOBI) {
#pragma unroll
        for (int i = 0; i < D; ++i) z[i] = minv[node * D + i] * r[i];
    } else {
#pragma unroll
        for (int i = 0; i < D; ++i) z[i] = r[i];
    }
}

// r = b - q; z = M r; p = z; partials of r.z, r.r, b.b.
template <int D>
__global__ __launch_bounds__(kB) void k_pcg_init(int N, int prec, const double* __restrict__ minv,
                                                 const double* __restrict__ b, const double* __restrict__ q,
                                                 double* __restrict__ r, double* __restrict__ zv,
                                                 double* __restrict__ p, double* __restrict__ part, int nparts) {
    __shared__ double sh[kB / 64];
    double rz = 0.0, rr = 0.0, bb = 0.0;
    for (long node = (long)blockIdx.x * kB + threadIdx.x; node < N; node += (long)gridDim.x * kB) {
        double rv[D], zz[D];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const double bv = b[node * D + i];
            rv[i] = bv - q[node * D + i];
            bb += bv * bv;
        }
        apply_prec<D>(prec, minv, node, rv, zz);
#pragma unroll
        for (int i = 0; i < D; ++i) {
            r[node * D + i] = rv[i];
            zv[node * D + i] = zz[i];
            p[node * D + i] = zz[i];
            rz += rv[i] * zz[i];
            rr += rv[i] * rv[i];
        }
    }
    const double a = block_sum(rz, sh), c = block_sum(rr, sh), e = block_sum(bb, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[nparts + blockIdx.x] = c;
        part[2 * nparts + blockIdx.x] = e;
    }
}

// Fixed-order sums of `count` partial arrays into scal[dst[k]].
__global__ void k_reduce(const double* __restrict__ part, int n, int nparts, int count, int d0, int d1, int d2,
                         int shift_rz, double* __restrict__ scal) {
    __shared__ double sh[1024 / 64];
    const int dst[3] = {d0, d1, d2};
    for (int k = 0; k < count; ++k) {
        double s = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[(long)k * nparts + i];
        const double t = block_sum(s, sh);
        if (threadIdx.x == 0) {
            if (shift_rz && dst[k] == S_RZ) scal[S_RZOLD] = scal[S_RZ];
            scal[dst[k]] = t;
        }
        __syncthreads();
    }
}

// x += a p; r -= a q; z = M r; partials of r.z and r.r (a = rz / pq).
// store_z = 0: z is not written (a V-cycle computes it next).
template <int D>
__global__ __launch_bounds__(kB) void k_pcg_update(int N, int prec, const double* __restrict__ minv,
                                                   const double* __restrict__ scal, const double* __restrict__ p,
                                                   const double* __restrict__ q, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ zv, int store_z,
                                                   double* __restrict__ part, int nparts) {
    __shared__ double sh[kB / 64];
    const double pq = scal[S_PQ];
    const double alpha = pq != 0.0 ? scal[S_RZ] / pq : 0.0;
    double rz = 0.0, rr = 0.0;
    for (long node = (long)blockIdx.x * kB + threadIdx.x; node < N; node += (long)gridDim.x * kB) {
        double rv[D], zz[D];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const long o = node * D + i;
            x[o] += alpha * p[o];
            rv[i] = r[o] - alpha * q[o];
        }
        apply_prec<D>(prec, minv, node, rv, zz);
#pragma unroll
        for (int i = 0; i < D; ++i) {
            r[node * D + i] = rv[i];
            if (store_z) zv[node * D + i] = zz[i];
            rz += rv[i] * zz[i];
            rr += rv[i] * rv[i];
        }
    }
    const double a = block_sum(rz, sh), c = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[nparts + blockIdx.x] = c;
    }
}

// p = z + beta p, beta = rz / rz_old.
__global__ __launch_bounds__(kB) void k_pcg_dir(long n, const double* __restrict__ scal, const double* __restrict__ zv,
                                                double* __restrict__ p) {
    const double rzo = scal[S_RZOLD];
    const double beta = rzo != 0.0 ? scal[S_RZ] / rzo : 0.0;
    for (long i = (long)blockIdx.x * kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) p[i] = zv[i] + beta * p[i];
}

// Inverse of the diagonal blocks (block Jacobi) or of the diagonal entries.
template <int D, class VT>
__global__ void k_prec_setup(int N, int prec, const int* __restrict__ diag, const VT* __restrict__ vals,
                             double* __restrict__ minv) {
    const long node = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= N) return;
    const VT* a = vals + (long)diag[node] * D * D;
    if (prec == PA_PRECOND_BLOCK_JACOBI && D == 3) {
        const double a00 = a[0], a01 = a[1], a02 = a[2], a10 = a[3], a11 = a[4], a12 = a[5], a20 = a[6], a21 = a[7],
                     a22 = a[8];
        const double c00 = a11 * a22 - a12 * a21, c01 = a02 * a21 - a01 * a22, c02 = a01 * a12 - a02 * a11;
        const double c10 = a12 * a20 - a10 * a22, c11 = a00 * a22 - a02 * a20, c12 = a02 * a10 - a00 * a12;
        const double c20 = a10 * a21 - a11 * a20, c21 = a01 * a20 - a00 * a21, c22 = a00 * a11 - a01 * a10;
        const double det = a00 * c00 + a01 * c10 + a02 * c20;
        const double id = det != 0.0 ? 1.0 / det : 0.0;
        double* m = minv + node * 9;
        m[0] = c00 * id; m[1] = c01 * id; m[2] = c02 * id;
        m[3] = c10 * id; m[4] = c11 * id; m[5] = c12 * id;
        m[6] = c20 * id; m[7] = c21 * id; m[8] = c22 * id;
    } else if (prec == PA_PRECOND_BLOCK_JACOBI) {
        const double d = a[0];
        minv[node] = d != 0.0 ? 1.0 / d : 0.0;
    } else {
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const double d = a[i * D + i];
            minv[node * D + i] = d != 0.0 ? 1.0 / d : 0.0;
        }
    }
}

// Mixed-mode residual fix-up: r = -R(x) on free dofs, g - x on constrained ones.
__global__ void k_neg(long n, double* __restrict__ v) {
    for (long i = (long)blockIdx.x * kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) v[i] = -v[i];
}
__global__ void k_bc_residual(int nbc, const int* __restrict__ dofs, const double* __restrict__ g,
                              const double* __restrict__ x, double* __restrict__ r) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nbc) r[dofs[b]] = g[b] - x[dofs[b]];
}
__global__ void k_axpy(long n, double a, const double* __restrict__ x, double* __restrict__ y) {
    for (long i = (long)blockIdx.x * kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) y[i] += a * x[i];
}
__global__ __launch_bounds__(kB) void k_dot_part(long n, const double* __restrict__ a, const double* __restrict__ b,
                                                 double* __restrict__ part) {
    __shared__ double sh[kB / 64];
    double s = 0.0;
    for (long i = (long)blockIdx.x * kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) s += a[i] * b[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

void check() { PA_HIP(hipGetLastError()); }

// y = A x and partial sums of x.y per thread block on the block stencil: G
// lanes per node, NS offsets per lane, FP64 accumulation; lane l < D writes
// dof l.
template <int D, class VT, int G, int NS, bool SYM>
__global__ __launch_bounds__(kB) void k_st_spmv_dot(Grid g, BlockStencil fs, const VT* __restrict__ A,
                                                    const double* __restrict__ x, double* __restrict__ y,
                                                    double* __restrict__ part) {
    constexpr int ROWS = kB / G, BS = D * D;
    __shared__ double sh[kB / 64];
    const int l = threadIdx.x & (G - 1);
    const int n = xcd_block() * ROWS + threadIdx.x / G;
    double d = 0.0;
    if (n < g.nodes()) {
        int i, j, k;
        stencil_ijk(g, fs, n, i, j, k);
        const int ld = l < D ? l : D - 1;
        const double xl = x[(long)n * D + ld];
        VT blk[NS][BS];
        double xc[NS][D];
        bool ok[NS], tr[NS];
#pragma unroll
        for (int m = 0; m < NS; ++m) {
            long c, base;
            stencil_slot<G, SYM>(g, fs, m, l, n, i, j, k, BS, ok[m], tr[m], c, base);
#pragma unroll
            for (int q = 0; q < BS; ++q) blk[m][q] = A[base + q];
#pragma unroll
            for (int q = 0; q < D; ++q) xc[m][q] = x[c * D + q];
        }
        double acc[D];
#pragma unroll
        for (int p = 0; p < D; ++p) {
            acc[p] = 0.0;
#pragma unroll
            for (int m = 0; m < NS; ++m) {
                double yn = 0.0, yt = 0.0;
#pragma unroll
                for (int q = 0; q < D; ++q) {
                    yn = fma((double)blk[m][p * D + q], xc[m][q], yn);
                    if constexpr (SYM) yt = fma((double)blk[m][q * D + p], xc[m][q], yt);
                }
                acc[p] += ok[m] ? ((SYM && tr[m]) ? yt : yn) : 0.0;
            }
            acc[p] = node_lanes_sum<G>(acc[p]);
        }
        double al = acc[0];
#pragma unroll
        for (int p = 1; p < D; ++p) al = ld == p ? acc[p] : al;
        if (l < D) {
            y[(long)n * D + l] = al;
            d = al * xl;
        }
    }
    const double t = block_sum(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// The stencil copy of the solve's matrix (once per solve).
template <int D, class VT>
void stencil_copy(pa_solver* s, const VT* vals, hipStream_t st) {
    const long nq = (long)s->mesh->N * s->bs.SS;
    VT* A;
    if constexpr (sizeof(VT) == 8) {
        if (s->st64.n != (size_t)(nq * D * D)) s->st64.alloc(nq * D * D);
        A = s->st64.p;
    } else {
        if (s->st32.n != (size_t)(nq * D * D)) s->st32.alloc(nq * D * D);
        A = s->st32.p;
    }
    k_crs_to_stencil<D, VT, VT><<<(int)((nq + kB - 1) / kB), kB, 0, st>>>(s->sgrid, s->bs, s->mesh->d_rowmap.p,
                                                                         s->mesh->d_colidx.p, vals, A);
    check();
    s->st_ready = true;
}

template <int D, class VT>
void spmv_dot(pa_solver* s, const VT* vals, const double* x, double* y, hipStream_t st) {
    const pa_mesh* m = s->mesh;
    if (s->st_ready) {
        const VT* A;
        if constexpr (sizeof(VT) == 8) A = s->st64.p;
        else A = s->st32.p;
        const int nb = (m->N + kB / 8 - 1) / (kB / 8);
        auto run = [&](auto kernel) { kernel<<<nb, kB, 0, st>>>(s->sgrid, s->bs, A, x, y, s->part.p); };
        if (s->bs.S <= 16) {
            if (s->bs.half) run(k_st_spmv_dot<D, VT, 8, 2, true>);
            else run(k_st_spmv_dot<D, VT, 8, 2, false>);
        } else {
            if (s->bs.half) run(k_st_spmv_dot<D, VT, 8, 4, true>);
            else run(k_st_spmv_dot<D, VT, 8, 4, false>);
        }
        s->nsp = nb;
    } else {
        k_spmv_dot<D, VT><<<s->nsb, kB, 0, st>>>(m->N, m->d_rowmap.p, m->d_colidx.p, vals, x, y, s->part.p);
        s->nsp = s->nsb;
    }
    check();
}

double read_scalar(pa_solver* s, int idx, hipStream_t st) {
    double v = 0.0;
    PA_HIP(hipMemcpyAsync(&v, s->scal.p + idx, sizeof(double), hipMemcpyDeviceToHost, st));
    PA_HIP(hipStreamSynchronize(st));
    return v;
}

// ||v||^2 with the fixed-order reduction.
double norm2(pa_solver* s, const double* v, hipStream_t st) {
    const long n = (long)s->mesh->N * s->dpn;
    k_dot_part<<<s->nvb, kB, 0, st>>>(n, v, v, s->part.p);
    check();
    k_reduce<<<1, 1024, 0, st>>>(s->part.p, s->nvb, s->nvb, 1, S_N - 1, 0, 0, 0, s->scal.p);
    check();
    return read_scalar(s, S_N - 1, st);
}

// z = M r and the r.z, r.r partial sums after the residual was replaced.
template <int D>
__global__ __launch_bounds__(kB) void k_pcg_reprec(int N, int prec, const double* __restrict__ minv,
                                                   const double* __restrict__ r, double* __restrict__ zv,
                                                   double* __restrict__ part, int nparts) {
    __shared__ double sh[kB / 64];
    double rz = 0.0, rr = 0.0;
    for (long node = (long)blockIdx.x * kB + threadIdx.x; node < N; node += (long)gridDim.x * kB) {
        double rv[D], zz[D];
#pragma unroll
        for (int i = 0; i < D; ++i) rv[i] = r[node * D + i];
        apply_prec<D>(prec, minv, node, rv, zz);
#pragma unroll
        for (int i = 0; i < D; ++i) {
            zv[node * D + i] = zz[i];
            rz += rv[i] * zz[i];
            rr += rv[i] * rv[i];
        }
    }
    const double a = block_sum(rz, sh), c = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[nparts + blockIdx.x] = c;
    }
}

// Exact residual supplier for reliable updates: writes b - A_exact x into r.
using ExactResidual = std::function<void(const double* x, double* r)>;

template <int D, class VT>
void pcg_impl(pa_solver* s, const VT* vals, const double* b, double* x, double tol, int max_iter, int check_every,
              pa_solve_stats* stats, hipStream_t st, const ExactResidual* exact = nullptr, int replace_every = 0) {
    const pa_mesh* m = s->mesh;
    const int N = m->N;
    const long n = (long)N * D;
    const int nvb = s->nvb;
    // multigrid: the fused kernels run unpreconditioned (z = r), then z = M r
    // by a V-cycle and r.z are recomputed
    const bool mgp = s->precond == PA_PRECOND_MULTIGRID;
    PA_REQUIRE(!mgp || s->mg, PA_ERR_STATE, "pcg: multigrid preconditioner selected but not set");
    const int kprec = mgp ? PA_PRECOND_NONE : s->precond;
    auto mg_rz = [&]() {
        multigrid_apply(s->mg, s->r.p, s->zv.p, st);
        k_dot_part<<<nvb, kB, 0, st>>>(n, s->r.p, s->zv.p, s->part.p);
        check();
        k_reduce<<<1, 1024, 0, st>>>(s->part.p, nvb, nvb, 1, S_RZ, 0, 0, 0, s->scal.p);
        check();
    };
    if (kprec != PA_PRECOND_NONE) {
        k_prec_setup<D, VT><<<(N + 255) / 256, 256, 0, st>>>(N, kprec, s->d_diag.p, vals, s->d_minv.p);
        check();
    }
    s->st_ready = false;
    if (s->stencil) stencil_copy<D, VT>(s, vals, st);
    spmv_dot<D, VT>(s, vals, x, s->q.p, st);
    k_pcg_init<D><<<nvb, kB, 0, st>>>(N, kprec, s->d_minv.p, b, s->q.p, s->r.p, s->zv.p, s->p.p, s->part.p, nvb);
    check();
    k_reduce<<<1, 1024, 0, st>>>(s->part.p, nvb, nvb, 3, S_RZ, S_RR, S_BB, 0, s->scal.p);
    check();
    if (mgp) {
        mg_rz();
        PA_HIP(hipMemcpyAsync(s->p.p, s->zv.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    const double bb = read_scalar(s, S_BB, st);
    const double bnorm = std::sqrt(bb);
    int it = 0;
    double rr = read_scalar(s, S_RR, st);
    stats->sweeps = 1;
    const double target = tol * bnorm;
    check_every = check_every > 0 ? check_every : 10;
    while (it < max_iter && std::sqrt(rr) > target && bnorm > 0.0) {
        const int burst = std::min(check_every, max_iter - it);
        for (int k = 0; k < burst; ++k) {
            spmv_dot<D, VT>(s, vals, s->p.p, s->q.p, st);
            k_reduce<<<1, 1024, 0, st>>>(s->part.p, s->nsp, s->nsp, 1, S_PQ, 0, 0, 0, s->scal.p);
            check();
            k_pcg_update<D><<<nvb, kB, 0, st>>>(N, kprec, s->d_minv.p, s->scal.p, s->p.p, s->q.p, x, s->r.p,
                                                s->zv.p, mgp ? 0 : 1, s->part.p, nvb);
            check();
            k_reduce<<<1, 1024, 0, st>>>(s->part.p, nvb, nvb, 2, S_RZ, S_RR, 0, 1, s->scal.p);
            check();
            if (mgp) mg_rz();
            k_pcg_dir<<<nvb, kB, 0, st>>>(n, s->scal.p, s->zv.p, s->p.p);
            check();
        }
        it += burst;
        rr = read_scalar(s, S_RR, st);
        if (!std::isfinite(rr)) break;
        const bool converged = std::sqrt(rr) <= target;
        if (exact && (converged || (replace_every > 0 && it % replace_every == 0))) {
            // residual replacement: swap the recursive residual for the exact
            // one and restart the directions from it (p = M r), since the
            // recursive residual of the FP32 operator may have drifted far
            // below the exact one
            (*exact)(x, s->r.p);
            k_pcg_reprec<D><<<nvb, kB, 0, st>>>(N, kprec, s->d_minv.p, s->r.p, s->zv.p, s->part.p, nvb);
            check();
            k_reduce<<<1, 1024, 0, st>>>(s->part.p, nvb, nvb, 2, S_RZ, S_RR, 0, 0, s->scal.p);
            check();
            if (mgp) mg_rz();
            PA_HIP(hipMemcpyAsync(s->p.p, s->zv.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            rr = read_scalar(s, S_RR, st);
            ++stats->sweeps;
        }
    }
    double rn;
    if (exact) {
        (*exact)(x, s->q.p);
        rn = std::sqrt(norm2(s, s->q.p, st));
    } else {
        // FP64 residual of the matrix in use
        spmv_dot<D, VT>(s, vals, x, s->q.p, st);
        k_axpy<<<nvb, kB, 0, st>>>(n, -1.0, b, s->q.p);  // q = A x - b
        check();
        rn = std::sqrt(norm2(s, s->q.p, st));
    }
    stats->iterations = it;
    stats->rel_residual = bnorm > 0.0 ? rn / bnorm : 0.0;
    stats->converged = stats->rel_residual <= tol ? 1 : 0;
}

// The matrices this solver is given are symmetric (the elliptic problems:
// assembled blocks (m, n) and (n, m) are transposes bit for bit, and stay so
// under the symmetric Dirichlet elimination): the SpMV reads half a stencil.
void solver_set_symmetric(pa_solver* s) {
    s->symmetric = true;
    if (s->stencil) s->bs = make_block_stencil(s->sgrid, s->soffsets, true, 8);
}

void pcg(pa_solver* s, const void* vals, int precision, const double* b, double* x, double tol, int max_iter,
         int check_every, pa_solve_stats* stats, hipStream_t st) {
    PA_REQUIRE(precision == PA_FP64 || precision == PA_FP32, PA_ERR_ARG, "pcg: unknown precision");
    if (s->dpn == 3) {
        if (precision == PA_FP64) pcg_impl<3, double>(s, (const double*)vals, b, x, tol, max_iter, check_every, stats, st);
        else pcg_impl<3, float>(s, (const float*)vals, b, x, tol, max_iter, check_every, stats, st);
    } else {
        if (precision == PA_FP64) pcg_impl<1, double>(s, (const double*)vals, b, x, tol, max_iter, check_every, stats, st);
        else pcg_impl<1, float>(s, (const float*)vals, b, x, tol, max_iter, check_every, stats, st);
    }
}

void check_status(int st) {
    if (st != PA_OK) throw Error(st, pa_last_error());
}

}  // namespace
}  // namespace pa

using pa::guarded;

extern "C" {

int pa_solver_create(const pa_mesh* mesh, int dpn, int precond, pa_solver** out) {
    return guarded([&] {
        PA_REQUIRE(mesh && out, PA_ERR_ARG, "pa_solver_create: null argument");
        PA_REQUIRE(dpn == 1 || dpn == 3, PA_ERR_ARG, "pa_solver_create: dpn must be 1 or 3");
        PA_REQUIRE(precond >= PA_PRECOND_NONE && precond <= PA_PRECOND_MULTIGRID, PA_ERR_ARG,
                   "pa_solver_create: unknown preconditioner");
        auto* s = new pa_solver();
        try {
            s->mesh = mesh;
            s->dpn = dpn;
            s->precond = precond;
            const int N = mesh->N;
            std::vector<int> diag(N);
            for (int i = 0; i < N; ++i) {
                diag[i] = -1;
                for (int e = mesh->h_rowmap[i]; e < mesh->h_rowmap[i + 1]; ++e)
                    if (mesh->h_colidx[e] == i) diag[i] = e;
                PA_REQUIRE(diag[i] >= 0, PA_ERR_ARG, "pa_solver_create: graph row without diagonal");
            }
            s->d_diag.upload(diag.data(), diag.size());
            const long n = (long)N * dpn;
            s->nvb = (int)std::min<long>(1024, (N + pa::kB - 1) / pa::kB);
            s->nsb = (N + pa::kB / 16 - 1) / (pa::kB / 16);
            s->d_minv.alloc(precond == PA_PRECOND_BLOCK_JACOBI ? (long)N * dpn * dpn : n);
            s->r.alloc(n);
            s->zv.alloc(n);
            s->p.alloc(n);
            s->q.alloc(n);
            s->part.alloc(4L * std::max(s->nvb, s->nsb));
            // a lexicographic node grid whose graph is a neighbourhood graph
            const int* gd = mesh->grid_dims[0] > 0 ? mesh->grid_dims : (mesh->has_grid ? mesh->kgrid : nullptr);
            if (gd) {
                s->sgrid = pa::Grid{gd[0] - 1, gd[1] - 1, gd[2] - 1};
                s->stencil = gd[0] > 1 && gd[1] > 1 && gd[2] > 1 && pa::grid_graph_offsets(mesh, s->sgrid, &s->soffsets);
                if (s->stencil) s->bs = pa::make_block_stencil(s->sgrid, s->soffsets, false, 8);
            }
            s->scal.alloc(pa::S_N);
            PA_HIP(hipMemset(s->scal.p, 0, pa::S_N * sizeof(double)));
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int pa_solver_destroy(pa_solver* s) {
    return guarded([&] { delete s; });
}

int pa_solver_set_multigrid(pa_solver* s, pa_multigrid* mg) {
    return guarded([&] {
        PA_REQUIRE(s && mg, PA_ERR_ARG, "pa_solver_set_multigrid: null argument");
        s->mg = mg;
        s->precond = PA_PRECOND_MULTIGRID;
    });
}

int pa_solver_pcg(pa_solver* s, const void* vals, int precision, const double* b, double* x, double tol,
                  int max_iter, int check_every, pa_solve_stats* stats, void* stream) {
    return guarded([&] {
        PA_REQUIRE(s && vals && b && x && stats, PA_ERR_ARG, "pa_solver_pcg: null argument");
        pa::pcg(s, vals, precision, b, x, tol, max_iter, check_every, stats, (hipStream_t)stream);
    });
}

int pa_problem_create(pa_vector_function* vf, pa_scalar_function* sf, int nbc, const int* dofs, const double* vals,
                      const pa_problem_options* opt, pa_problem** out) {
    return guarded([&] {
        PA_REQUIRE(vf && sf && opt && out, PA_ERR_ARG, "pa_problem_create: null argument");
        PA_REQUIRE(sf->mesh == vf->mesh && sf->dpn == vf->dpn, PA_ERR_ARG,
                   "pa_problem_create: criterion and PDE differ in mesh or physics");
        PA_REQUIRE(nbc >= 0 && (nbc == 0 || (dofs && vals)), PA_ERR_ARG, "pa_problem_create: bad constraints");
        const long n = (long)vf->mesh->N * vf->dpn;
        for (int i = 0; i < nbc; ++i)
            PA_REQUIRE(dofs[i] >= 0 && dofs[i] < n, PA_ERR_ARG, "pa_problem_create: constrained dof out of range");
        auto* P = new pa_problem();
        try {
            P->vf = vf;
            P->sf = sf;
            P->opt = *opt;
            P->nbc = nbc;
            P->bc_dofs.upload(dofs, nbc);
            P->bc_vals.upload(vals, nbc);
            if (nbc) {
                P->bc_dense.alloc(n);
                P->bc_touch.alloc(vf->mesh->N);
            }
            pa::check_status(pa_solver_create(vf->mesh, vf->dpn, opt->precond, &P->solver));
            pa::solver_set_symmetric(P->solver);
            P->rhs.alloc(n);
            P->res.alloc(n);
            P->dx.alloc(n);
        } catch (...) {
            if (P->solver) pa_solver_destroy(P->solver);
            delete P;
            throw;
        }
        *out = P;
    });
}

int pa_problem_destroy(pa_problem* P) {
    return guarded([&] {
        if (!P) return;
        if (P->solver) pa_solver_destroy(P->solver);
        if (P->mg) pa_multigrid_destroy(P->mg);
        delete P;
    });
}

int pa_problem_set_options(pa_problem* P, const pa_problem_options* opt) {
    return guarded([&] {
        PA_REQUIRE(P && opt, PA_ERR_ARG, "pa_problem_set_options: null argument");
        PA_REQUIRE(opt->precond != PA_PRECOND_MULTIGRID || P->mg, PA_ERR_ARG,
                   "pa_problem_set_options: multigrid needs pa_problem_set_multigrid");
        if (opt->precond != P->opt.precond) {
            pa_solver* s = nullptr;
            pa::check_status(pa_solver_create(P->vf->mesh, P->vf->dpn, opt->precond, &s));
            pa::solver_set_symmetric(s);
            if (opt->precond == PA_PRECOND_MULTIGRID) s->mg = P->mg;
            pa_solver_destroy(P->solver);
            P->solver = s;
        }
        P->opt = *opt;
    });
}

int pa_problem_set_multigrid(pa_problem* P, int nx, int ny, int nz, const pa_multigrid_options* mopt) {
    return guarded([&] {
        PA_REQUIRE(P && mopt, PA_ERR_ARG, "pa_problem_set_multigrid: null argument");
        pa_multigrid* mg = nullptr;
        pa::check_status(pa_multigrid_create(P->vf->mesh, nx, ny, nz, P->vf->dpn, mopt, &mg));
        // elastostatic and thermal Jacobians are symmetric, and so are they
        // after the symmetric Dirichlet elimination
        pa::multigrid_set_elliptic(mg);
        if (P->mg) pa_multigrid_destroy(P->mg);
        P->mg = mg;
        pa_solver* s = nullptr;
        pa::check_status(pa_solver_create(P->vf->mesh, P->vf->dpn, PA_PRECOND_MULTIGRID, &s));
        pa::solver_set_symmetric(s);
        s->mg = mg;
        pa_solver_destroy(P->solver);
        P->solver = s;
        P->opt.precond = PA_PRECOND_MULTIGRID;
    });
}

int pa_problem_solution(pa_problem* P, const double* z, double* u, pa_solve_stats* stats, void* stream) {
    return guarded([&] {
        PA_REQUIRE(P && z && u && stats, PA_ERR_ARG, "pa_problem_solution: null argument");
        const hipStream_t st = (hipStream_t)stream;
        pa_vector_function* vf = P->vf;
        const pa_mesh* m = vf->mesh;
        const int dpn = vf->dpn;
        const long n = (long)m->N * dpn;
        const long nv = (long)m->nnzb * dpn * dpn;
        const int nvb = P->solver->nvb;
        PA_HIP(hipMemsetAsync(u, 0, n * sizeof(double), st));
        const bool mixed = P->opt.mode == PA_FP32;
        void* K;
        if (mixed) {
            if (P->K32.n != (size_t)nv) P->K32.alloc(nv);
            K = P->K32.p;
        } else {
            if (P->K64.n != (size_t)nv) P->K64.alloc(nv);
            K = P->K64.p;
        }
        // R(0, z), K(z); rhs = -R; symmetric Dirichlet elimination
        pa::check_status(pa_vector_function_value_gradient_u(vf, u, z, P->rhs.p, K, P->opt.mode, stream));
        pa::k_neg<<<nvb, pa::kB, 0, st>>>(n, P->rhs.p);
        pa::check();
        if (P->nbc)
            pa::apply_block_constraints(m, dpn, K, P->opt.mode, P->rhs.p, P->nbc, P->bc_dofs.p, P->bc_vals.p, 1.0,
                                        P->bc_dense.p, P->bc_touch.p, st);
        if (P->opt.precond == PA_PRECOND_MULTIGRID) {
            PA_REQUIRE(P->mg, PA_ERR_STATE, "pa_problem_solution: multigrid selected but not set");
            pa::check_status(pa_multigrid_setup(P->mg, K, P->opt.mode, P->nbc, P->bc_dofs.p, stream));
        }
        const double tol = P->opt.tol;
        if (!mixed) {
            pa::pcg(P->solver, K, PA_FP64, P->rhs.p, u, tol, P->opt.max_iter, P->opt.check_every, stats, st);
            return;
        }
        // mixed: CG on the FP32 matrix with reliable updates, the recursive
        // residual replaced by the exact FP64 one, -R(u) from the
        // matrix-free assembly kernels (constrained rows g - u)
        const pa::ExactResidual exact = [&](const double* x, double* r) {
            pa::check_status(pa_vector_function_value(vf, x, z, r, stream));
            pa::k_neg<<<nvb, pa::kB, 0, st>>>(n, r);
            pa::check();
            if (P->nbc) {
                pa::k_bc_residual<<<(P->nbc + 255) / 256, 256, 0, st>>>(P->nbc, P->bc_dofs.p, P->bc_vals.p, x, r);
                pa::check();
            }
        };
        const int every = P->opt.replace_every > 0 ? P->opt.replace_every : 0;
        if (dpn == 3)
            pa::pcg_impl<3, float>(P->solver, (const float*)K, P->rhs.p, u, tol, P->opt.max_iter,
                                   P->opt.check_every, stats, st, &exact, every);
        else
            pa::pcg_impl<1, float>(P->solver, (const float*)K, P->rhs.p, u, tol, P->opt.max_iter,
                                   P->opt.check_every, stats, st, &exact, every);
    });
}

int pa_problem_criterion_value(pa_problem* P, const double* z, const double* u, double* value, void* stream) {
    return guarded([&] {
        PA_REQUIRE(P && z && u && value, PA_ERR_ARG, "pa_problem_criterion_value: null argument");
        pa::check_status(pa_scalar_function_value(P->sf, u, z, value, stream));
    });
}

int pa_problem_criterion_gradient(pa_problem* P, const double* z, const double* u, double* grad, void* stream) {
    return guarded([&] {
        PA_REQUIRE(P && z && u && grad, PA_ERR_ARG, "pa_problem_criterion_gradient: null argument");
        pa::self_adjoint_gradient(P->vf, P->sf, u, z, grad, (hipStream_t)stream);
    });
}

int pa_problem_criterion_gradient_x(pa_problem* P, const double* z, const double* u, double* grad, void* stream) {
    return guarded([&] {
        PA_REQUIRE(P && z && u && grad, PA_ERR_ARG, "pa_problem_criterion_gradient_x: null argument");
        // self-adjoint: dF/dx = -(df/dx), src/elliptic/Problem.hpp:512-515
        pa::criterion_gradient_x(P->sf, u, z, -1.0, grad, (hipStream_t)stream);
    });
}

}  // extern "C"
