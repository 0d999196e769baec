// PlasticityProblem on the device: pseudo-time steps, the condensed Newton
// solve of every step, the projection update, the plastic-work criterion and
// its path-dependent adjoint gradient.
//
// Linear solves: restarted flexible GMRES (default) or BiCGStab, with the
// inverses of the D x D nodal diagonal blocks of the tangent (D = 4: ux, uy,
// uz, p of the saddle-point form; D = 3 in the pure-displacement form) or the
// geometric multigrid on D x D blocks as the preconditioner.  Reductions are
// fixed-order, so runs are deterministic.
//
// Step states (u, c, pi) live in slots on the device.  With checkpoint
// stride 1 every step has its own slot (a 64x32x32-hex problem with 10 steps
// holds about 0.5 GB).  With stride m > 1 only the kept steps (m-1, 2m-1, ...,
// and the last: each ends a segment of m steps) have their own slot, and the
// m - 1 other steps of a segment share m - 1 slots; the criterion and the
// backward recursion recompute a segment's unkept steps from the kept step
// before it by the same Newton solves, which reproduce the states bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pa_j2.h"

struct pa_plasticity_problem {
    pa_plasticity* p = nullptr;
    pa_plasticity_options opt{};
    int nbc = 0;
    std::vector<int> h_dofs;
    std::vector<double> h_vals;
    pa::DevBuf<int> bc_dofs;
    pa::DevBuf<double> bc_inc, bc_zero;
    pa::DevBuf<double> bc_dense;  // constraint work vectors (N * dpn, N)
    pa::DevBuf<unsigned char> bc_touch;
    // step state slots (see slot())
    pa::DevBuf<double> U, Cs, PI;
    int nslots = 0, nkept = 0;
    pa::DevBuf<double> zeroU, zeroC;
    // Newton / Krylov workspace
    pa::DevBuf<double> vals, R, rhs, du;
    pa::DevBuf<double> r, rh, pv, v, y, s, zz, t, minv, part, scal;
    pa::DevBuf<double> gV, gZ, gpart, gh, gy;  // GMRES basis, preconditioned basis, partials, Hessenberg column, y
    // adjoint workspace
    pa::DevBuf<double> wu, wu_next, gc, gcp, gcp_next, zc, zc_out, blocal, mu, mu_next, lam, gam, gam_next;
    std::vector<int> newton_its, linear_its;
    long reorth = 0;  // Gram-Schmidt passes repeated (all GMRES solves)
    // outcome of the last Krylov solve
    bool lin_converged = true;
    double lin_rel = 0.0;
    std::vector<double> res_norms;
    bool solved = false;
    bool verbose = false;  // PA_PLASTICITY_VERBOSE set: one stderr line per Newton iteration
    pa_multigrid* mg = nullptr;  // owned; preconditioner of the linear solves when set
    ~pa_plasticity_problem() {
        if (mg) pa_multigrid_destroy(mg);
    }
};

namespace pa {
namespace j2 {
namespace {

constexpr int kB = 256;
void check() { PA_HIP(hipGetLastError()); }
int blocks(long n) { return (int)((n + kB - 1) / kB); }

enum { S_RHO = 0, S_RHO_OLD, S_RV, S_TS, S_TT, S_RR, S_BB, S_NORM, S_N = 16 };

__device__ inline double block_sum(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = v;
    __syncthreads();
    double tt = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / 64); ++i) tt += sh[i];
    __syncthreads();
    return tt;
}

// Fixed-order sums of `count` partial arrays (stride nparts) into scal[dst_k];
// with shift, scal[S_RHO_OLD] = scal[S_RHO] first.
__global__ void k_reduce(const double* __restrict__ part, int n, int nparts, int count, int d0, int d1, int shift,
                         double* __restrict__ scal) {
    __shared__ double sh[1024 / 64];
    const int dst[2] = {d0, d1};
    for (int kk = 0; kk < count; ++kk) {
        double a = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) a += part[(long)kk * nparts + i];
        const double tt = block_sum(a, sh);
        if (threadIdx.x == 0) {
            if (shift && kk == 0) scal[S_RHO_OLD] = scal[S_RHO];
            scal[dst[kk]] = tt;
        }
        __syncthreads();
    }
}

// y = A x (D x D blocks, 16 lanes per block row); partial sums of y.a1 and y.a2.
template <int D>
__global__ __launch_bounds__(kB) void k_spmv(int N, const int* __restrict__ rowmap, const int* __restrict__ colidx,
                                             const double* __restrict__ vals, const double* __restrict__ x,
                                             double* __restrict__ y, const double* __restrict__ a1,
                                             const double* __restrict__ a2, double* __restrict__ part,
                                             int nparts) {
    constexpr int G = 16, ROWS = kB / G;
    __shared__ double sh[kB / 64];
    const int grp = threadIdx.x / G, lane = threadIdx.x % G;
    const int row = blockIdx.x * ROWS + grp;
    double acc[D] = {};
    if (row < N) {
        for (int e = rowmap[row] + lane; e < rowmap[row + 1]; e += G) {
            const long c = colidx[e];
            double xv[D];
#pragma unroll
            for (int j = 0; j < D; ++j) xv[j] = x[c * D + j];
            const double* b = vals + (long)e * (D * D);
#pragma unroll
            for (int i = 0; i < D; ++i)
#pragma unroll
                for (int j = 0; j < D; ++j) acc[i] += b[i * D + j] * xv[j];
        }
    }
    double d1 = 0.0, d2 = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) acc[i] += __shfl_xor(acc[i], o, G);
        if (row < N && lane == 0) {
            const long o = (long)row * D + i;
            y[o] = acc[i];
            if (a1) d1 += acc[i] * a1[o];
            if (a2) d2 += acc[i] * a2[o];
        }
    }
    const double s1 = block_sum(d1, sh), s2 = block_sum(d2, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s1;
        part[nparts + blockIdx.x] = s2;
    }
}

// Inverse of each D x D diagonal block, Gauss-Jordan with partial pivoting.
template <int D>
__global__ void k_block_inverse(int N, const int* __restrict__ rowmap, const int* __restrict__ colidx,
                                const double* __restrict__ vals, double* __restrict__ minv) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    int e = rowmap[n];
    while (colidx[e] != n) ++e;
    double A[D][2 * D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            A[i][j] = vals[(long)e * (D * D) + i * D + j];
            A[i][D + j] = (i == j) ? 1.0 : 0.0;
        }
#pragma unroll
    for (int k = 0; k < D; ++k) {
        int piv = k;
        double best = fabs(A[k][k]);
#pragma unroll
        for (int i = k + 1; i < D; ++i)
            if (fabs(A[i][k]) > best) {
                best = fabs(A[i][k]);
                piv = i;
            }
#pragma unroll
        for (int i = k + 1; i < D; ++i)
            if (i == piv)
#pragma unroll
                for (int j = 0; j < 2 * D; ++j) {
                    const double tmp = A[k][j];
                    A[k][j] = A[i][j];
                    A[i][j] = tmp;
                }
        const double inv = A[k][k] != 0.0 ? 1.0 / A[k][k] : 0.0;
#pragma unroll
        for (int j = 0; j < 2 * D; ++j) A[k][j] *= inv;
#pragma unroll
        for (int i = 0; i < D; ++i)
            if (i != k) {
                const double f = A[i][k];
#pragma unroll
                for (int j = 0; j < 2 * D; ++j) A[i][j] -= f * A[k][j];
            }
    }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) minv[(long)n * (D * D) + i * D + j] = A[i][D + j];
}

template <int D>
__device__ inline void prec_block(const double* __restrict__ minv, long n, const double in[D], double out[D]) {
    const double* m = minv + n * (D * D);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) s += m[i * D + j] * in[j];
        out[i] = s;
    }
}

// r = b - q, rh = r, p = v = 0; partials of r.r (-> rho and rr) and b.b.
__global__ __launch_bounds__(kB) void k_bicg_init(long n, const double* __restrict__ b, const double* __restrict__ q,
                                                  double* __restrict__ r, double* __restrict__ rh,
                                                  double* __restrict__ p, double* __restrict__ v,
                                                  double* __restrict__ part, int nparts) {
    __shared__ double sh[kB / 64];
    double rr = 0.0, bb = 0.0;
    for (long i = (long)blockIdx.x * kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) {
        const double ri = b[i] - q[i];
        r[i] = ri;
        rh[i] = ri;
        p[i] = 0.0;
        v[i] = 0.0;
        rr += ri * ri;
        bb += b[i] * b[i];
    }
    const double s1 = block_sum(rr, sh), s2 = block_sum(bb, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s1;
        part[nparts + blockIdx.x] = s2;
    }
}

// p = r + beta (p - omega v); y = M p; beta = (rho / rho_old)(alpha / omega).
template <int D>
__global__ __launch_bounds__(kB) void k_bicg_dir(int N, const double* __restrict__ scal,
                                                 const double* __restrict__ minv, const double* __restrict__ r,
                                                 const double* __restrict__ v, double* __restrict__ p,
                                                 double* __restrict__ y) {
    const double alpha = scal[S_RHO_OLD] / scal[S_RV];
    const double omega = scal[S_TS] / scal[S_TT];
    const double beta = (scal[S_RHO] / scal[S_RHO_OLD]) * (alpha / omega);
    for (long n = (long)blockIdx.x * kB + threadIdx.x; n < N; n += (long)gridDim.x * kB) {
        double pn[D], yn[D];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const long o = n * D + i;
            pn[i] = r[o] + beta * (p[o] - omega * v[o]);
            p[o] = pn[i];
        }
        if (!minv) continue;  // multigrid: y = M p is applied afterwards
        prec_block<D>(minv, n, pn, yn);
#pragma unroll
        for (int i = 0; i < D; ++i) y[n * D + i] = yn[i];
    }
}

// s = r - alpha v; z = M s; alpha = rho / (rh.v).
template <int D>
__global__ __launch_bounds__(kB) void k_bicg_s(int N, const double* __restrict__ scal, const double* __restrict__ minv,
                                               const double* __restrict__ r, const double* __restrict__ v,
                                               double* __restrict__ s, double* __restrict__ zz) {
    const double alpha = scal[S_RHO] / scal[S_RV];
    for (long n = (long)blockIdx.x * kB + threadIdx.x; n < N; n += (long)gridDim.x * kB) {
        double sn[D], zn[D];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const long o = n * D + i;
            sn[i] = r[o] - alpha * v[o];
            s[o] = sn[i];
        }
        if (!minv) continue;  // multigrid: z = M s is applied afterwards
        prec_block<D>(minv, n, sn, zn);
#pragma unroll
        for (int i = 0; i < D; ++i) zz[n * D + i] = zn[i];
    }
}

// x += alpha y + omega z; r = s - omega t; partials of rh.r and r.r.
__global__ __launch_bounds__(kB) void k_bicg_update(long n, const double* __restrict__ scal,
                                                    const double* __restrict__ y, const double* __restrict__ zz,
                                                    const double* __restrict__ s, const double* __restrict__ t,
                                                    const double* __restrict__ rh, double* __restrict__ x,
                                                    double* __restrict__ r, double* __restrict__ part, int nparts) {
    __shared__ double sh[kB / 64];
    const double alpha = scal[S_RHO] / scal[S_RV];
    const double omega = scal[S_TS] / scal[S_TT];
    double d1 = 0.0, d2 = 0.0;
    for (long i = (long)blockIdx.x * kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) {
        x[i] += alpha * y[i] + omega * zz[i];
        const double ri = s[i] - omega * t[i];
        r[i] = ri;
        d1 += rh[i] * ri;
        d2 += ri * ri;
    }
    const double a = block_sum(d1, sh), b = block_sum(d2, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[nparts + blockIdx.x] = b;
    }
}

__global__ void k_set_scalars(double* scal) {
    scal[S_RHO_OLD] = 1.0;
    scal[S_RV] = 1.0;
    scal[S_TS] = 1.0;
    scal[S_TT] = 1.0;
}

__global__ void k_neg_copy(long n, const double* __restrict__ a, double* __restrict__ b) {
    for (long i = (long)blockIdx.x * kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) b[i] = -a[i];
}

__global__ void k_axpy(long n, double a, const double* __restrict__ x, double* __restrict__ y) {
    for (long i = (long)blockIdx.x * kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) y[i] += a * x[i];
}

__global__ __launch_bounds__(kB) void k_dot_part(long n, const double* __restrict__ a, double* __restrict__ part) {
    __shared__ double sh[kB / 64];
    double s = 0.0;
    for (long i = (long)blockIdx.x * kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) s += a[i] * a[i];
    const double tt = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tt;
}

double read_scal(pa_plasticity_problem* P, int idx, hipStream_t st) {
    double v = 0.0;
    PA_HIP(hipMemcpyAsync(&v, P->scal.p + idx, sizeof(double), hipMemcpyDeviceToHost, st));
    PA_HIP(hipStreamSynchronize(st));
    return v;
}

// global unknowns of the formulation (dpn per node)
long num_dofs(const pa_plasticity_problem* P) { return (long)P->p->dpn * P->p->mesh->N; }
int vec_blocks(const pa_plasticity_problem* P) { return std::min(blocks(num_dofs(P)), 2048); }
int row_blocks(const pa_plasticity_problem* P) { return (P->p->mesh->N + 15) / 16; }

double norm(pa_plasticity_problem* P, const double* x, hipStream_t st) {
    const long n = num_dofs(P);
    const int nb = vec_blocks(P);
    k_dot_part<<<nb, kB, 0, st>>>(n, x, P->part.p);
    check();
    k_reduce<<<1, 1024, 0, st>>>(P->part.p, nb, nb, 1, S_NORM, 0, 0, P->scal.p);
    check();
    return std::sqrt(read_scal(P, S_NORM, st));
}

template <int D>
__global__ void k_block_prec(int N, const double* __restrict__ minv, const double* __restrict__ r,
                             double* __restrict__ z);

// Kernels of the block size D of the formulation.
struct BlockKernels {
    decltype(&k_spmv<4>) spmv;
    decltype(&k_block_inverse<4>) inverse;
    decltype(&k_bicg_dir<4>) dir;
    decltype(&k_bicg_s<4>) s;
    decltype(&k_block_prec<4>) prec;
};
BlockKernels block_kernels(const pa_plasticity_problem* P) {
    if (P->p->dpn == 4) return {k_spmv<4>, k_block_inverse<4>, k_bicg_dir<4>, k_bicg_s<4>, k_block_prec<4>};
    return {k_spmv<3>, k_block_inverse<3>, k_bicg_dir<3>, k_bicg_s<3>, k_block_prec<3>};
}

// BiCGStab with block-Jacobi preconditioning on the D x D block matrix P->vals:
// solves A x = b from the x given; returns the iterations.
int bicgstab(pa_plasticity_problem* P, const double* b, double* x, hipStream_t st) {
    const pa_mesh* m = P->p->mesh;
    const int N = m->N;
    const long n = num_dofs(P);
    const BlockKernels K = block_kernels(P);
    const int nvb = vec_blocks(P), nsb = row_blocks(P);
    const int np = std::max(nvb, nsb);
    const double* A = P->vals.p;
    const double* minv = P->mg ? nullptr : P->minv.p;
    if (P->mg) {
        const int rc = pa_multigrid_setup(P->mg, A, PA_FP64, P->nbc, P->bc_dofs.p, st);
        if (rc != PA_OK) throw Error(rc, pa_last_error());
    } else {
        K.inverse<<<blocks(N), kB, 0, st>>>(N, m->d_rowmap.p, m->d_colidx.p, A, P->minv.p);
        check();
    }
    K.spmv<<<nsb, kB, 0, st>>>(N, m->d_rowmap.p, m->d_colidx.p, A, x, P->t.p, nullptr, nullptr, P->part.p, np);
    check();
    k_bicg_init<<<nvb, kB, 0, st>>>(n, b, P->t.p, P->r.p, P->rh.p, P->pv.p, P->v.p, P->part.p, np);
    check();
    k_set_scalars<<<1, 1, 0, st>>>(P->scal.p);
    check();
    k_reduce<<<1, 1024, 0, st>>>(P->part.p, nvb, np, 2, S_RHO, S_BB, 0, P->scal.p);
    check();
    const double bnorm = std::sqrt(read_scal(P, S_BB, st));
    double rr = read_scal(P, S_RHO, st);
    const double target = P->opt.linear_tol * bnorm;
    const int check_every = 10;
    int it = 0;
    while (it < P->opt.linear_max_iter && bnorm > 0.0 && std::sqrt(rr) > target) {
        const int burst = std::min(check_every, P->opt.linear_max_iter - it);
        for (int q = 0; q < burst; ++q) {
            K.dir<<<nvb, kB, 0, st>>>(N, P->scal.p, minv, P->r.p, P->v.p, P->pv.p, P->y.p);
            check();
            if (P->mg) multigrid_apply(P->mg, P->pv.p, P->y.p, st);
            K.spmv<<<nsb, kB, 0, st>>>(N, m->d_rowmap.p, m->d_colidx.p, A, P->y.p, P->v.p, P->rh.p, nullptr,
                                       P->part.p, np);
            check();
            k_reduce<<<1, 1024, 0, st>>>(P->part.p, nsb, np, 1, S_RV, 0, 0, P->scal.p);
            check();
            K.s<<<nvb, kB, 0, st>>>(N, P->scal.p, minv, P->r.p, P->v.p, P->s.p, P->zz.p);
            check();
            if (P->mg) multigrid_apply(P->mg, P->s.p, P->zz.p, st);
            K.spmv<<<nsb, kB, 0, st>>>(N, m->d_rowmap.p, m->d_colidx.p, A, P->zz.p, P->t.p, P->s.p, P->t.p,
                                       P->part.p, np);
            check();
            k_reduce<<<1, 1024, 0, st>>>(P->part.p, nsb, np, 2, S_TS, S_TT, 0, P->scal.p);
            check();
            k_bicg_update<<<nvb, kB, 0, st>>>(n, P->scal.p, P->y.p, P->zz.p, P->s.p, P->t.p, P->rh.p, x, P->r.p,
                                              P->part.p, np);
            check();
            k_reduce<<<1, 1024, 0, st>>>(P->part.p, nvb, np, 2, S_RHO, S_RR, 1, P->scal.p);
            check();
        }
        it += burst;
        rr = read_scal(P, S_RR, st);
        if (!std::isfinite(rr)) break;
    }
    P->lin_rel = bnorm > 0.0 ? std::sqrt(rr) / bnorm : 0.0;
    P->lin_converged = std::isfinite(rr) && (bnorm == 0.0 || std::sqrt(rr) <= target);
    return it;
}

// ---------------------------------------------------------------------------
// flexible restarted GMRES (right preconditioning; the FP32 multigrid cycle
// need not be an exactly linear operator), classical Gram-Schmidt repeated
// once when the first pass loses too much of the vector's norm; the
// (m+1) x m Hessenberg least-squares problem and its Givens rotations live on
// the host
// ---------------------------------------------------------------------------
// part[i*nparts + block] = partial sums of w . V_i, i < k (V_i at V + i*n);
// with self, row k holds the partial sums of w . w.  blockIdx.y picks a
// group of kMdotRows rows, so the groups' loads are in flight at once.
constexpr int kMdotRows = 4;
__global__ __launch_bounds__(kB) void k_mdot(long n, int k, int self, const double* __restrict__ V,
                                             const double* __restrict__ w, double* __restrict__ part, int nparts) {
    constexpr int Q = kMdotRows;
    __shared__ double sh[Q][kB / 64];
    const int rows = k + (self ? 1 : 0);
    const int i0 = blockIdx.y * Q;
    double s[Q] = {};
    for (long t = (long)blockIdx.x * kB + threadIdx.x; t < n; t += (long)gridDim.x * kB) {
        const double wt = w[t];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int i = i0 + q;
            if (i < k) s[q] += wt * V[(long)i * n + t];
            else if (i == k && self) s[q] += wt * wt;
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        double v = s[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (threadIdx.x % 64 == 0) sh[q][threadIdx.x / 64] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < Q && i0 + (int)threadIdx.x < rows) {
        double tt = 0.0;
        for (int wv = 0; wv < kB / 64; ++wv) tt += sh[threadIdx.x][wv];
        part[(long)(i0 + threadIdx.x) * nparts + blockIdx.x] = tt;
    }
}

// h[i] (+)= sum over the nb partials of array i; one block per i, fixed order.
__global__ void k_reduce_rows(const double* __restrict__ part, int nb, int nparts, int accumulate,
                              double* __restrict__ h) {
    __shared__ double sh[1024 / 64];
    const int i = blockIdx.x;
    double s = 0.0;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) s += part[(long)i * nparts + t];
    const double tt = block_sum(s, sh);
    if (threadIdx.x == 0) h[i] = accumulate ? h[i] + tt : tt;
}

// w -= sum_{i<k} h[i] V_i; part[block] = partial sums of the new w . w
__global__ __launch_bounds__(kB) void k_mupdate(long n, int k, const double* __restrict__ V,
                                                const double* __restrict__ h, double* __restrict__ w,
                                                double* __restrict__ part) {
    __shared__ double sh[kB / 64];
    double q = 0.0;
    for (long t = (long)blockIdx.x * kB + threadIdx.x; t < n; t += (long)gridDim.x * kB) {
        double s = w[t];
        for (int i = 0; i < k; ++i) s -= h[i] * V[(long)i * n + t];
        w[t] = s;
        q += s * s;
    }
    const double tt = block_sum(q, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = tt;
}

// x += sum_{i<k} y[i] Z_i
__global__ __launch_bounds__(kB) void k_madd(long n, int k, const double* __restrict__ Z,
                                             const double* __restrict__ y, double* __restrict__ x) {
    for (long t = (long)blockIdx.x * kB + threadIdx.x; t < n; t += (long)gridDim.x * kB) {
        double s = x[t];
        for (int i = 0; i < k; ++i) s += y[i] * Z[(long)i * n + t];
        x[t] = s;
    }
}

__global__ void k_scale(long n, double f, double* v) {
    for (long t = (long)blockIdx.x * kB + threadIdx.x; t < n; t += (long)gridDim.x * kB) v[t] *= f;
}

__global__ void k_sub(long n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ c) {
    for (long t = (long)blockIdx.x * kB + threadIdx.x; t < n; t += (long)gridDim.x * kB) c[t] = a[t] - b[t];
}

template <int D>
__global__ void k_block_prec(int N, const double* __restrict__ minv, const double* __restrict__ r,
                             double* __restrict__ z) {
    for (long n = (long)blockIdx.x * kB + threadIdx.x; n < N; n += (long)gridDim.x * kB) {
        double in[D], out[D];
#pragma unroll
        for (int i = 0; i < D; ++i) in[i] = r[n * D + i];
        prec_block<D>(minv, n, in, out);
#pragma unroll
        for (int i = 0; i < D; ++i) z[n * D + i] = out[i];
    }
}

void preconditioner_setup(pa_plasticity_problem* P, hipStream_t st) {
    const pa_mesh* m = P->p->mesh;
    if (P->mg) {
        const int rc = pa_multigrid_setup(P->mg, P->vals.p, PA_FP64, P->nbc, P->bc_dofs.p, st);
        if (rc != PA_OK) throw Error(rc, pa_last_error());
    } else {
        block_kernels(P).inverse<<<blocks(m->N), kB, 0, st>>>(m->N, m->d_rowmap.p, m->d_colidx.p, P->vals.p,
                                                               P->minv.p);
        check();
    }
}

void apply_preconditioner(pa_plasticity_problem* P, const double* r, double* z, hipStream_t st) {
    if (P->mg) {
        multigrid_apply(P->mg, r, z, st);
    } else {
        block_kernels(P).prec<<<vec_blocks(P), kB, 0, st>>>(P->p->mesh->N, P->minv.p, r, z);
        check();
    }
}

int gmres(pa_plasticity_problem* P, const double* b, double* x, hipStream_t st) {
    const pa_mesh* m = P->p->mesh;
    const int N = m->N;
    const long n = num_dofs(P);
    const int M = P->opt.restart;
    const auto spmv = block_kernels(P).spmv;
    const int nvb = vec_blocks(P), nsb = row_blocks(P), np = std::max(nvb, nsb);
    double* V = P->gV.p;
    double* Z = P->gZ.p;
    preconditioner_setup(P, st);
    const double bnorm = norm(P, b, st);
    P->lin_converged = true;
    P->lin_rel = 0.0;
    if (!(bnorm > 0.0)) return 0;
    const double target = P->opt.linear_tol * bnorm;
    P->lin_converged = false;
    P->lin_rel = 1.0;
    std::vector<double> H((M + 1) * M), cs(M), sn(M), g(M + 1), y(M), hcol(M + 1), h1(M + 3), h2(M + 3);
    int it = 0;
    while (it < P->opt.linear_max_iter) {
        // r = b - A x into V_0
        spmv<<<nsb, kB, 0, st>>>(N, m->d_rowmap.p, m->d_colidx.p, P->vals.p, x, P->t.p, nullptr, nullptr,
                                 P->part.p, np);
        check();
        k_sub<<<nvb, kB, 0, st>>>(n, b, P->t.p, V);
        check();
        const double beta = norm(P, V, st);
        if (!std::isfinite(beta)) return it;
        P->lin_rel = beta / bnorm;
        if (beta <= target) {
            P->lin_converged = true;
            break;
        }
        k_scale<<<nvb, kB, 0, st>>>(n, 1.0 / beta, V);  // V_0 = r / beta
        check();
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        bool done = false;
        for (; j < M && it < P->opt.linear_max_iter; ++j, ++it) {
            double* vj = V + (long)j * n;
            double* zj = Z + (long)j * n;
            double* w = V + (long)(j + 1) * n;
            apply_preconditioner(P, vj, zj, st);
            spmv<<<nsb, kB, 0, st>>>(N, m->d_rowmap.p, m->d_colidx.p, P->vals.p, zj, w, nullptr, nullptr,
                                     P->part.p, np);
            check();
            // classical Gram-Schmidt against V_0..V_j: h = V^T w and |w|^2 in
            // one pass over the basis, w -= V h and |w|^2 in another.  The
            // pair is repeated when the new norm is below 1/sqrt(2) of the
            // old (the criterion of Daniel, Gragg, Kaufman and Stewart,
            // Math. Comp. 30 (1976)); one repetition restores orthogonality
            // to working precision.
            double* g1 = P->gh.p;
            double* g2 = P->gh.p + (M + 3);
            k_mdot<<<dim3(nvb, (j + 2 + kMdotRows - 1) / kMdotRows), kB, 0, st>>>(n, j + 1, 1, V, w, P->gpart.p, nvb);
            check();
            k_reduce_rows<<<j + 2, 1024, 0, st>>>(P->gpart.p, nvb, nvb, 0, g1);  // h, |w|^2
            check();
            k_mupdate<<<nvb, kB, 0, st>>>(n, j + 1, V, g1, w, P->part.p);
            check();
            k_reduce_rows<<<1, 1024, 0, st>>>(P->part.p, nvb, nvb, 0, g1 + j + 2);  // |w - V h|^2
            check();
            PA_HIP(hipMemcpyAsync(h1.data(), g1, (j + 3) * sizeof(double), hipMemcpyDeviceToHost, st));
            PA_HIP(hipStreamSynchronize(st));
            double hn2 = h1[j + 2];
            for (int i = 0; i <= j; ++i) hcol[i] = h1[i];
            if (!(hn2 > 0.5 * h1[j + 1])) {
                k_mdot<<<dim3(nvb, (j + 1 + kMdotRows - 1) / kMdotRows), kB, 0, st>>>(n, j + 1, 0, V, w, P->gpart.p,
                                                                                      nvb);
                check();
                k_reduce_rows<<<j + 1, 1024, 0, st>>>(P->gpart.p, nvb, nvb, 0, g2);
                check();
                k_mupdate<<<nvb, kB, 0, st>>>(n, j + 1, V, g2, w, P->part.p);
                check();
                k_reduce_rows<<<1, 1024, 0, st>>>(P->part.p, nvb, nvb, 0, g2 + j + 1);
                check();
                PA_HIP(hipMemcpyAsync(h2.data(), g2, (j + 2) * sizeof(double), hipMemcpyDeviceToHost, st));
                PA_HIP(hipStreamSynchronize(st));
                for (int i = 0; i <= j; ++i) hcol[i] += h2[i];
                hn2 = h2[j + 1];
                ++P->reorth;
            }
            const double hn = std::sqrt(hn2);
            hcol[j + 1] = hn;
            if (!std::isfinite(hn)) return it;
            if (hn > 0.0) {
                k_scale<<<nvb, kB, 0, st>>>(n, 1.0 / hn, w);  // V_{j+1} = w / hn
                check();
            }
            // apply the previous rotations, then a new one zeroing H[j+1][j]
            for (int i = 0; i < j; ++i) {
                const double a = hcol[i], c = hcol[i + 1];
                hcol[i] = cs[i] * a + sn[i] * c;
                hcol[i + 1] = -sn[i] * a + cs[i] * c;
            }
            const double a = hcol[j], c = hcol[j + 1];
            const double rr = std::hypot(a, c);
            cs[j] = rr > 0.0 ? a / rr : 1.0;
            sn[j] = rr > 0.0 ? c / rr : 0.0;
            hcol[j] = rr;
            hcol[j + 1] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            for (int i = 0; i <= j; ++i) H[i * M + j] = hcol[i];
            P->lin_rel = std::fabs(g[j + 1]) / bnorm;
            if (std::fabs(g[j + 1]) <= target || hn == 0.0) {
                ++j;
                ++it;
                done = true;
                break;
            }
        }
        // y = R^-1 g (upper triangular j x j), x += Z y
        for (int i = j - 1; i >= 0; --i) {
            double s = g[i];
            for (int k2 = i + 1; k2 < j; ++k2) s -= H[i * M + k2] * y[k2];
            y[i] = s / H[i * M + i];
        }
        PA_HIP(hipMemcpyAsync(P->gy.p, y.data(), j * sizeof(double), hipMemcpyHostToDevice, st));
        k_madd<<<nvb, kB, 0, st>>>(n, j, Z, P->gy.p, x);
        check();
        PA_HIP(hipStreamSynchronize(st));
        if (done) {
            P->lin_converged = std::isfinite(P->lin_rel);
            break;
        }
    }
    return it;
}

__global__ void k_copy_bc(int nbc, const int* __restrict__ dofs, const double* __restrict__ b, double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nbc) x[dofs[i]] = b[dofs[i]];
}

// A x = b for the constrained tangent.  The Dirichlet rows are identity rows
// (their columns zeroed) whose scale is unrelated to the stiffness rows, so a
// relative residual test cannot see their error; x starts at b on those dofs
// and every search direction is zero there (both preconditioners and A map
// such vectors to such vectors), so they stay exact.
int linear_solve(pa_plasticity_problem* P, const double* b, double* x, hipStream_t st) {
    if (P->nbc) {
        k_copy_bc<<<(P->nbc + 255) / 256, 256, 0, st>>>(P->nbc, P->bc_dofs.p, b, x);
        check();
    }
    return P->opt.krylov == PA_KRYLOV_BICGSTAB ? bicgstab(P, b, x, st) : gmres(P, b, x, st);
}

// Checkpoint layout: stride m, kept steps m-1, 2m-1, ... and K-1.
int stride(const pa_plasticity_problem* P) { return std::max(1, P->opt.checkpoint_stride); }
bool kept(const pa_plasticity_problem* P, int k) {
    const int m = stride(P);
    return m == 1 || k == P->opt.num_steps - 1 || k % m == m - 1;
}
// Slot of step k: its own for a kept step (k / m), else one of the m - 1
// slots shared by the segments' unkept steps (nkept + k % m).
int slot(const pa_plasticity_problem* P, int k) {
    const int m = stride(P);
    if (m == 1) return k;
    return kept(P, k) ? k / m : P->nkept + k % m;
}
double* step_u(pa_plasticity_problem* P, int k) { return P->U.p + (long)slot(P, k) * num_dofs(P); }
double* step_c(pa_plasticity_problem* P, int k) { return P->Cs.p + (long)slot(P, k) * 14 * P->p->mesh->C; }
double* step_pi(pa_plasticity_problem* P, int k) { return P->PI.p + (long)slot(P, k) * 3 * P->p->mesh->N; }
const double* prev_u(pa_plasticity_problem* P, int k) { return k ? step_u(P, k - 1) : P->zeroU.p; }
const double* prev_c(pa_plasticity_problem* P, int k) { return k ? step_c(P, k - 1) : P->zeroC.p; }

void constraints(pa_plasticity_problem* P, double* rhs, const double* bcv, hipStream_t st) {
    if (P->nbc == 0) return;
    apply_block_constraints(P->p->mesh, P->p->dpn, P->vals.p, PA_FP64, rhs, P->nbc, P->bc_dofs.p, bcv, 1.0,
                            P->bc_dense.p, P->bc_touch.p, st);
}

// The projected pressure gradient step k uses (src/PlasticityProblem.hpp:872-891):
// zero in the first step and in the pure-displacement form, else the
// projection of the previous step's pressure.
void step_projection(pa_plasticity_problem* P, int k, const double* z, hipStream_t st) {
    if (P->p->dpn == 4 && k > 0) projection(P->p, step_u(P, k - 1), z, step_pi(P, k), st);
    else PA_HIP(hipMemsetAsync(step_pi(P, k), 0, 3L * P->p->mesh->N * sizeof(double), st));
}

// Newton solve of step k (src/NewtonRaphsonSolver.hpp:555-610) from the states
// of step k - 1 and the step's projection; returns convergence and, with
// record, keeps the iteration counts.
bool forward_step(pa_plasticity_problem* P, int k, const double* z, bool record, hipStream_t st) {
    const int K = P->opt.num_steps;
    const long nu = num_dofs(P);
    double* uk = step_u(P, k);
    double* ck = step_c(P, k);
    const double* up = prev_u(P, k);
    const double* cp = prev_c(P, k);
    const double* pik = step_pi(P, k);
    // Dirichlet increment of this step: g(t_k) - g(t_{k-1}), t_k = (k+1)/K
    const double dt = 1.0 / K, tk = dt * (k + 1), tp = dt * k;
    std::vector<double> inc(P->nbc);
    for (int j = 0; j < P->nbc; ++j) inc[j] = P->h_vals[j] * tk - (k > 0 ? P->h_vals[j] * tp : 0.0);
    if (P->nbc) PA_HIP(hipMemcpyAsync(P->bc_inc.p, inc.data(), P->nbc * sizeof(double), hipMemcpyHostToDevice, st));
    PA_HIP(hipMemcpyAsync(uk, up, nu * sizeof(double), hipMemcpyDeviceToDevice, st));
    update_local_state(P->p, uk, cp, z, ck, st);
    int it = 0, lin = 0;
    double ref = 0.0, measure = HUGE_VAL, nrm = 0.0;
    bool converged = false;
    while (true) {
        condensed_system(P->p, uk, ck, cp, pik, z, P->vals.p, P->R.p, false, st);
        k_neg_copy<<<vec_blocks(P), kB, 0, st>>>(nu, P->R.p, P->rhs.p);
        check();
        constraints(P, P->rhs.p, it == 0 ? P->bc_inc.p : P->bc_zero.p, st);
        nrm = norm(P, P->rhs.p, st);
        if (it == 0) {
            ref = nrm;
        } else {
            measure = std::fabs(ref - nrm);
            ref = nrm;
        }
        if (measure < P->opt.newton_stop_tol || nrm < P->opt.newton_current_norm_tol) {
            converged = true;
            break;
        }
        if (nrm >= 1e20 || it >= P->opt.max_newton || !std::isfinite(nrm)) break;
        PA_HIP(hipMemsetAsync(P->du.p, 0, nu * sizeof(double), st));
        const int li = linear_solve(P, P->rhs.p, P->du.p, st);
        lin += li;
        if (P->verbose)
            std::fprintf(stderr, "plasticity: step %d newton %d |r| %.6e linear %d (Gram-Schmidt repeated %ld times so far)\n",
                         k, it, nrm, li, P->reorth);
        k_axpy<<<vec_blocks(P), kB, 0, st>>>(nu, 1.0, P->du.p, uk);
        check();
        update_local_state(P->p, uk, cp, z, ck, st);
        ++it;
    }
    if (record) {
        P->newton_its[k] = it;
        P->linear_its[k] = lin;
        P->res_norms[k] = nrm;
    }
    return converged;
}

// Forward solve (src/PlasticityProblem.hpp:784-868).
bool solve_forward(pa_plasticity_problem* P, const double* z, hipStream_t st) {
    const int K = P->opt.num_steps;
    P->newton_its.assign(K, 0);
    P->linear_its.assign(K, 0);
    P->res_norms.assign(K, 0.0);
    bool all_converged = true;
    for (int k = 0; k < K; ++k) {
        step_projection(P, k, z, st);
        all_converged = forward_step(P, k, z, true, st) && all_converged;
    }
    PA_HIP(hipStreamSynchronize(st));
    P->solved = true;
    return all_converged;
}

// Segment s: steps [first, last], last a kept step.
int num_segments(const pa_plasticity_problem* P) { return (P->opt.num_steps + stride(P) - 1) / stride(P); }
int seg_first(const pa_plasticity_problem* P, int s) { return s * stride(P); }
int seg_last(const pa_plasticity_problem* P, int s) {
    return std::min(s * stride(P) + stride(P) - 1, P->opt.num_steps - 1);
}

// States of a segment's unkept steps, recomputed from the kept step before it.
void recompute_segment(pa_plasticity_problem* P, int s, const double* z, hipStream_t st) {
    for (int k = seg_first(P, s); k < seg_last(P, s); ++k) {
        step_projection(P, k, z, st);
        forward_step(P, k, z, false, st);
    }
}

// Plastic-work criterion over the steps (src/PlasticityProblem.hpp:925-956);
// with checkpoints each segment is recomputed first.
double criterion_value(pa_plasticity_problem* P, const double* z, double mult, hipStream_t st) {
    double total = 0.0;
    for (int s = 0; s < num_segments(P); ++s) {
        if (stride(P) > 1) recompute_segment(P, s, z, st);
        for (int k = seg_first(P, s); k <= seg_last(P, s); ++k) {
            plastic_work(P->p, step_u(P, k), prev_u(P, k), step_c(P, k), prev_c(P, k), z, P->scal.p + S_NORM, st);
            total += read_scal(P, S_NORM, st);
        }
    }
    return mult * total;
}

// One step of the backward recursion (src/PathDependentAdjointSolver.hpp:
// 277-917): the adjoint of step k from the adjoint of step k + 1 (kept in the
// *_next buffers), its control terms added to grad.  Returns false when the
// global adjoint solve stopped short of the tolerance.
bool backward_step(pa_plasticity_problem* P, int k, const double* z, double mult, double* grad, hipStream_t st) {
    const pa_mesh* m = P->p->mesh;
    const int K = P->opt.num_steps;
    const long nu = num_dofs(P);
    const bool mixed = P->p->dpn == 4;
    const bool last = (k == K - 1);
    const double* u = step_u(P, k);
    const double* c = step_c(P, k);
    const double* up = prev_u(P, k);
    const double* cp = prev_c(P, k);
    const double* pi = step_pi(P, k);
    AdjointWork w{P->wu.p, P->gc.p, P->gcp.p, P->zc.p};
    work_partials(P->p, u, up, c, cp, z, mult, w, st);
    adjoint_local_rhs(P->p, u, c, cp, z, P->gc.p, last ? nullptr : P->gcp_next.p, last ? nullptr : P->mu_next.p,
                      P->wu.p, last ? nullptr : P->wu_next.p, last ? nullptr : P->gam_next.p, P->blocal.p, st);
    gather_vcell(P->p, -1.0, P->rhs.p, st);
    condensed_system(P->p, u, c, cp, pi, z, P->vals.p, nullptr, true, st);
    constraints(P, P->rhs.p, P->bc_zero.p, st);
    PA_HIP(hipMemsetAsync(P->lam.p, 0, nu * sizeof(double), st));
    linear_solve(P, P->rhs.p, P->lam.p, st);
    const bool ok = P->lin_converged;
    adjoint_local_update(P->p, u, c, cp, pi, z, P->lam.p, P->blocal.p, P->zc.p, P->mu.p, P->zc_out.p, st);
    if (k == 0 || !mixed) {  // no projection adjoint in the first step or the displacement form
        PA_HIP(hipMemsetAsync(P->gam.p, 0, 3L * m->N * sizeof(double), st));
    } else {
        gather_uniform(P->p, 3, P->p->vcell.p, 16, -1.0, false, P->gam.p, st);
        divide_rowsum(P->p, P->gam.p, z, st);
        if (P->p->prm.projection_penalty) projection_control_term(P->p, u, pi, P->gam.p, z, P->zc_out.p, st);
    }
    gather_uniform(P->p, 1, P->zc_out.p, 1.0, true, grad, st);
    std::swap(P->wu.p, P->wu_next.p);
    std::swap(P->gcp.p, P->gcp_next.p);
    std::swap(P->mu.p, P->mu_next.p);
    std::swap(P->gam.p, P->gam_next.p);
    return ok;
}

// Backward recursion (src/PlasticityProblem.hpp:1035-1066) over the segments
// from the last, each recomputed first when checkpointed.  The recompute
// touches only the forward workspace, not the adjoint carried between steps.
void criterion_gradient(pa_plasticity_problem* P, const double* z, double mult, double* grad, hipStream_t st) {
    PA_HIP(hipMemsetAsync(grad, 0, P->p->mesh->N * sizeof(double), st));
    int failed = -1;
    double worst = 0.0;
    for (int s = num_segments(P) - 1; s >= 0; --s) {
        if (stride(P) > 1) recompute_segment(P, s, z, st);
        for (int k = seg_last(P, s); k >= seg_first(P, s); --k) {
            if (!backward_step(P, k, z, mult, grad, st) && failed < 0) {
                failed = k;
                worst = P->lin_rel;
            }
        }
    }
    PA_HIP(hipStreamSynchronize(st));
    if (failed >= 0) {
        char msg[160];
        std::snprintf(msg, sizeof msg,
                      "criterion_gradient: the adjoint linear solve of step %d stopped at relative residual %.3e "
                      "(linear_tol %.1e, linear_max_iter %d)",
                      failed, worst, P->opt.linear_tol, P->opt.linear_max_iter);
        throw Error(PA_ERR_STATE, msg);
    }
}

}  // namespace
}  // namespace j2
}  // namespace pa

using pa::guarded;

extern "C" {

void pa_plasticity_options_default(pa_plasticity_options* o) {
    if (!o) return;
    o->num_steps = 20;
    o->max_newton = 10;
    o->newton_stop_tol = 1e-8;
    o->newton_current_norm_tol = 1e-8;
    o->linear_max_iter = 20000;
    o->linear_tol = 1e-12;
    o->krylov = PA_KRYLOV_GMRES;
    o->restart = 50;
    o->checkpoint_stride = 1;
}

int pa_plasticity_problem_create(pa_plasticity* p, int nbc, const int* dofs, const double* vals,
                                 const pa_plasticity_options* opt, pa_plasticity_problem** out) {
    return guarded([&] {
        PA_REQUIRE(p && out && (nbc == 0 || (dofs && vals)) && nbc >= 0, PA_ERR_ARG,
                   "pa_plasticity_problem_create: bad argument");
        auto* P = new pa_plasticity_problem();
        try {
            P->p = p;
            P->verbose = std::getenv("PA_PLASTICITY_VERBOSE") != nullptr;
            if (opt) P->opt = *opt;
            else pa_plasticity_options_default(&P->opt);
            PA_REQUIRE(P->opt.num_steps >= 1 && P->opt.max_newton >= 0 && P->opt.linear_max_iter >= 1 &&
                           P->opt.restart >= 1 && P->opt.restart <= 1000 && P->opt.checkpoint_stride >= 0 &&
                           P->opt.checkpoint_stride <= P->opt.num_steps &&
                           (P->opt.krylov == PA_KRYLOV_GMRES || P->opt.krylov == PA_KRYLOV_BICGSTAB), PA_ERR_ARG,
                       "pa_plasticity_problem_create: bad options");
            const pa_mesh* m = p->mesh;
            const long nu = (long)p->dpn * m->N;
            for (int j = 0; j < nbc; ++j)
                PA_REQUIRE(dofs[j] >= 0 && dofs[j] < nu, PA_ERR_ARG, "pa_plasticity_problem_create: dof out of range");
            P->nbc = nbc;
            P->h_dofs.assign(dofs, dofs + nbc);
            P->h_vals.assign(vals, vals + nbc);
            if (nbc) {
                P->bc_dofs.upload(dofs, nbc);
                P->bc_inc.alloc(nbc);
                P->bc_zero.alloc(nbc);
                PA_HIP(hipMemset(P->bc_zero.p, 0, nbc * sizeof(double)));
                P->bc_dense.alloc((size_t)p->dpn * p->mesh->N);
                P->bc_touch.alloc(p->mesh->N);
            }
            const int K = P->opt.num_steps;
            const int mstride = pa::j2::stride(P);
            P->nkept = mstride == 1 ? K : (K + mstride - 1) / mstride;
            P->nslots = mstride == 1 ? K : P->nkept + mstride - 1;
            P->U.alloc(nu * P->nslots);
            P->Cs.alloc(14L * m->C * P->nslots);
            P->PI.alloc(3L * m->N * P->nslots);
            P->zeroU.alloc(nu);
            P->zeroC.alloc(14L * m->C);
            PA_HIP(hipMemset(P->zeroU.p, 0, nu * sizeof(double)));
            PA_HIP(hipMemset(P->zeroC.p, 0, 14L * m->C * sizeof(double)));
            P->vals.alloc((long)p->dpn * p->dpn * m->nnzb);
            for (auto* b : {&P->R, &P->rhs, &P->du, &P->r, &P->rh, &P->pv, &P->v, &P->y, &P->s, &P->zz, &P->t,
                            &P->lam})
                b->alloc(nu);
            P->minv.alloc((long)p->dpn * p->dpn * m->N);
            {
                const long M = P->opt.restart;
                P->gV.alloc((M + 1) * nu);
                P->gZ.alloc(M * nu);
                P->gpart.alloc((M + 1) * pa::j2::vec_blocks(P));
                P->gh.alloc(2 * (M + 3));
                P->gy.alloc(M);
            }
            const int np = std::max(pa::j2::vec_blocks(P), pa::j2::row_blocks(P));
            P->part.alloc(2L * np);
            P->scal.alloc(pa::j2::S_N);
            for (auto* b : {&P->wu, &P->wu_next}) b->alloc(16L * m->C);
            for (auto* b : {&P->gc, &P->gcp, &P->gcp_next, &P->blocal, &P->mu, &P->mu_next}) b->alloc(14L * m->C);
            P->zc.alloc(m->C);
            P->zc_out.alloc(m->C);
            P->gam.alloc(3L * m->N);
            P->gam_next.alloc(3L * m->N);
        } catch (...) {
            delete P;
            throw;
        }
        *out = P;
    });
}

int pa_plasticity_problem_destroy(pa_plasticity_problem* P) {
    return guarded([&] { delete P; });
}

int pa_plasticity_problem_set_multigrid(pa_plasticity_problem* P, int nx, int ny, int nz,
                                        const pa_multigrid_options* opt) {
    return guarded([&] {
        PA_REQUIRE(P, PA_ERR_ARG, "pa_plasticity_problem_set_multigrid: null argument");
        pa_multigrid_options o{};
        if (opt) o = *opt;
        else {
            o.pre_sweeps = 2;
            o.post_sweeps = 2;
            o.omega = 0.6;
        }
        pa_multigrid* mg = nullptr;
        const int rc = pa_multigrid_create(P->p->mesh, nx, ny, nz, P->p->dpn, &o, &mg);
        if (rc != PA_OK) throw pa::Error(rc, pa_last_error());
        if (P->mg) pa_multigrid_destroy(P->mg);
        P->mg = mg;
    });
}

int pa_plasticity_problem_solution(pa_plasticity_problem* P, const double* z, int* converged, void* stream) {
    return guarded([&] {
        PA_REQUIRE(P && z, PA_ERR_ARG, "pa_plasticity_problem_solution: null argument");
        const bool ok = pa::j2::solve_forward(P, z, (hipStream_t)stream);
        if (converged) *converged = ok ? 1 : 0;
    });
}

int pa_plasticity_problem_state(const pa_plasticity_problem* P, int step, const double** u, const double** c,
                                const double** pi) {
    return guarded([&] {
        PA_REQUIRE(P && step >= 0 && step < P->opt.num_steps, PA_ERR_ARG, "pa_plasticity_problem_state: bad step");
        PA_REQUIRE(P->solved, PA_ERR_STATE, "pa_plasticity_problem_state: no solution yet");
        PA_REQUIRE(pa::j2::kept(P, step), PA_ERR_STATE,
                   "pa_plasticity_problem_state: step not kept (checkpoint stride " +
                       std::to_string(pa::j2::stride(P)) + ")");
        auto* Q = const_cast<pa_plasticity_problem*>(P);
        if (u) *u = pa::j2::step_u(Q, step);
        if (c) *c = pa::j2::step_c(Q, step);
        if (pi) *pi = pa::j2::step_pi(Q, step);
    });
}

int pa_plasticity_problem_state_bytes(const pa_plasticity_problem* P, size_t* bytes) {
    return guarded([&] {
        PA_REQUIRE(P && bytes, PA_ERR_ARG, "pa_plasticity_problem_state_bytes: null argument");
        *bytes = (P->U.n + P->Cs.n + P->PI.n) * sizeof(double);
    });
}

int pa_plasticity_problem_stats(const pa_plasticity_problem* P, int step, int* newton, int* linear, double* res) {
    return guarded([&] {
        PA_REQUIRE(P && step >= 0 && step < P->opt.num_steps, PA_ERR_ARG, "pa_plasticity_problem_stats: bad step");
        PA_REQUIRE(P->solved, PA_ERR_STATE, "pa_plasticity_problem_stats: no solution yet");
        if (newton) *newton = P->newton_its[step];
        if (linear) *linear = P->linear_its[step];
        if (res) *res = P->res_norms[step];
    });
}

int pa_plasticity_problem_criterion_value(pa_plasticity_problem* P, const double* z, double mult, double* value,
                                          void* stream) {
    return guarded([&] {
        PA_REQUIRE(P && z && value, PA_ERR_ARG, "pa_plasticity_problem_criterion_value: null argument");
        PA_REQUIRE(P->solved, PA_ERR_STATE, "pa_plasticity_problem_criterion_value: no solution yet");
        *value = pa::j2::criterion_value(P, z, mult, (hipStream_t)stream);
    });
}

int pa_plasticity_problem_criterion_gradient(pa_plasticity_problem* P, const double* z, double mult, double* grad,
                                             void* stream) {
    return guarded([&] {
        PA_REQUIRE(P && z && grad, PA_ERR_ARG, "pa_plasticity_problem_criterion_gradient: null argument");
        PA_REQUIRE(P->solved, PA_ERR_STATE, "pa_plasticity_problem_criterion_gradient: no solution yet");
        pa::j2::criterion_gradient(P, z, mult, grad, (hipStream_t)stream);
    });
}

}  // extern "C"
