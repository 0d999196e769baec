// J2 plasticity on Tet4 cells: the stabilized mixed u-p global residual
// (src/InfinitesimalStrainPlasticityResidual.hpp:393-521), the 14-variable
// local J2 residual and its closed-form return map
// (src/J2PlasticityLocalResidual.hpp:310-543), the Schur-condensed Newton
// linearization (src/NewtonRaphsonSolver.hpp:204-330), the pressure-gradient
// projection (src/PressureGradientProjectionResidual.hpp:111-162) and the
// plastic-work criterion (src/PlasticWorkCriterion.hpp:163-265), with every
// partial derivative written out by hand in place of the reference's FAD
// evaluation types.
//
// Work decomposition:
//   * thread-per-cell kernels for the return map, the local residual, the
//     global cell residual, the projection sources and the plastic work;
//   * the condensed system: one 16-lane group per cell, lane r owning row r
//     of the local system [dH/dc | dH/du | H]; Gauss-Jordan elimination
//     without pivoting (the reference inverts dH/dc by LU without pivoting,
//     src/BLAS3.hpp:154-183) with the pivot rows broadcast by lane shuffles,
//     then lane r forms row r of K_t = dR/du - dR/dc (dH/dc)^-1 dH/du and of
//     R_t; the 16x16 cell matrix is staged in LDS and stored contiguously in
//     block layout [a][b][i][j];
//   * a node-ordered gather (one 16-lane group per graph row, lane = entry of
//     the 4x4 block) sums the cell blocks into the block-CRS values in a fixed
//     order: no atomics, deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "pa_j2.h"

namespace pa {
namespace j2 {
namespace {

constexpr int kB = 256;
constexpr double kSqrt32 = 1.2247448713915890491;   // sqrt(3/2)
constexpr double kSqrt23 = 0.81649658092772603273;  // sqrt(2/3)

void check() { PA_HIP(hipGetLastError()); }

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ inline double powp(double x, double p) {
    if (p == 1.0) return x;
    if (p == 2.0) return x * x;
    if (p == 3.0) return x * x * x;
    return pow(x, p);
}

// MSIMP (src/Simp.hpp:130-134) and its derivative.
__device__ inline void msimp(double rho, double p, double emin, double& v, double& dv) {
    v = emin + (1.0 - emin) * powp(rho, p);
    dv = (rho == 0.0) ? 0.0 : (1.0 - emin) * p * powp(rho, p - 1.0);
}

// v[i] with i known only at run time, by bit masks (a select chain between
// elements of one array is folded back into a dynamic index, i.e. scratch).
template <int M>
__device__ inline double sel(const double* v, int i) {
    unsigned long long b = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) b |= (unsigned long long)__double_as_longlong(v[k]) & (0ULL - (unsigned long long)(i == k));
    return __longlong_as_double((long long)b);
}

// Deviatoric stress, src/ComputeDeviatoricStress.hpp:107-127.
__device__ inline void dev(double mu, const double e[6], double s[6]) {
    const double tr = (e[0] + e[1] + e[2]) / 3.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) s[i] = (2.0 * mu) * (e[i] - tr);
#pragma unroll
    for (int i = 3; i < 6; ++i) s[i] = mu * e[i];
}

// Entry (i, j) of the deviatoric operator d s / d e with shear modulus mu.
__device__ inline double dv_entry(double mu, int i, int j) {
    if (i < 3 && j < 3) return 2.0 * mu * ((i == j ? 1.0 : 0.0) - 1.0 / 3.0);
    return (i == j) ? mu : 0.0;
}

// B^T q for one node: sum_v q_v d eps_v / d u_{b,j} (Voigt xx yy zz yz xz xy).
__device__ inline void bt(const double g[3], const double q[6], double out[3]) {
    out[0] = q[0] * g[0] + q[5] * g[1] + q[4] * g[2];
    out[1] = q[1] * g[1] + q[5] * g[0] + q[3] * g[2];
    out[2] = q[2] * g[2] + q[4] * g[0] + q[3] * g[1];
}

// Row i (run-time) of the symmetric stress tensor stored in Voigt order.
__device__ inline void tensor_row(const double s[6], int i, double out[3]) {
    const double r0[3] = {s[0], s[5], s[4]}, r1[3] = {s[5], s[1], s[3]}, r2[3] = {s[4], s[3], s[2]};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const double t[3] = {r0[j], r1[j], r2[j]};
        out[j] = sel<3>(t, i);
    }
}

// ---------------------------------------------------------------------------
// cell input: geometry (ComputeGradientWorkset<3>, src/ImplicitFunctors.hpp:169-221),
// total strain (Strain<3,4>, src/Strain.hpp:56-95), nodal pressures, density
// ---------------------------------------------------------------------------
struct CellIn {
    double g[4][3];
    double V;  // |det J| / 6
    double eps[6];
    double p[4];
    double rho;
};

__device__ inline void cell_geometry(const double4* __restrict__ X, const int nn[4], double g[4][3], double& V) {
    double x[4][3];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const double4 q = X[nn[n]];
        x[n][0] = q.x;
        x[n][1] = q.y;
        x[n][2] = q.z;
    }
    const double j11 = x[0][0] - x[3][0], j12 = x[1][0] - x[3][0], j13 = x[2][0] - x[3][0];
    const double j21 = x[0][1] - x[3][1], j22 = x[1][1] - x[3][1], j23 = x[2][1] - x[3][1];
    const double j31 = x[0][2] - x[3][2], j32 = x[1][2] - x[3][2], j33 = x[2][2] - x[3][2];
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
#pragma unroll
    for (int d = 0; d < 3; ++d) g[3][d] = -(g[0][d] + g[1][d] + g[2][d]);
    V = fabs(det) * (1.0 / 6.0);
}

// u holds D dofs per node: (ux, uy, uz, p) for D = 4, (ux, uy, uz) and p = 0 for D = 3.
__device__ inline void strain_of(const double g[4][3], const double* __restrict__ u, int D, const int nn[4],
                                 double eps[6], double p[4]) {
    double uu[4][3];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const double* un = u + (long)nn[n] * D;
        uu[n][0] = un[0];
        uu[n][1] = un[1];
        uu[n][2] = un[2];
        p[n] = D == 4 ? un[3] : 0.0;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double s = 0.0;
#pragma unroll
        for (int n = 0; n < 4; ++n) s += uu[n][d] * g[n][d];
        eps[d] = s;
    }
    double s3 = 0.0, s4 = 0.0, s5 = 0.0;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        s3 += uu[n][1] * g[n][2] + uu[n][2] * g[n][1];
        s4 += uu[n][0] * g[n][2] + uu[n][2] * g[n][0];
        s5 += uu[n][0] * g[n][1] + uu[n][1] * g[n][0];
    }
    eps[3] = s3;
    eps[4] = s4;
    eps[5] = s5;
}

__device__ inline void load_cell(const int* __restrict__ conn, int e, const double4* __restrict__ X,
                                 const double* __restrict__ u, int D, const double* __restrict__ z, CellIn& C,
                                 int nn[4]) {
    const int4 q = reinterpret_cast<const int4*>(conn)[e];
    nn[0] = q.x;
    nn[1] = q.y;
    nn[2] = q.z;
    nn[3] = q.w;
    cell_geometry(X, nn, C.g, C.V);
    strain_of(C.g, u, D, nn, C.eps, C.p);
    C.rho = (z[nn[0]] + z[nn[1]] + z[nn[2]] + z[nn[3]]) / 4.0;
}

__device__ inline void load14(const double* __restrict__ src, int e, double out[14]) {
    const double* s = src + (long)e * 14;
#pragma unroll
    for (int k = 0; k < 14; ++k) out[k] = s[k];
}

// ---------------------------------------------------------------------------
// local J2 residual: context, residual, rows of its Jacobians
// ---------------------------------------------------------------------------
struct Loc {
    bool plastic;
    double c0, c1;
    double pe, dpe, pp, dpp;
    double muL;
    double s0[6];  // Dev(mu)(eps - eps_p), unpenalized
    double n[6];   // (s - alpha) / ||s - alpha||
    double wn[6];  // W n, W = diag(1, 1, 1, 2, 2, 2)
    double wd[6];  // (W n)^T d s / d eps_e
    double nrm;
    double hiso, hkin, sy, yield;
};

__device__ inline void local_ctx(const KP& k, const CellIn& C, const double c[14], Loc& L) {
    L.c0 = c[0];
    L.c1 = c[1];
    L.plastic = c[1] > 0.0;
    msimp(C.rho, k.pe_e, k.pe_m, L.pe, L.dpe);
    msimp(C.rho, k.pp_e, k.pp_m, L.pp, L.dpp);
    L.muL = L.pe * k.mu;
    double ee[6], s[6], eta[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) ee[i] = C.eps[i] - c[2 + i];
    dev(L.muL, ee, s);
    dev(k.mu, ee, L.s0);
#pragma unroll
    for (int i = 0; i < 6; ++i) eta[i] = s[i] - c[8 + i];
    L.nrm = sqrt(eta[0] * eta[0] + eta[1] * eta[1] + eta[2] * eta[2] + 2.0 * (eta[3] * eta[3]) +
                 2.0 * (eta[4] * eta[4]) + 2.0 * (eta[5] * eta[5]));
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        L.n[i] = eta[i] / L.nrm;
        L.wn[i] = (i < 3 ? 1.0 : 2.0) * L.n[i];
    }
    const double t = (L.wn[0] + L.wn[1] + L.wn[2]) / 3.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) L.wd[j] = 2.0 * L.muL * (L.wn[j] - t);
#pragma unroll
    for (int j = 3; j < 6; ++j) L.wd[j] = L.muL * L.wn[j];
    L.hiso = L.pp * k.hiso;
    L.hkin = L.pp * k.hkin;
    L.sy = L.pp * k.sy;
    L.yield = L.sy + L.hiso * c[0];
}

// H, src/J2PlasticityLocalResidual.hpp:397-424.
__device__ inline double local_residual_row(const Loc& L, const double c[14], const double cp[14], int r) {
    const double cr = sel<14>(c, r), cpr = sel<14>(cp, r);
    if (!L.plastic) return r == 1 ? L.c1 : cr - cpr;
    if (r == 0) return c[0] - cp[0] - c[1];
    if (r == 1) return kSqrt32 * L.nrm - L.yield;
    if (r < 8) {
        const int i = r - 2;
        const double f = i < 3 ? 1.0 : 2.0;
        return cr - cpr - f * kSqrt32 * c[1] * sel<6>(L.n, i);
    }
    return cr - cpr - kSqrt23 * c[1] * L.hkin * sel<6>(L.n, r - 8);
}

// Row r of dH/dc (14) and the strain-space vector q with dH_r/du = B^T q.
__device__ inline void dh_row(const Loc& L, int r, double row[14], double q[6]) {
#pragma unroll
    for (int k = 0; k < 14; ++k) row[k] = (!L.plastic && k == r) ? 1.0 : 0.0;
#pragma unroll
    for (int v = 0; v < 6; ++v) q[v] = 0.0;
    if (!L.plastic) return;
    if (r == 0) {
        row[0] = 1.0;
        row[1] = -1.0;
        return;
    }
    if (r == 1) {
        row[0] = -L.hiso;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            row[2 + j] = -kSqrt32 * L.wd[j];
            row[8 + j] = -kSqrt32 * L.wn[j];
            q[j] = kSqrt32 * L.wd[j];
        }
        return;
    }
    if (r >= 14) return;
    const bool ep = r < 8;
    const int i = ep ? r - 2 : r - 8;
    const double ni = sel<6>(L.n, i);
    const double Fi = i < 3 ? 1.0 : 2.0;
    const double a = ep ? kSqrt32 * Fi * L.c1 : kSqrt23 * L.c1 * L.hkin;
    row[1] = ep ? -kSqrt32 * Fi * ni : -kSqrt23 * L.hkin * ni;
    const double inv = 1.0 / L.nrm;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const double ndv = (dv_entry(L.muL, i, j) - ni * L.wd[j]) * inv;
        const double nn = ((i == j ? 1.0 : 0.0) - ni * L.wn[j]) * inv;
        row[2 + j] = ((ep && i == j) ? 1.0 : 0.0) + a * ndv;
        row[8 + j] = ((!ep && i == j) ? 1.0 : 0.0) + a * nn;
        q[j] = -a * ndv;
    }
}

// d H_r / d rho (the control partial of node a is 1/4 of it).
__device__ inline double dh_drho(const KP& k, const Loc& L, int r) {
    if (!L.plastic || r == 0 || r >= 14) return 0.0;
    double wx = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) wx += L.wn[j] * (L.dpe * L.s0[j]);
    if (r == 1) return kSqrt32 * wx - (L.dpp * k.sy + L.dpp * k.hiso * L.c0);
    const bool ep = r < 8;
    const int i = ep ? r - 2 : r - 8;
    const double ni = sel<6>(L.n, i), xi = L.dpe * sel<6>(L.s0, i);
    const double Nx = (xi - ni * wx) / L.nrm;
    if (ep) return -kSqrt32 * (i < 3 ? 1.0 : 2.0) * L.c1 * Nx;
    return -kSqrt23 * L.c1 * (L.dpp * k.hkin * ni + L.hkin * Nx);
}

// ---------------------------------------------------------------------------
// global residual: context, residual rows, rows of its Jacobians
// ---------------------------------------------------------------------------
struct Glob {
    double pg, dpg, kG;
    double sG[6];  // Dev(pg mu)(eps - eps_p)
    double s0[6];  // Dev(mu)(eps - eps_p)
    double pgp;    // mixed: ps * interpolated pressure; displacement: kG tr(eps - eps_p)
    double tre;    // tr(eps - eps_p)
    double stab[3];
    double tau, volstrain;
};

__device__ inline void global_ctx(const KP& k, const CellIn& C, const double c[14], const double pig[3], Glob& G) {
    double ee[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) ee[i] = C.eps[i] - c[2 + i];
    msimp(C.rho, k.pg_e, k.pg_m, G.pg, G.dpg);
    dev(G.pg * k.mu, ee, G.sG);
    dev(k.mu, ee, G.s0);
    double gradp[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double s = 0.0;
#pragma unroll
        for (int n = 0; n < 4; ++n) s += C.p[n] * C.g[n][d];
        gradp[d] = s;
    }
    double p = 0.0;
#pragma unroll
    for (int n = 0; n < 4; ++n) p += 0.25 * C.p[n];
    G.kG = G.pg * k.kappa;
    G.tre = (ee[0] + ee[1]) + ee[2];
    if (!k.mixed) {  // the volumetric stress takes the pressure's place; no stabilization
        G.pgp = G.kG * G.tre;
        G.volstrain = G.tau = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) G.stab[d] = 0.0;
        return;
    }
    G.pgp = p * k.ps;
    const double div = C.eps[0] + C.eps[1] + C.eps[2];
    G.volstrain = k.ps * (div - G.pgp / G.kG);
    G.tau = pow(C.V, 2.0 / 3.0) / (2.0 * k.mu);
#pragma unroll
    for (int d = 0; d < 3; ++d) G.stab[d] = k.ps * G.tau * ((k.ps * gradp[d]) - pig[d]);
}

// R_{a,i}, row r = a*4 + i (the pressure row i = 3 is empty in the displacement formulation).
__device__ inline double global_residual_row(const KP& k, const CellIn& C, const Glob& G, const double ga[3], int i) {
    if (i < 3) {
        double srow[3];
        tensor_row(G.sG, i, srow);
        double R = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) R += C.V * srow[j] * ga[j];
        R += C.V * G.pgp * sel<3>(ga, i);
        return R;
    }
    if (!k.mixed) return 0.0;
    double R = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) R += -1.0 * G.stab[d] * ga[d] * C.V;
    R += 0.25 * G.volstrain * C.V;
    return R;
}

// Row (a, i) of dR/du (16 columns b*4 + j).
__device__ inline void dr_du_row(const KP& k, const CellIn& C, const Glob& G, const double ga[3], int i,
                                 double out[16]) {
    const double muG = G.pg * k.mu;
    const double gai = sel<3>(ga, i);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const double* gb = C.g[b];
        const double dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
        if (!k.mixed) {
            // V (muG (dot delta_ij + g_aj g_bi - 2/3 g_ai g_bj) + kG g_ai g_bj); no pressure row or column
            const double gbi = sel<3>(C.g[b], i);
#pragma unroll
            for (int j = 0; j < 3; ++j)
                out[b * 4 + j] =
                    i < 3 ? C.V * muG * ((i == j ? dot : 0.0) + ga[j] * gbi - (2.0 / 3.0) * gai * gb[j]) +
                                C.V * G.kG * gai * gb[j]
                          : 0.0;
            out[b * 4 + 3] = 0.0;
        } else if (i < 3) {
            const double gbi = sel<3>(C.g[b], i);
#pragma unroll
            for (int j = 0; j < 3; ++j)
                out[b * 4 + j] = C.V * muG * ((i == j ? dot : 0.0) + ga[j] * gbi - (2.0 / 3.0) * gai * gb[j]);
            out[b * 4 + 3] = C.V * k.ps * 0.25 * gai;
        } else {
#pragma unroll
            for (int j = 0; j < 3; ++j) out[b * 4 + j] = 0.25 * C.V * k.ps * gb[j];
            out[b * 4 + 3] = -C.V * k.ps * k.ps * G.tau * dot - C.V * k.ps * k.ps / (16.0 * G.kG);
        }
    }
}

// Row (a, i) of dR/d eps_p (the only local states the global residual sees).
__device__ inline void dr_dep_row(const KP& k, const CellIn& C, const Glob& G, const double ga[3], int i,
                                  double out[6]) {
    const double muG = G.pg * k.mu;
    if (i >= 3) {
#pragma unroll
        for (int w = 0; w < 6; ++w) out[w] = 0.0;
        return;
    }
    const double gai = sel<3>(ga, i);
#pragma unroll
    for (int w = 0; w < 3; ++w) out[w] = -C.V * gai * 2.0 * muG * ((i == w ? 1.0 : 0.0) - 1.0 / 3.0);
    // shear w: the j with voigt(i, j) = w
    out[3] = -C.V * muG * (i == 1 ? ga[2] : (i == 2 ? ga[1] : 0.0));
    out[4] = -C.V * muG * (i == 0 ? ga[2] : (i == 2 ? ga[0] : 0.0));
    out[5] = -C.V * muG * (i == 0 ? ga[1] : (i == 1 ? ga[0] : 0.0));
    if (!k.mixed)
#pragma unroll
        for (int w = 0; w < 3; ++w) out[w] -= C.V * G.kG * gai;
}

// d R_{a,i} / d rho.
__device__ inline double dr_drho_row(const KP& k, const CellIn& C, const Glob& G, const double ga[3], int i) {
    if (i < 3) {
        double srow[3];
        tensor_row(G.s0, i, srow);
        const double d = C.V * G.dpg * (srow[0] * ga[0] + srow[1] * ga[1] + srow[2] * ga[2]);
        return k.mixed ? d : d + C.V * G.dpg * k.kappa * G.tre * sel<3>(ga, i);
    }
    if (!k.mixed) return 0.0;
    return 0.25 * C.V * k.ps * G.pgp * G.dpg / (G.pg * G.pg * k.kappa);
}

__device__ inline void pick_node(const double g[4][3], int a, double out[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double t[4] = {g[0][d], g[1][d], g[2][d], g[3][d]};
        out[d] = sel<4>(t, a);
    }
}

__device__ inline void interp_pi(const double* __restrict__ pi, const int nn[4], double pig[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double s = 0.0;
#pragma unroll
        for (int n = 0; n < 4; ++n) s += 0.25 * pi[(long)nn[n] * 3 + d];
        pig[d] = s;
    }
}

// ---------------------------------------------------------------------------
// thread-per-cell kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kB) void k_update(int C, KP k, const int* __restrict__ conn,
                                               const double4* __restrict__ X, const double* __restrict__ u,
                                               const double* __restrict__ cp, const double* __restrict__ z,
                                               double* __restrict__ cout) {
    const int e = blockIdx.x * kB + threadIdx.x;
    if (e >= C) return;
    CellIn Ci;
    int nn[4];
    load_cell(conn, e, X, u, k.dpn, z, Ci, nn);
    double c[14];
    load14(cp, e, c);
    c[1] = 0.0;
    Loc L;
    local_ctx(k, Ci, c, L);
    const double ftrial = kSqrt32 * L.nrm - L.yield;
    if (ftrial > 1.0e-10) {
        const double cpv0 = c[0];
        c[1] = ftrial / (3.0 * L.muL + L.hiso + L.hkin);
        c[0] = cpv0 + c[1];
        const double m1 = c[1] * kSqrt32;
#pragma unroll
        for (int i = 0; i < 6; ++i) c[2 + i] = c[2 + i] + (i < 3 ? 1.0 : 2.0) * m1 * L.n[i];
        const double m2 = c[1] * kSqrt23 * L.hkin;
#pragma unroll
        for (int i = 0; i < 6; ++i) c[8 + i] = c[8 + i] + m2 * L.n[i];
    }
    double* o = cout + (long)e * 14;
#pragma unroll
    for (int q = 0; q < 14; ++q) o[q] = c[q];
}

// Local residual (MODE 0) or one row-block of a local workset Jacobian.
// Thread per (cell, row); wrt: PA_WRT_C, PA_WRT_U, PA_WRT_Z, PA_WRT_C_PREV.
__global__ __launch_bounds__(kB) void k_local(int C, int mode, int wrt, KP k, const int* __restrict__ conn,
                                              const double4* __restrict__ X, const double* __restrict__ u,
                                              const double* __restrict__ c_in, const double* __restrict__ cp_in,
                                              const double* __restrict__ z, double* __restrict__ out) {
    const long t = (long)blockIdx.x * kB + threadIdx.x;
    const int e = (int)(t / 14), r = (int)(t % 14);
    if (e >= C) return;
    CellIn Ci;
    int nn[4];
    load_cell(conn, e, X, u, k.dpn, z, Ci, nn);
    double c[14], cp[14];
    load14(c_in, e, c);
    load14(cp_in, e, cp);
    Loc L;
    local_ctx(k, Ci, c, L);
    if (mode == 0) {
        out[(long)e * 14 + r] = local_residual_row(L, c, cp, r);
        return;
    }
    if (wrt == PA_WRT_C) {
        double row[14], q[6];
        dh_row(L, r, row, q);
        double* o = out + ((long)e * 14 + r) * 14;
#pragma unroll
        for (int j = 0; j < 14; ++j) o[j] = row[j];
    } else if (wrt == PA_WRT_U) {
        double row[14], q[6];
        dh_row(L, r, row, q);
        double* o = out + ((long)e * 14 + r) * 16;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            double v[3];
            bt(Ci.g[b], q, v);
            o[b * 4 + 0] = v[0];
            o[b * 4 + 1] = v[1];
            o[b * 4 + 2] = v[2];
            o[b * 4 + 3] = 0.0;
        }
    } else if (wrt == PA_WRT_Z) {
        const double d = 0.25 * dh_drho(k, L, r);
        double* o = out + ((long)e * 14 + r) * 4;
#pragma unroll
        for (int b = 0; b < 4; ++b) o[b] = d;
    } else {  // PA_WRT_C_PREV: -I except the plastic-multiplier row
        double* o = out + ((long)e * 14 + r) * 14;
#pragma unroll
        for (int j = 0; j < 14; ++j) o[j] = (j == r && r != 1) ? -1.0 : 0.0;
    }
}

// Global cell residual into vcell (mode 0) or one row of a global workset
// Jacobian (mode 1; thread per (cell, row)); wrt: PA_WRT_U, PA_WRT_C, PA_WRT_Z, PA_WRT_PI.
__global__ __launch_bounds__(kB) void k_global(int C, int mode, int wrt, KP k, const int* __restrict__ conn,
                                               const double4* __restrict__ X, const double* __restrict__ u,
                                               const double* __restrict__ c_in, const double* __restrict__ pi,
                                               const double* __restrict__ z, double* __restrict__ out) {
    const long t = (long)blockIdx.x * kB + threadIdx.x;
    const int e = (int)(t / 16), r = (int)(t % 16);
    if (e >= C) return;
    const int a = r / 4, i = r % 4;
    CellIn Ci;
    int nn[4];
    load_cell(conn, e, X, u, k.dpn, z, Ci, nn);
    double c[14];
    load14(c_in, e, c);
    double pig[3];
    interp_pi(pi, nn, pig);
    Glob G;
    global_ctx(k, Ci, c, pig, G);
    double ga[3];
    pick_node(Ci.g, a, ga);
    if (mode == 0) {
        out[(long)e * 16 + r] = global_residual_row(k, Ci, G, ga, i);
        return;
    }
    if (wrt == PA_WRT_U) {
        double row[16];
        dr_du_row(k, Ci, G, ga, i, row);
        double* o = out + ((long)e * 16 + r) * 16;
#pragma unroll
        for (int j = 0; j < 16; ++j) o[j] = row[j];
    } else if (wrt == PA_WRT_C) {
        double dep[6];
        dr_dep_row(k, Ci, G, ga, i, dep);
        double* o = out + ((long)e * 16 + r) * 14;
        o[0] = 0.0;
        o[1] = 0.0;
#pragma unroll
        for (int w = 0; w < 6; ++w) {
            o[2 + w] = dep[w];
            o[8 + w] = 0.0;
        }
    } else if (wrt == PA_WRT_Z) {
        const double d = 0.25 * dr_drho_row(k, Ci, G, ga, i);
        double* o = out + ((long)e * 16 + r) * 4;
#pragma unroll
        for (int b = 0; b < 4; ++b) o[b] = d;
    } else {  // PA_WRT_PI: only the pressure rows see the projected gradient
        double* o = out + ((long)e * 16 + r) * 12;
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int d = 0; d < 3; ++d) o[b * 3 + d] = (i == 3 && k.mixed) ? Ci.V * k.ps * G.tau * 0.25 * ga[d] : 0.0;
    }
}

// ---------------------------------------------------------------------------
// 16-lane-group kernels: the local 14x14 system, one lane per row
// ---------------------------------------------------------------------------
constexpr int G16 = 16;
constexpr int CPB = kB / G16;  // cells per block

// Gauss-Jordan elimination without pivoting across the lanes of a 16-lane
// group: lane r holds row r of [A | B] (NC columns, A = the first 14); lanes
// 14 and 15 hold zero rows.  Afterwards lane r holds row r of A^-1 B.
template <int NC>
__device__ inline void gauss_jordan14(double (&A)[NC], int lane) {
#pragma unroll
    for (int k = 0; k < 14; ++k) {
        const double inv = 1.0 / __shfl(A[k], k, G16);
        const double f = A[k];
        const bool me = lane == k;
#pragma unroll
        for (int j = k + 1; j < NC; ++j) {
            const double pj = __shfl(A[j], k, G16) * inv;
            A[j] = me ? pj : A[j] - f * pj;
        }
        A[k] = me ? 1.0 : 0.0;
    }
}

__device__ inline double group_sum16(double v) {
#pragma unroll
    for (int o = G16 / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G16);
    return v;
}

__device__ inline int sel_int4(const int nn[4], int a) {
    return (a == 0) * nn[0] + (a == 1) * nn[1] + (a == 2) * nn[2] + (a == 3) * nn[3];
}

// Condensed cell system (NewtonRaphsonSolver::assembleTangentMatrix and
// ::assembleResidual per cell): kcell[e] = K_t (16x16, block layout), vcell[e] = R_t.
__global__ __launch_bounds__(kB) void k_condensed(int C, KP k, const int* __restrict__ conn,
                                                  const double4* __restrict__ X, const double* __restrict__ u,
                                                  const double* __restrict__ c_in, const double* __restrict__ cp_in,
                                                  const double* __restrict__ pi, const double* __restrict__ z,
                                                  double* __restrict__ kcell, double* __restrict__ vcell) {
    __shared__ double stage[CPB][272];
    const int grp = threadIdx.x / G16, lane = threadIdx.x % G16;
    const int e0 = blockIdx.x * CPB + grp;
    const bool valid = e0 < C;
    const int e = valid ? e0 : C - 1;
    CellIn Ci;
    int nn[4];
    load_cell(conn, e, X, u, k.dpn, z, Ci, nn);
    double c[14], cp[14];
    load14(c_in, e, c);
    load14(cp_in, e, cp);
    Loc L;
    local_ctx(k, Ci, c, L);
    // lane r: row r of [dH/dc | dH/du (12 displacement columns) | H]
    double A[27];
    {
        double row[14], q[6];
        dh_row(L, lane, row, q);
#pragma unroll
        for (int j = 0; j < 14; ++j) A[j] = row[j];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            double v[3];
            bt(Ci.g[b], q, v);
            A[14 + b * 3 + 0] = v[0];
            A[14 + b * 3 + 1] = v[1];
            A[14 + b * 3 + 2] = v[2];
        }
        A[26] = lane < 14 ? local_residual_row(L, c, cp, lane) : 0.0;
    }
    gauss_jordan14<27>(A, lane);
    // lane r = (a, i): row r of K_t and R_t; the Schur term uses the plastic
    // strain rows (lanes 2..7) of (dH/dc)^-1 [dH/du | H]
    const int a = lane / 4, i = lane % 4;
    double pig[3];
    interp_pi(pi, nn, pig);
    Glob Gl;
    global_ctx(k, Ci, c, pig, Gl);
    double ga[3];
    pick_node(Ci.g, a, ga);
    double K[16];
    dr_du_row(k, Ci, Gl, ga, i, K);
    double Rt = global_residual_row(k, Ci, Gl, ga, i);
    double dep[6];
    dr_dep_row(k, Ci, Gl, ga, i, dep);
#pragma unroll
    for (int col = 0; col < 13; ++col) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < 6; ++w) s += dep[w] * __shfl(A[14 + col], 2 + w, G16);
        if (col < 12)
            K[(col / 3) * 4 + col % 3] -= s;
        else
            Rt -= s;
    }
    double* sg = stage[grp];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int j = 0; j < 4; ++j) sg[((a * 4 + b) * 4 + i) * 4 + j] = K[b * 4 + j];
    sg[256 + lane] = Rt;
    __syncthreads();
    if (valid) {
        double* kd = kcell + (long)e0 * 256;
#pragma unroll
        for (int t = 0; t < 16; ++t) kd[t * G16 + lane] = sg[t * G16 + lane];
        vcell[(long)e0 * 16 + lane] = sg[256 + lane];
    }
}

// Rows of dH/dc of the cell into LDS (lane r writes row r), then each lane
// reads its column: lane l gets row l of (dH/dc)^T.
__device__ inline void transposed_rows(const Loc& L, int lane, double (*sJ)[14], double A[15], double q[6]) {
    double row[14];
    dh_row(L, lane, row, q);
    if (lane < 14)
#pragma unroll
        for (int j = 0; j < 14; ++j) sJ[lane][j] = row[j];
    __syncthreads();
    const int col = lane < 14 ? lane : 0;
#pragma unroll
    for (int r = 0; r < 14; ++r) {
        const double v = sJ[r][col];
        A[r] = lane < 14 ? v : 0.0;
    }
}

// Local adjoint right-hand side of step k (PathDependentAdjointSolver::
// computeLocalAdjointRHS, src/PathDependentAdjointSolver.hpp:277-351) and the
// cell vectors of the global adjoint right-hand side (:394-469):
//   b_local = gc_k + [gcp_{k+1} + (dH_{k+1}/dc_k)^T mu_{k+1}],
//   vcell   = wu_k + [wu_{k+1} + DpDn^T gamma_{k+1}] - (dH_k/du_k)^T (dH_k/dc_k)^-T b_local.
// (dH/dc_prev = -I except the plastic-multiplier row, so its transpose times
// mu is -mu with that component dropped.)
__global__ __launch_bounds__(kB) void k_adj_local_rhs(int C, KP k, const int* __restrict__ conn,
                                                      const double4* __restrict__ X, const double* __restrict__ u,
                                                      const double* __restrict__ c_in,
                                                      const double* __restrict__ cp_in, const double* __restrict__ z,
                                                      const double* __restrict__ gc,
                                                      const double* __restrict__ gcp_next,
                                                      const double* __restrict__ mu_next,
                                                      const double* __restrict__ wu,
                                                      const double* __restrict__ wu_next,
                                                      const double* __restrict__ gamma_next,
                                                      double* __restrict__ blocal, double* __restrict__ vcell) {
    __shared__ double sJ[CPB][14][14];
    const int grp = threadIdx.x / G16, lane = threadIdx.x % G16;
    const int e0 = blockIdx.x * CPB + grp;
    const bool valid = e0 < C;
    const int e = valid ? e0 : C - 1;
    CellIn Ci;
    int nn[4];
    load_cell(conn, e, X, u, k.dpn, z, Ci, nn);
    double c[14];
    load14(c_in, e, c);
    Loc L;
    local_ctx(k, Ci, c, L);
    double A[15], q[6];
    transposed_rows(L, lane, sJ[grp], A, q);
    double b = 0.0;
    if (lane < 14) {
        b = gc[(long)e * 14 + lane];
        if (gcp_next) b += gcp_next[(long)e * 14 + lane] - (lane == 1 ? 0.0 : mu_next[(long)e * 14 + lane]);
        if (valid) blocal[(long)e0 * 14 + lane] = b;
    }
    A[14] = b;
    gauss_jordan14<15>(A, lane);
    const double t = A[14];  // component `lane` of (dH/dc)^-T b_local
    double Q[6];
#pragma unroll
    for (int v = 0; v < 6; ++v) Q[v] = group_sum16(t * q[v]);
    const int a = lane / 4, i = lane % 4;
    double ga[3];
    pick_node(Ci.g, a, ga);
    double F = 0.0;
    if (i < 3) {
        double v3[3];
        bt(ga, Q, v3);
        F = sel<3>(v3, i);
    }
    double out = wu[(long)e * 16 + lane] - F;
    if (wu_next) {
        out += wu_next[(long)e * 16 + lane];
        if (i == 3 && k.mixed) {
            // DpDn^T gamma: dQ_{a',d}/dp_b = -1/4 V w ps dN_b/dx_d (src/PressureGradientProjectionResidual.hpp:136-160)
            double w = 1.0, dw;
            if (k.proj) msimp(Ci.rho, k.pj_e, k.pj_m, w, dw);
            double sg = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double s = 0.0;
#pragma unroll
                for (int n = 0; n < 4; ++n) s += gamma_next[(long)nn[n] * 3 + d];
                sg += ga[d] * s;
            }
            out += -0.25 * Ci.V * w * k.ps * sg;
        }
    }
    if (valid) vcell[(long)e0 * 16 + lane] = out;
}

// Local adjoint of step k and the control contributions
// (PathDependentAdjointSolver::updateLocalAdjointVars, :824-880, and
// addPDEpartialDerivativeZ, :517-551):
//   mu_k = -(dH/dc)^-T (b_local + (dR/dc)^T lambda_k),
//   zc_out[e] = 1/4 (dR/drho . lambda + dH/drho . mu) + zc[e],
//   vcell[e*16 + d] = sum_a lambda_{a,3} V ps tau / 4 dN_a/dx_d   ((dR/dPi)^T lambda per node of the cell).
__global__ __launch_bounds__(kB) void k_adj_local_update(int C, KP k, const int* __restrict__ conn,
                                                         const double4* __restrict__ X,
                                                         const double* __restrict__ u,
                                                         const double* __restrict__ c_in,
                                                         const double* __restrict__ pi,
                                                         const double* __restrict__ z,
                                                         const double* __restrict__ lambda,
                                                         const double* __restrict__ blocal,
                                                         const double* __restrict__ zc, double* __restrict__ mu,
                                                         double* __restrict__ zc_out, double* __restrict__ vcell) {
    __shared__ double sJ[CPB][14][14];
    const int grp = threadIdx.x / G16, lane = threadIdx.x % G16;
    const int e0 = blockIdx.x * CPB + grp;
    const bool valid = e0 < C;
    const int e = valid ? e0 : C - 1;
    CellIn Ci;
    int nn[4];
    load_cell(conn, e, X, u, k.dpn, z, Ci, nn);
    double c[14];
    load14(c_in, e, c);
    Loc L;
    local_ctx(k, Ci, c, L);
    double pig[3];
    interp_pi(pi, nn, pig);
    Glob Gl;
    global_ctx(k, Ci, c, pig, Gl);
    // global adjoint pieces, lane r = (a, i)
    const int a = lane / 4, i = lane % 4;
    double ga[3];
    pick_node(Ci.g, a, ga);
    const double lam = i < k.dpn ? lambda[(long)sel_int4(nn, a) * k.dpn + i] : 0.0;
    double dep[6];
    dr_dep_row(k, Ci, Gl, ga, i, dep);
    double drl[6];
#pragma unroll
    for (int w = 0; w < 6; ++w) drl[w] = group_sum16(lam * dep[w]);
    const double zR = group_sum16(lam * dr_drho_row(k, Ci, Gl, ga, i));
    double pv[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) pv[d] = group_sum16(i == 3 ? lam * Ci.V * k.ps * Gl.tau * 0.25 * ga[d] : 0.0);
    // local adjoint
    double A[15], q[6];
    transposed_rows(L, lane, sJ[grp], A, q);
    double b = 0.0;
    if (lane < 14) {
        b = blocal[(long)e * 14 + lane];
        if (lane >= 2 && lane < 8) b += sel<6>(drl, lane - 2);
    }
    A[14] = b;
    gauss_jordan14<15>(A, lane);
    const double m = lane < 14 ? -A[14] : 0.0;
    if (valid && lane < 14) mu[(long)e0 * 14 + lane] = m;
    const double zH = group_sum16(m * dh_drho(k, L, lane));
    if (valid) {
        if (lane == 0) zc_out[e0] = 0.25 * (zR + zH) + zc[e0];
        if (lane < 3) vcell[(long)e0 * 16 + lane] = sel<3>(pv, lane);
    }
}

// ---------------------------------------------------------------------------
// node-ordered gathers (fixed summation order: incident cells ascending)
// ---------------------------------------------------------------------------
// Block-CRS values (D x D blocks of the 4x4 cell blocks: D = 3 drops the
// pressure row and column) and the residual of each graph row; lane = 4x4 block entry.
template <int D>
__global__ __launch_bounds__(kB) void k_gather_matrix(int N, const int* __restrict__ rowmap,
                                                      const int* __restrict__ n2c_off,
                                                      const int2* __restrict__ pairs,
                                                      const double* __restrict__ kcell,
                                                      const double* __restrict__ vcell, int transpose, int maxlen,
                                                      int rpb, double* __restrict__ values, double* __restrict__ R) {
    extern __shared__ double acc[];
    const int grp = threadIdx.x / G16, lane = threadIdx.x % G16;
    const int row = blockIdx.x * rpb + grp;
    if (grp >= rpb || row >= N) return;  // whole 16-lane groups exit together
    double* my = acc + (long)grp * maxlen * 16;
    const int len = rowmap[row + 1] - rowmap[row];
    for (int s = 0; s < len; ++s) my[s * 16 + lane] = 0.0;
    const int i = lane / 4, j = lane % 4;
    const int tsrc = transpose ? j * 4 + i : lane;
    double racc = 0.0;
    for (int p = n2c_off[row]; p < n2c_off[row + 1]; ++p) {
        const int2 pr = pairs[p];
        const long e = pr.x >> 2;
        const int a = pr.x & 3;
        const double* ke = kcell + e * 256;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int slot = (pr.y >> (8 * b)) & 0xff;
            const int blk = transpose ? b * 4 + a : a * 4 + b;
            my[slot * 16 + lane] += ke[blk * 16 + tsrc];
        }
        if (lane < 4) racc += vcell[e * 16 + a * 4 + lane];
    }
    double* out = values + (long)rowmap[row] * (D * D);
    if (i < D && j < D)
        for (int s = 0; s < len; ++s) out[s * (D * D) + i * D + j] = my[s * 16 + lane];
    if (R && lane < D) R[(long)row * D + lane] = racc;
}

// out[node*D + d] = scale * sum over incident cells of vcell[cell*16 + a*4 + d], d < D.
__global__ __launch_bounds__(kB) void k_gather_vcell(int N, int D, const int* __restrict__ n2c_off,
                                                     const int2* __restrict__ pairs,
                                                     const double* __restrict__ vcell, double scale,
                                                     double* __restrict__ out) {
    const long t = (long)blockIdx.x * kB + threadIdx.x;
    if (t >= (long)N * D) return;
    const int node = (int)(t / D), d = (int)(t % D);
    double s = 0.0;
    for (int p = n2c_off[node]; p < n2c_off[node + 1]; ++p) {
        const int2 pr = pairs[p];
        s += vcell[(long)(pr.x >> 2) * 16 + (pr.x & 3) * 4 + d];
    }
    out[t] = scale * s;
}

// out[node*W + k] (+)= scale * sum over incident cells of ws[cell*stride + k].
__global__ __launch_bounds__(kB) void k_gather_uniform(int N, int W, int stride, const int* __restrict__ n2c_off,
                                                       const int2* __restrict__ pairs,
                                                       const double* __restrict__ ws, double scale, int accumulate,
                                                       double* __restrict__ out) {
    const long t = (long)blockIdx.x * kB + threadIdx.x;
    if (t >= (long)N * W) return;
    const int node = (int)(t / W), kk = (int)(t % W);
    double s = 0.0;
    for (int p = n2c_off[node]; p < n2c_off[node + 1]; ++p) s += ws[(long)(pairs[p].x >> 2) * stride + kk];
    out[t] = accumulate ? out[t] + scale * s : scale * s;
}

// Projection sources per cell: vcell[e*16 + d] = 1/4 V w ps grad p (d < 3), vcell[e*16 + 3] = 1/4 V w.
__global__ __launch_bounds__(kB) void k_proj_cells(int C, KP k, const int* __restrict__ conn,
                                                   const double4* __restrict__ X, const double* __restrict__ u,
                                                   const double* __restrict__ z, double* __restrict__ vcell) {
    const int e = blockIdx.x * kB + threadIdx.x;
    if (e >= C) return;
    CellIn Ci;
    int nn[4];
    load_cell(conn, e, X, u, k.dpn, z, Ci, nn);
    double w = 1.0, dw;
    if (k.proj) msimp(Ci.rho, k.pj_e, k.pj_m, w, dw);
    double* o = vcell + (long)e * 16;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double s = 0.0;
#pragma unroll
        for (int n = 0; n < 4; ++n) s += k.ps * Ci.p[n] * Ci.g[n][d];
        o[d] = 0.25 * (s * w) * Ci.V;
    }
    o[3] = 0.25 * w * Ci.V;
}

// pi[node*3 + d] = acc[node*4 + d] / acc[node*4 + 3]; rowsum[node] = acc[node*4 + 3].
__global__ void k_proj_finish(int N, const double* __restrict__ acc, double* __restrict__ pi,
                              double* __restrict__ rowsum) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const double m = acc[(long)n * 4 + 3];
    if (pi)
#pragma unroll
        for (int d = 0; d < 3; ++d) pi[(long)n * 3 + d] = acc[(long)n * 4 + d] / m;
    if (rowsum) rowsum[n] = m;
}

__global__ void k_divide_rowsum(int N, const double* __restrict__ rowsum, double* __restrict__ g) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (long)N * 3) g[t] = g[t] / rowsum[t / 3];
}

// Projection-penalty control term (addPDEpartialDerivativeZ, :534-540, with
// the step's own pressure as the reference passes it):
// zc_out[e] += 1/16 V dw (Pi_gp - ps grad p) . sum_a gamma_a.
__global__ __launch_bounds__(kB) void k_proj_control(int C, KP k, const int* __restrict__ conn,
                                                     const double4* __restrict__ X, const double* __restrict__ u,
                                                     const double* __restrict__ pi,
                                                     const double* __restrict__ gamma, const double* __restrict__ z,
                                                     double* __restrict__ zc_out) {
    const int e = blockIdx.x * kB + threadIdx.x;
    if (e >= C) return;
    CellIn Ci;
    int nn[4];
    load_cell(conn, e, X, u, k.dpn, z, Ci, nn);
    double w, dw;
    msimp(Ci.rho, k.pj_e, k.pj_m, w, dw);
    double pig[3];
    interp_pi(pi, nn, pig);
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double gp = 0.0, gs = 0.0;
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            gp += Ci.p[n] * Ci.g[n][d];
            gs += gamma[(long)nn[n] * 3 + d];
        }
        s += (pig[d] - k.ps * gp) * gs;
    }
    zc_out[e] += 0.25 * 0.25 * Ci.V * dw * s;
}

// ---------------------------------------------------------------------------
// plastic work (src/PlasticWorkCriterion.hpp:163-265) and its partials
// ---------------------------------------------------------------------------
struct Work {
    double V, pw, dpw;
    double delta[6];      // eps_p - eps_p,prev
    double sigsum[6];     // sigma_c + sigma_p (Cauchy, pressure on the normals)
    double s0sum[6];      // unpenalized deviatoric stresses summed
};

__device__ inline void work_ctx(const KP& k, const CellIn& Cc, const double epsp_prev[6], const double pp[4],
                                const double c[14], const double cp[14], Work& W) {
    W.V = Cc.V;
    msimp(Cc.rho, k.pw_e, k.pw_m, W.pw, W.dpw);
    const double mu = W.pw * k.mu;
    double eec[6], eep[6], sc[6], sp[6], s0c[6], s0p[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        eec[i] = Cc.eps[i] - c[2 + i];
        eep[i] = epsp_prev[i] - cp[2 + i];
        W.delta[i] = c[2 + i] - cp[2 + i];
    }
    dev(mu, eep, sp);
    dev(mu, eec, sc);
    dev(k.mu, eep, s0p);
    dev(k.mu, eec, s0c);
    double Pp = 0.0, Pc = 0.0;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        Pp += 0.25 * pp[n];
        Pc += 0.25 * Cc.p[n];
    }
    Pp = Pp * k.ps;
    Pc = Pc * k.ps;
    // displacement formulation: the volumetric stress kappa~_w tr(eps_e) takes the pressure's place
    const double trc = (eec[0] + eec[1]) + eec[2], trp = (eep[0] + eep[1]) + eep[2];
    if (!k.mixed) {
        Pp = (W.pw * k.kappa) * trp;
        Pc = (W.pw * k.kappa) * trc;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double sigc = i < 3 ? sc[i] + Pc : sc[i];
        const double sigp = i < 3 ? sp[i] + Pp : sp[i];
        W.sigsum[i] = sigc + sigp;
        W.s0sum[i] = s0c[i] + s0p[i] + ((!k.mixed && i < 3) ? k.kappa * (trc + trp) : 0.0);
    }
}

__device__ inline void load_work_cell(const KP& k, const int* __restrict__ conn, int e,
                                      const double4* __restrict__ X, const double* __restrict__ u,
                                      const double* __restrict__ up, const double* __restrict__ c_in,
                                      const double* __restrict__ cp_in, const double* __restrict__ z, CellIn& Ci,
                                      double c[14], double cp[14], Work& W) {
    int nn[4];
    load_cell(conn, e, X, u, k.dpn, z, Ci, nn);
    double epsp[6], pp[4];
    strain_of(Ci.g, up, k.dpn, nn, epsp, pp);
    load14(c_in, e, c);
    load14(cp_in, e, cp);
    work_ctx(k, Ci, epsp, pp, c, cp, W);
}

__device__ inline double block_sum(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / 64); ++i) t += sh[i];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(kB) void k_work_cells(int C, KP k, const int* __restrict__ conn,
                                                   const double4* __restrict__ X, const double* __restrict__ u,
                                                   const double* __restrict__ up, const double* __restrict__ c_in,
                                                   const double* __restrict__ cp_in, const double* __restrict__ z,
                                                   double* __restrict__ part) {
    __shared__ double sh[kB / 64];
    const int e = blockIdx.x * kB + threadIdx.x;
    double f = 0.0;
    if (e < C) {
        CellIn Ci;
        double c[14], cp[14];
        Work W;
        load_work_cell(k, conn, e, X, u, up, c_in, cp_in, z, Ci, c, cp, W);
#pragma unroll
        for (int i = 0; i < 6; ++i) f += (0.5 * W.sigsum[i]) * W.delta[i];
        f *= W.V;
    }
    const double t = block_sum(f, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_reduce_to(const double* __restrict__ part, int n, double* __restrict__ dst) {
    __shared__ double sh[1024 / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) *dst = t;
}

// Partials of mult * f_k (thread per cell): wu = d/du_k (equal to d/du_{k-1}),
// gc = d/dc_k, gcp = d/dc_{k-1}, zc = 1/4 d/drho.
__global__ __launch_bounds__(kB) void k_work_partials(int C, KP k, double mult, const int* __restrict__ conn,
                                                      const double4* __restrict__ X, const double* __restrict__ u,
                                                      const double* __restrict__ up,
                                                      const double* __restrict__ c_in,
                                                      const double* __restrict__ cp_in,
                                                      const double* __restrict__ z, double* __restrict__ wu,
                                                      double* __restrict__ gc, double* __restrict__ gcp,
                                                      double* __restrict__ zc) {
    const int e = blockIdx.x * kB + threadIdx.x;
    if (e >= C) return;
    CellIn Ci;
    double c[14], cp[14];
    Work W;
    load_work_cell(k, conn, e, X, u, up, c_in, cp_in, z, Ci, c, cp, W);
    // m = (d sigma / d eps_e)^T delta (the elastic operator is symmetric)
    double m[6];
    dev(W.pw * k.mu, W.delta, m);
    const double hv = 0.5 * W.V * mult;
    const double trd = W.delta[0] + W.delta[1] + W.delta[2];
    if (!k.mixed)
#pragma unroll
        for (int i = 0; i < 3; ++i) m[i] += (W.pw * k.kappa) * trd;
    double* o = wu + (long)e * 16;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        double v[3];
        bt(Ci.g[b], m, v);
        o[b * 4 + 0] = hv * v[0];
        o[b * 4 + 1] = hv * v[1];
        o[b * 4 + 2] = hv * v[2];
        o[b * 4 + 3] = k.mixed ? hv * k.ps * 0.25 * trd : 0.0;
    }
    double* oc = gc + (long)e * 14;
    double* op = gcp + (long)e * 14;
    oc[0] = oc[1] = op[0] = op[1] = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        oc[2 + i] = W.V * mult * (0.5 * W.sigsum[i] - 0.5 * m[i]);
        op[2 + i] = W.V * mult * (-0.5 * W.sigsum[i] - 0.5 * m[i]);
        oc[8 + i] = 0.0;
        op[8 + i] = 0.0;
    }
    double sd = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) sd += W.s0sum[i] * W.delta[i];
    zc[e] = 0.25 * hv * W.dpw * sd;
}

int blocks(long n) { return (int)((n + kB - 1) / kB); }

}  // namespace

KP constants(const pa_j2_params& p) {
    KP k{};
    k.mu = p.youngs / (2.0 * (1.0 + p.poisson));
    k.kappa = p.youngs / (3.0 * (1.0 - 2.0 * p.poisson));
    k.hiso = p.hardening_isotropic;
    k.hkin = p.hardening_kinematic;
    k.sy = p.initial_yield_stress;
    k.pe_e = p.elastic_penalty_exponent;
    k.pe_m = p.elastic_min_ersatz;
    k.pp_e = p.plastic_penalty_exponent;
    k.pp_m = p.plastic_min_ersatz;
    k.pg_e = p.global_penalty_exponent;
    k.pg_m = p.global_min_ersatz;
    k.ps = p.pressure_scaling;
    k.proj = p.projection_penalty;
    k.pj_e = p.projection_penalty_exponent;
    k.pj_m = p.projection_min_ersatz;
    k.pw_e = p.work_penalty_exponent;
    k.pw_m = p.work_min_ersatz;
    k.mixed = p.formulation == PA_J2_MIXED ? 1 : 0;
    k.dpn = k.mixed ? 4 : 3;
    return k;
}

namespace {
const double4* coords4(const pa_plasticity* p) { return reinterpret_cast<const double4*>(p->mesh->d_coords4.p); }
}  // namespace

void update_local_state(pa_plasticity* p, const double* u, const double* cp, const double* z, double* c,
                        hipStream_t st) {
    const pa_mesh* m = p->mesh;
    k_update<<<blocks(m->C), kB, 0, st>>>(m->C, constants(p->prm), m->d_conn.p, coords4(p), u, cp, z, c);
    check();
}

void gather_vcell(pa_plasticity* p, double scale, double* out, hipStream_t st) {
    const pa_mesh* m = p->mesh;
    k_gather_vcell<<<blocks((long)p->dpn * m->N), kB, 0, st>>>(m->N, p->dpn, m->d_n2c_off.p, m->d_pairs.p,
                                                               p->vcell.p, scale, out);
    check();
}

void gather_uniform(pa_plasticity* p, int W, const double* ws, int stride, double scale, bool accumulate,
                    double* out, hipStream_t st) {
    const pa_mesh* m = p->mesh;
    k_gather_uniform<<<blocks((long)W * m->N), kB, 0, st>>>(m->N, W, stride, m->d_n2c_off.p, m->d_pairs.p, ws, scale,
                                                            accumulate ? 1 : 0, out);
    check();
}

void gather_uniform(pa_plasticity* p, int W, const double* ws, double scale, bool accumulate, double* out,
                    hipStream_t st) {
    gather_uniform(p, W, ws, W, scale, accumulate, out, st);
}

void condensed_system(pa_plasticity* p, const double* u, const double* c, const double* cp, const double* pi,
                      const double* z, double* values, double* R, bool transpose, hipStream_t st) {
    const pa_mesh* m = p->mesh;
    k_condensed<<<(m->C + CPB - 1) / CPB, kB, 0, st>>>(m->C, constants(p->prm), m->d_conn.p, coords4(p), u, c, cp,
                                                       pi, z, p->kcell.p, p->vcell.p);
    check();
    // rows per workgroup: as many 16-lane groups as 64 KB of row accumulators allow
    const int maxlen = m->max_row_len;
    const int rpb = std::min(CPB, (int)(65536 / ((size_t)maxlen * 16 * sizeof(double))));
    const size_t lds = (size_t)rpb * maxlen * 16 * sizeof(double);
    auto* gather = p->dpn == 4 ? k_gather_matrix<4> : k_gather_matrix<3>;
    gather<<<(m->N + rpb - 1) / rpb, kB, lds, st>>>(m->N, m->d_rowmap.p, m->d_n2c_off.p, m->d_pairs.p, p->kcell.p,
                                                    p->vcell.p, transpose ? 1 : 0, maxlen, rpb, values,
                                                    transpose ? nullptr : R);
    check();
}

void projection(pa_plasticity* p, const double* u, const double* z, double* pi, hipStream_t st) {
    const pa_mesh* m = p->mesh;
    k_proj_cells<<<blocks(m->C), kB, 0, st>>>(m->C, constants(p->prm), m->d_conn.p, coords4(p), u, z, p->vcell.p);
    check();
    gather_uniform(p, 4, p->vcell.p, 16, 1.0, false, p->tmp4.p, st);
    k_proj_finish<<<blocks(m->N), kB, 0, st>>>(m->N, p->tmp4.p, pi, p->rowsum.p);
    check();
}

void divide_rowsum(pa_plasticity* p, double* gamma, const double* z, hipStream_t st) {
    const pa_mesh* m = p->mesh;
    // the lumped mass depends on z only; u = 0 gives the sources zero
    k_proj_cells<<<blocks(m->C), kB, 0, st>>>(m->C, constants(p->prm), m->d_conn.p, coords4(p), p->zero4.p, z,
                                              p->vcell.p);
    check();
    gather_uniform(p, 4, p->vcell.p, 16, 1.0, false, p->tmp4.p, st);
    k_proj_finish<<<blocks(m->N), kB, 0, st>>>(m->N, p->tmp4.p, nullptr, p->rowsum.p);
    check();
    k_divide_rowsum<<<blocks(3L * m->N), kB, 0, st>>>(m->N, p->rowsum.p, gamma);
    check();
}

void plastic_work(pa_plasticity* p, const double* u, const double* up, const double* c, const double* cp,
                  const double* z, double* dst, hipStream_t st) {
    const pa_mesh* m = p->mesh;
    const int nb = blocks(m->C);
    k_work_cells<<<nb, kB, 0, st>>>(m->C, constants(p->prm), m->d_conn.p, coords4(p), u, up, c, cp, z, p->part.p);
    check();
    k_reduce_to<<<1, 1024, 0, st>>>(p->part.p, nb, dst);
    check();
}

void work_partials(pa_plasticity* p, const double* u, const double* up, const double* c, const double* cp,
                   const double* z, double mult, const AdjointWork& w, hipStream_t st) {
    const pa_mesh* m = p->mesh;
    k_work_partials<<<blocks(m->C), kB, 0, st>>>(m->C, constants(p->prm), mult, m->d_conn.p, coords4(p), u, up, c,
                                                 cp, z, w.wu, w.gc, w.gcp, w.zc);
    check();
}

void adjoint_local_rhs(pa_plasticity* p, const double* u, const double* c, const double* cp, const double* z,
                       const double* gc, const double* gcp_next, const double* mu_next, const double* wu,
                       const double* wu_next, const double* gamma_next, double* blocal, hipStream_t st) {
    (void)cp;
    const pa_mesh* m = p->mesh;
    k_adj_local_rhs<<<(m->C + CPB - 1) / CPB, kB, 0, st>>>(m->C, constants(p->prm), m->d_conn.p, coords4(p), u, c,
                                                           cp, z, gc, gcp_next, mu_next, wu, wu_next, gamma_next,
                                                           blocal, p->vcell.p);
    check();
}

void adjoint_local_update(pa_plasticity* p, const double* u, const double* c, const double* cp, const double* pi,
                          const double* z, const double* lambda, const double* blocal, const double* zc,
                          double* mu, double* zc_out, hipStream_t st) {
    (void)cp;
    const pa_mesh* m = p->mesh;
    k_adj_local_update<<<(m->C + CPB - 1) / CPB, kB, 0, st>>>(m->C, constants(p->prm), m->d_conn.p, coords4(p), u,
                                                              c, pi, z, lambda, blocal, zc, mu, zc_out, p->vcell.p);
    check();
}

void projection_control_term(pa_plasticity* p, const double* u, const double* pi, const double* gamma,
                             const double* z, double* zc_out, hipStream_t st) {
    const pa_mesh* m = p->mesh;
    k_proj_control<<<blocks(m->C), kB, 0, st>>>(m->C, constants(p->prm), m->d_conn.p, coords4(p), u, pi, gamma, z,
                                                zc_out);
    check();
}

}  // namespace j2
}  // namespace pa

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using pa::guarded;

extern "C" {

void pa_j2_params_default(pa_j2_params* p) {
    if (!p) return;
    *p = pa_j2_params{};
    p->elastic_penalty_exponent = 1.0;
    p->elastic_min_ersatz = 1e-9;
    p->plastic_penalty_exponent = 0.5;
    p->plastic_min_ersatz = 1e-4;
    p->global_penalty_exponent = 1.0;
    p->global_min_ersatz = 1e-9;
    p->pressure_scaling = 1.0;
    p->projection_penalty = 0;
    p->projection_penalty_exponent = 3.0;
    p->projection_min_ersatz = 1e-9;
    p->work_penalty_exponent = 3.0;
    p->work_min_ersatz = 1e-9;
    p->formulation = PA_J2_MIXED;
}

int pa_plasticity_create(const pa_mesh* mesh, const pa_j2_params* params, pa_plasticity** out) {
    return guarded([&] {
        PA_REQUIRE(mesh && params && out, PA_ERR_ARG, "pa_plasticity_create: null argument");
        PA_REQUIRE(params->youngs > 0.0 && params->poisson > -1.0 && params->poisson < 0.5, PA_ERR_ARG,
                   "pa_plasticity_create: elastic constants out of range");
        PA_REQUIRE(params->formulation == PA_J2_MIXED || params->formulation == PA_J2_DISPLACEMENT, PA_ERR_ARG,
                   "pa_plasticity_create: formulation must be PA_J2_MIXED or PA_J2_DISPLACEMENT");
        PA_REQUIRE(mesh->max_row_len <= 64, PA_ERR_LIMIT,
                   "pa_plasticity_create: graph rows above 64 blocks are not supported by the gather kernel");
        auto* p = new pa_plasticity();
        try {
            p->mesh = mesh;
            p->prm = *params;
            p->dpn = params->formulation == PA_J2_MIXED ? 4 : 3;
            p->kcell.alloc(256L * mesh->C);
            p->vcell.alloc(16L * mesh->C);
            p->part.alloc((size_t)std::max(1L, (mesh->C + 255L) / 256));
            p->scal.alloc(8);
            p->rowsum.alloc(mesh->N);
            p->tmp4.alloc(4L * mesh->N);
            p->zero4.alloc(4L * mesh->N);
            PA_HIP(hipMemset(p->zero4.p, 0, 4L * mesh->N * sizeof(double)));
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

int pa_plasticity_destroy(pa_plasticity* p) {
    return guarded([&] { delete p; });
}

int pa_j2_update_local_state(pa_plasticity* p, const double* u, const double* cp, const double* z, double* c,
                             void* stream) {
    return guarded([&] {
        PA_REQUIRE(p && u && cp && z && c, PA_ERR_ARG, "pa_j2_update_local_state: null argument");
        pa::j2::update_local_state(p, u, cp, z, c, (hipStream_t)stream);
    });
}

int pa_j2_local_value(pa_plasticity* p, const double* u, const double* c, const double* cp, const double* z,
                      double* H, void* stream) {
    return guarded([&] {
        PA_REQUIRE(p && u && c && cp && z && H, PA_ERR_ARG, "pa_j2_local_value: null argument");
        const pa_mesh* m = p->mesh;
        pa::j2::k_local<<<pa::j2::blocks(14L * m->C), pa::j2::kB, 0, (hipStream_t)stream>>>(
            m->C, 0, 0, pa::j2::constants(p->prm), m->d_conn.p, pa::j2::coords4(p), u, c, cp, z, H);
        pa::j2::check();
    });
}

int pa_j2_local_gradient(pa_plasticity* p, int wrt, const double* u, const double* c, const double* cp,
                         const double* z, double* out, void* stream) {
    return guarded([&] {
        PA_REQUIRE(p && u && c && cp && z && out, PA_ERR_ARG, "pa_j2_local_gradient: null argument");
        PA_REQUIRE(wrt == PA_WRT_C || wrt == PA_WRT_U || wrt == PA_WRT_Z || wrt == PA_WRT_C_PREV, PA_ERR_ARG,
                   "pa_j2_local_gradient: wrt must be PA_WRT_C, PA_WRT_U, PA_WRT_Z or PA_WRT_C_PREV");
        const pa_mesh* m = p->mesh;
        pa::j2::k_local<<<pa::j2::blocks(14L * m->C), pa::j2::kB, 0, (hipStream_t)stream>>>(
            m->C, 1, wrt, pa::j2::constants(p->prm), m->d_conn.p, pa::j2::coords4(p), u, c, cp, z, out);
        pa::j2::check();
    });
}

int pa_j2_global_value(pa_plasticity* p, const double* u, const double* c, const double* pi, const double* z,
                       double* R, void* stream) {
    return guarded([&] {
        PA_REQUIRE(p && u && c && pi && z && R, PA_ERR_ARG, "pa_j2_global_value: null argument");
        const pa_mesh* m = p->mesh;
        const hipStream_t st = (hipStream_t)stream;
        pa::j2::k_global<<<pa::j2::blocks(16L * m->C), pa::j2::kB, 0, st>>>(
            m->C, 0, 0, pa::j2::constants(p->prm), m->d_conn.p, pa::j2::coords4(p), u, c, pi, z, p->vcell.p);
        pa::j2::check();
        pa::j2::gather_vcell(p, 1.0, R, st);
    });
}

int pa_j2_global_gradient(pa_plasticity* p, int wrt, const double* u, const double* c, const double* pi,
                          const double* z, double* out, void* stream) {
    return guarded([&] {
        PA_REQUIRE(p && u && c && pi && z && out, PA_ERR_ARG, "pa_j2_global_gradient: null argument");
        PA_REQUIRE(wrt == PA_WRT_U || wrt == PA_WRT_C || wrt == PA_WRT_Z || wrt == PA_WRT_PI, PA_ERR_ARG,
                   "pa_j2_global_gradient: wrt must be PA_WRT_U, PA_WRT_C, PA_WRT_Z or PA_WRT_PI");
        const pa_mesh* m = p->mesh;
        pa::j2::k_global<<<pa::j2::blocks(16L * m->C), pa::j2::kB, 0, (hipStream_t)stream>>>(
            m->C, 1, wrt, pa::j2::constants(p->prm), m->d_conn.p, pa::j2::coords4(p), u, c, pi, z, out);
        pa::j2::check();
    });
}

int pa_j2_condensed_system(pa_plasticity* p, const double* u, const double* c, const double* cp, const double* pi,
                           const double* z, double* values, double* R, int transpose, void* stream) {
    return guarded([&] {
        PA_REQUIRE(p && u && c && cp && pi && z && values && (R || transpose), PA_ERR_ARG,
                   "pa_j2_condensed_system: null argument");
        pa::j2::condensed_system(p, u, c, cp, pi, z, values, R, transpose != 0, (hipStream_t)stream);
    });
}

int pa_j2_projection(pa_plasticity* p, const double* u, const double* z, double* pi, void* stream) {
    return guarded([&] {
        PA_REQUIRE(p && u && z && pi, PA_ERR_ARG, "pa_j2_projection: null argument");
        pa::j2::projection(p, u, z, pi, (hipStream_t)stream);
    });
}

int pa_j2_plastic_work(pa_plasticity* p, const double* u, const double* up, const double* c, const double* cp,
                       const double* z, double* value, void* stream) {
    return guarded([&] {
        PA_REQUIRE(p && u && up && c && cp && z && value, PA_ERR_ARG, "pa_j2_plastic_work: null argument");
        pa::j2::plastic_work(p, u, up, c, cp, z, value, (hipStream_t)stream);
    });
}

}  // extern "C"
