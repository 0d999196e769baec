// CPU FP64 oracle for the J2 plasticity path (stabilized mixed u-p global
// residual, 14-variable local J2 residual, closed-form return map, Schur
// condensation, pressure-gradient projection, plastic-work criterion).
// TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the CPU-baseline
// leg of the benchmark load it; the product library never links or calls it.
//
// Structure follows the reference: per-cell evaluation in forward-mode dual
// numbers (the Sacado SFad equivalent) for every partial derivative, the
// batched 14x14 inverse by LU without pivoting, the Schur complement formed
// from the separate workset Jacobians, and an atomic scatter by linear search
// of the CRS row.
//
// Pinning:
//   local residual        -> unit_tests/PlasticityTests.cpp:1833-1836 (1e-5)
//   local state update    -> unit_tests/PlasticityTests.cpp:2038-2041 (1e-5)
//   global residual       -> unit_tests/ElastoPlasticityTests.cpp:2069-2094 (1e-4)
//   batched inverse       -> A * A^-1 = I and the 2x2 case of
//                            unit_tests/ElastoPlasticityTests.cpp:1334-1372
//   every dual-number partial -> central finite differences of the values
//   closed-form radial return -> H(update(u)) = 0 on plastic cells
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

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
        v *= o.v;
        return *this;
    }
    Fad& operator/=(const Fad& o) {
        const double inv = 1.0 / o.v, q = v * inv;
        for (int i = 0; i < W; ++i) d[i] = (d[i] - q * o.d[i]) * inv;
        v = q;
        return *this;
    }
    Fad& operator*=(double s) { v *= s; for (int i = 0; i < W; ++i) d[i] *= s; return *this; }
    Fad& operator/=(double s) { v /= s; for (int i = 0; i < W; ++i) d[i] /= s; return *this; }
};
template <int W> Fad<W> operator+(Fad<W> a, const Fad<W>& b) { return a += b; }
template <int W> Fad<W> operator-(Fad<W> a, const Fad<W>& b) { return a -= b; }
template <int W> Fad<W> operator*(Fad<W> a, const Fad<W>& b) { return a *= b; }
template <int W> Fad<W> operator/(Fad<W> a, const Fad<W>& b) { return a /= b; }
template <int W> Fad<W> operator*(Fad<W> a, double s) { return a *= s; }
template <int W> Fad<W> operator*(double s, Fad<W> a) { return a *= s; }
template <int W> Fad<W> operator/(Fad<W> a, double s) { return a /= s; }
template <int W> Fad<W> operator+(Fad<W> a, double s) { a.v += s; return a; }
template <int W> Fad<W> operator+(double s, Fad<W> a) { a.v += s; return a; }
template <int W> Fad<W> operator-(Fad<W> a, double s) { a.v -= s; return a; }
template <int W> Fad<W> operator-(double s, const Fad<W>& a) { Fad<W> r; r.v = s - a.v; for (int i = 0; i < W; ++i) r.d[i] = -a.d[i]; return r; }
template <int W> Fad<W> operator-(const Fad<W>& a) { Fad<W> r; r.v = -a.v; for (int i = 0; i < W; ++i) r.d[i] = -a.d[i]; return r; }
template <int W> Fad<W> fsqrt(const Fad<W>& a) {
    Fad<W> r;
    r.v = std::sqrt(a.v);
    const double h = 0.5 / r.v;
    for (int i = 0; i < W; ++i) r.d[i] = h * a.d[i];
    return r;
}
template <int W> Fad<W> fpow(const Fad<W>& a, double p) {
    Fad<W> r;
    r.v = std::pow(a.v, p);
    const double dp = (a.v == 0.0) ? 0.0 : p * std::pow(a.v, p - 1.0);
    for (int i = 0; i < W; ++i) r.d[i] = dp * a.d[i];
    return r;
}
inline double fsqrt(double a) { return std::sqrt(a); }
inline double fpow(double a, double p) { return std::pow(a, p); }
inline double val(double x) { return x; }
template <int W> inline double val(const Fad<W>& x) { return x.v; }
inline double dv(double, int) { return 0.0; }
template <int W> inline double dv(const Fad<W>& x, int k) { return x.d[k]; }
inline void seed(double&, int) {}
template <int W> inline void seed(Fad<W>& x, int k) { x.d[k] = 1.0; }

// ComputeGradientWorkset<3>, src/ImplicitFunctors.hpp:169-221.
inline void compute_gradient(const double cfg[4][3], double g[4][3], double& vol) {
    double j11 = cfg[0][0] - cfg[3][0], j12 = cfg[1][0] - cfg[3][0], j13 = cfg[2][0] - cfg[3][0];
    double j21 = cfg[0][1] - cfg[3][1], j22 = cfg[1][1] - cfg[3][1], j23 = cfg[2][1] - cfg[3][1];
    double j31 = cfg[0][2] - cfg[3][2], j32 = cfg[1][2] - cfg[3][2], j33 = cfg[2][2] - cfg[3][2];
    double detj = j11 * j22 * j33 + j12 * j23 * j31 + j13 * j21 * j32 - j11 * j23 * j32 - j12 * j21 * j33 -
                  j13 * j22 * j31;
    double i11 = (j22 * j33 - j23 * j32) / detj, i12 = (j13 * j32 - j12 * j33) / detj, i13 = (j12 * j23 - j13 * j22) / detj;
    double i21 = (j23 * j31 - j21 * j33) / detj, i22 = (j11 * j33 - j13 * j31) / detj, i23 = (j13 * j21 - j11 * j23) / detj;
    double i31 = (j21 * j32 - j22 * j31) / detj, i32 = (j12 * j31 - j11 * j32) / detj, i33 = (j11 * j22 - j12 * j21) / detj;
    vol = std::fabs(detj);
    g[0][0] = i11; g[0][1] = i12; g[0][2] = i13;
    g[1][0] = i21; g[1][1] = i22; g[1][2] = i23;
    g[2][0] = i31; g[2][1] = i32; g[2][2] = i33;
    g[3][0] = -(i11 + i21 + i31); g[3][1] = -(i12 + i22 + i32); g[3][2] = -(i13 + i23 + i33);
}

const int kVoigt[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};

// MSIMP, src/Simp.hpp:130-134.
template <class T> inline T msimp(const T& x, double p, double emin) { return emin + (1.0 - emin) * fpow(x, p); }

}  // namespace

extern "C" {

typedef struct {
    int num_nodes;
    int num_cells;
    const double* coords;  // N*3
    const int* conn;       // C*4
    double youngs, poisson;
    // J2 Plasticity sublist (src/J2PlasticityLocalResidual.hpp:195-220)
    double hardening_isotropic, hardening_kinematic, initial_yield_stress;
    double elastic_penalty_exponent, elastic_min_ersatz;  // defaults 1.0, 1e-9
    double plastic_penalty_exponent, plastic_min_ersatz;  // defaults 0.5, 1e-4
    // Elliptic / Penalty Function of the global residual
    // (src/InfinitesimalStrainPlasticityResidual.hpp:142-160): defaults 1.0, 1e-9
    double global_penalty_exponent, global_min_ersatz;
    double pressure_scaling;  // Material Models / Pressure Scaling, default 1
    // State Gradient Projection / Penalty Function (src/Projection.hpp:30-62):
    // 0 = NoPenalty, 1 = SIMP with the two values below
    int projection_penalty;
    double projection_penalty_exponent, projection_min_ersatz;
    // plastic work criterion penalty (src/PlasticWorkCriterion.hpp:295-303)
    double work_penalty_exponent, work_min_ersatz;
    // 0: the reference's stabilized mixed u-p formulation (4 dof per node);
    // 1: pure-displacement J2 (3 dof per node), which has no reference
    //    counterpart: sigma = Dev(mu)(eps - eps_p) + kappa tr(eps - eps_p) I,
    //    no pressure unknown, stabilization or projection.
    int formulation;
} orc_j2_problem;

}  // extern "C"

namespace {

struct CellGeom {
    double g[4][3];
    double vol;  // |det J| / 6 (cell volume times the cubature weight)
};

inline CellGeom cell_geom(const orc_j2_problem* P, int c) {
    double cfg[4][3];
    for (int n = 0; n < 4; ++n)
        for (int d = 0; d < 3; ++d) cfg[n][d] = P->coords[(size_t)P->conn[(size_t)c * 4 + n] * 3 + d];
    CellGeom G;
    double detj;
    compute_gradient(cfg, G.g, detj);
    G.vol = detj * (1.0 / 6.0);
    return G;
}

inline double shear_modulus(const orc_j2_problem* P) { return P->youngs / (2.0 * (1.0 + P->poisson)); }
inline int dofs(const orc_j2_problem* P) { return P->formulation ? 3 : 4; }
inline double bulk_modulus(const orc_j2_problem* P) { return P->youngs / (3.0 * (1.0 - 2.0 * P->poisson)); }

// Strain<3, 4>, src/Strain.hpp:56-95: Voigt (xx, yy, zz, yz, xz, xy),
// engineering shears; local ordinal node*4 + dim.
template <class T>
inline void total_strain(const double g[4][3], const T u[16], T eps[6]) {
    for (int d = 0; d < 3; ++d) {
        eps[d] = 0.0;
        for (int n = 0; n < 4; ++n) eps[d] += u[n * 4 + d] * g[n][d];
    }
    int v = 3;
    for (int j = 2; j >= 1; --j)
        for (int i = j - 1; i >= 0; --i) {
            eps[v] = 0.0;
            for (int n = 0; n < 4; ++n) eps[v] += u[n * 4 + i] * g[n][j] + u[n * 4 + j] * g[n][i];
            ++v;
        }
}

// ComputeDeviatoricStress<3>, src/ComputeDeviatoricStress.hpp:107-127.
template <class T, class M>
inline void deviatoric_stress(const M& mu, const T e[6], T s[6]) {
    T tr = (e[0] + e[1] + e[2]) / 3.0;
    for (int i = 0; i < 3; ++i) s[i] = (2.0 * mu) * (e[i] - tr);
    for (int i = 3; i < 6; ++i) s[i] = mu * e[i];
}

// computeDeviatoricStressMinusBackstressNormalized<3>,
// src/J2PlasticityUtilities.hpp:372-405.
template <class T>
inline T normalized_relative_stress(const T s[6], const T c[14], T n[6]) {
    for (int i = 0; i < 6; ++i) n[i] = s[i] - c[8 + i];
    T nrm = fsqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2] + 2.0 * (n[3] * n[3]) + 2.0 * (n[4] * n[4]) +
                  2.0 * (n[5] * n[5]));
    for (int i = 0; i < 6; ++i) n[i] = n[i] / nrm;
    return nrm;
}

template <class T>
inline T cell_density(const T z[4]) { return (z[0] + z[1] + z[2] + z[3]) / 4.0; }

// J2PlasticityLocalResidual::evaluate, src/J2PlasticityLocalResidual.hpp:310-428.
template <class T>
void local_residual_cell(const orc_j2_problem* P, const CellGeom& G, const T u[16], const T c[14], const T cp[14],
                         const T z[4], T H[14]) {
    const double sqrt32 = std::sqrt(3.0 / 2.0), sqrt23 = std::sqrt(2.0 / 3.0);
    T eps[6], ee[6], s[6], n[6];
    total_strain(G.g, u, eps);
    for (int i = 0; i < 6; ++i) ee[i] = eps[i] - c[2 + i];  // ThermoPlasticityUtilities::computeElasticStrain
    T rho = cell_density(z);
    T epen = msimp(rho, P->elastic_penalty_exponent, P->elastic_min_ersatz);
    T mu = epen * shear_modulus(P);
    deviatoric_stress(mu, ee, s);
    T nrm = normalized_relative_stress(s, c, n);
    T ppen = msimp(rho, P->plastic_penalty_exponent, P->plastic_min_ersatz);
    T hiso = ppen * P->hardening_isotropic, hkin = ppen * P->hardening_kinematic, sy = ppen * P->initial_yield_stress;
    T yield = sy + hiso * c[0];
    H[0] = c[0] - cp[0];
    H[1] = c[1];
    for (int i = 0; i < 6; ++i) H[2 + i] = c[2 + i] - cp[2 + i];
    for (int i = 0; i < 6; ++i) H[8 + i] = c[8 + i] - cp[8 + i];
    if (val(c[1]) > 0.0) {
        H[0] = c[0] - cp[0] - c[1];
        H[1] = sqrt32 * nrm - yield;
        for (int i = 0; i < 6; ++i) {
            const double f = i < 3 ? 1.0 : 2.0;  // fillPlasticStrainTensorResidualPlasticStep, :514-540
            H[2 + i] = c[2 + i] - cp[2 + i] - f * sqrt32 * c[1] * n[i];
        }
        for (int i = 0; i < 6; ++i)  // fillBackstressTensorResidualPlasticStep, :581-609
            H[8 + i] = c[8 + i] - cp[8 + i] - sqrt23 * c[1] * hkin * n[i];
    }
}

// J2PlasticityLocalResidual::updateLocalState, src/J2PlasticityLocalResidual.hpp:440-543.
void update_cell(const orc_j2_problem* P, const CellGeom& G, const double u[16], const double cp[14],
                 const double z[4], double c[14]) {
    const double sqrt32 = std::sqrt(3.0 / 2.0), sqrt23 = std::sqrt(2.0 / 3.0);
    c[0] = cp[0];
    c[1] = 0.0;
    for (int i = 2; i < 14; ++i) c[i] = cp[i];
    double eps[6], ee[6], s[6], n[6];
    total_strain(G.g, u, eps);
    for (int i = 0; i < 6; ++i) ee[i] = eps[i] - c[2 + i];
    double rho = cell_density(z);
    double mu = msimp(rho, P->elastic_penalty_exponent, P->elastic_min_ersatz) * shear_modulus(P);
    deviatoric_stress(mu, ee, s);
    double nrm = normalized_relative_stress(s, c, n);
    double ppen = msimp(rho, P->plastic_penalty_exponent, P->plastic_min_ersatz);
    double hiso = ppen * P->hardening_isotropic, hkin = ppen * P->hardening_kinematic;
    double sy = ppen * P->initial_yield_stress;
    double yield = sy + hiso * c[0];
    double ftrial = sqrt32 * nrm - yield;
    if (ftrial > 1.0e-10) {
        c[1] = ftrial / (3.0 * mu + hiso + hkin);
        c[0] = cp[0] + c[1];
        // updatePlasticStrainAndBackstressPlasticStep<3>, src/J2PlasticityUtilities.hpp:250-275
        const double m1 = c[1] * sqrt32;
        for (int i = 0; i < 6; ++i) c[2 + i] = cp[2 + i] + (i < 3 ? 1.0 : 2.0) * m1 * n[i];
        const double m2 = c[1] * sqrt23 * hkin;
        for (int i = 0; i < 6; ++i) c[8 + i] = cp[8 + i] + m2 * n[i];
    }
}

// InfinitesimalStrainPlasticityResidual::evaluate,
// src/InfinitesimalStrainPlasticityResidual.hpp:393-521 (no body loads).
// pi: projected pressure gradient of the cell's four nodes (node*3 + dim).
template <class T>
void global_residual_cell(const orc_j2_problem* P, const CellGeom& G, const T u[16], const T c[14], const T pi[12],
                          const T z[4], T R[16]) {
    if (P->formulation) {
        // pure displacement: R_{n,i} = V sum_j sigma_ij dN_n/dx_j, with the
        // global residual's penalty on both moduli
        const double V = G.vol;
        T eps[6], ee[6], s[6];
        total_strain(G.g, u, eps);
        for (int i = 0; i < 6; ++i) ee[i] = eps[i] - c[2 + i];
        T pen = msimp(cell_density(z), P->global_penalty_exponent, P->global_min_ersatz);
        T mu = pen * shear_modulus(P);
        deviatoric_stress(mu, ee, s);
        T vol = (pen * bulk_modulus(P)) * ((ee[0] + ee[1]) + ee[2]);
        for (int i = 0; i < 3; ++i) s[i] = s[i] + vol;
        for (int k = 0; k < 16; ++k) R[k] = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int n = 0; n < 4; ++n)
                for (int j = 0; j < 3; ++j) R[n * 4 + i] += V * s[kVoigt[i][j]] * G.g[n][j];
        return;
    }
    const double ps = P->pressure_scaling;
    const double V = G.vol;
    T eps[6], ee[6], s[6];
    total_strain(G.g, u, eps);
    for (int i = 0; i < 6; ++i) ee[i] = eps[i] - c[2 + i];
    T pgrad[3];  // ScalarGrad with dof offset 3
    for (int d = 0; d < 3; ++d) {
        pgrad[d] = 0.0;
        for (int n = 0; n < 4; ++n) pgrad[d] += u[n * 4 + 3] * G.g[n][d];
    }
    T p = 0.0;  // InterpolateFromNodal, basis 1/4
    for (int n = 0; n < 4; ++n) p += 0.25 * u[n * 4 + 3];
    T pig[3];
    for (int d = 0; d < 3; ++d) {
        pig[d] = 0.0;
        for (int n = 0; n < 4; ++n) pig[d] += 0.25 * pi[n * 3 + d];
    }
    T rho = cell_density(z);
    T pen = msimp(rho, P->global_penalty_exponent, P->global_min_ersatz);
    T mu = pen * shear_modulus(P);
    deviatoric_stress(mu, ee, s);
    T div = eps[0] + eps[1] + eps[2];  // StrainDivergence<3> of the total strain
    T kappa = pen * bulk_modulus(P);
    p = p * ps;
    T volstrain = ps * (div - p / kappa);
    // ComputeStabilization<3>, src/ComputeStabilization.hpp:93-106 (unpenalized shear modulus)
    const double tau = std::pow(V, 2.0 / 3.0) / (2.0 * shear_modulus(P));
    T stab[3];
    for (int d = 0; d < 3; ++d) stab[d] = ps * tau * ((ps * pgrad[d]) - pig[d]);
    for (int k = 0; k < 16; ++k) R[k] = 0.0;
    for (int i = 0; i < 3; ++i)  // StressDivergence<3,4,0>
        for (int n = 0; n < 4; ++n)
            for (int j = 0; j < 3; ++j) R[n * 4 + i] += V * s[kVoigt[i][j]] * G.g[n][j];
    for (int i = 0; i < 3; ++i)  // PressureDivergence<3,4>
        for (int n = 0; n < 4; ++n) R[n * 4 + i] += V * p * G.g[n][i];
    for (int n = 0; n < 4; ++n)  // FluxDivergence<3,4,3> with scale -1
        for (int d = 0; d < 3; ++d) R[n * 4 + 3] += -1.0 * stab[d] * G.g[n][d] * V;
    for (int n = 0; n < 4; ++n)  // ProjectToNode<3,4,3>
        R[n * 4 + 3] += 0.25 * volstrain * V;
}

// PlasticWorkCriterion::evaluate, src/PlasticWorkCriterion.hpp:163-265.
template <class T>
T plastic_work_cell(const orc_j2_problem* P, const CellGeom& G, const T u[16], const T up[16], const T c[14],
                    const T cp[14], const T z[4]) {
    const double ps = P->pressure_scaling;
    T pp = 0.0, pc = 0.0;
    for (int n = 0; n < 4; ++n) {
        pp += 0.25 * up[n * 4 + 3];
        pc += 0.25 * u[n * 4 + 3];
    }
    pp = pp * ps;
    pc = pc * ps;
    T misfit[6];
    for (int i = 0; i < 6; ++i) misfit[i] = c[2 + i] - cp[2 + i];
    T epsp[6], epsc[6], eep[6], eec[6], sp[6], sc[6];
    total_strain(G.g, up, epsp);
    total_strain(G.g, u, epsc);
    for (int i = 0; i < 6; ++i) {
        eep[i] = epsp[i] - cp[2 + i];
        eec[i] = epsc[i] - c[2 + i];
    }
    T rho = cell_density(z);
    const T wpen = msimp(rho, P->work_penalty_exponent, P->work_min_ersatz);
    if (P->formulation) {  // the volumetric stress takes the pressure's place
        pp = (wpen * bulk_modulus(P)) * ((eep[0] + eep[1]) + eep[2]);
        pc = (wpen * bulk_modulus(P)) * ((eec[0] + eec[1]) + eec[2]);
    }
    T mu = wpen * shear_modulus(P);
    deviatoric_stress(mu, eep, sp);
    deviatoric_stress(mu, eec, sc);
    T w = 0.0;
    for (int i = 0; i < 6; ++i) {
        T sig_p = i < 3 ? sp[i] + pp : sp[i];  // ComputeStabilizedCauchyStress<3>
        T sig_c = i < 3 ? sc[i] + pc : sc[i];
        w += (0.5 * (sig_c + sig_p)) * misfit[i];
    }
    return w * G.vol;
}

// gathers of one cell
// cell layout node*4 + dim; the pressure slot is zero in the displacement formulation
inline void gather_u(const orc_j2_problem* P, const double* u, int c, double out[16]) {
    const int D = dofs(P);
    for (int n = 0; n < 4; ++n)
        for (int d = 0; d < 4; ++d) out[n * 4 + d] = d < D ? u[(size_t)P->conn[(size_t)c * 4 + n] * D + d] : 0.0;
}
inline void gather_z(const orc_j2_problem* P, const double* z, int c, double out[4]) {
    for (int n = 0; n < 4; ++n) out[n] = z[P->conn[(size_t)c * 4 + n]];
}

template <class F>
void for_cells(int C, F&& f) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < C; ++c) f(c);
}

// Dual-number Jacobian of the local residual with respect to one input.
// which: 0 = c (14), 1 = u (16), 2 = z (4), 3 = c_prev (14). Output row-major
// [cell][H index][input index].
template <int W>
void local_jacobian_cell(const orc_j2_problem* P, int which, const CellGeom& G, const double u[16],
                         const double c[14], const double cp[14], const double z[4], double* out) {
    using T = Fad<W>;
    T tu[16], tc[14], tcp[14], tz[4], H[14];
    for (int k = 0; k < 16; ++k) tu[k] = u[k];
    for (int k = 0; k < 14; ++k) { tc[k] = c[k]; tcp[k] = cp[k]; }
    for (int k = 0; k < 4; ++k) tz[k] = z[k];
    T* seedv = which == 0 ? tc : which == 1 ? tu : which == 2 ? tz : tcp;
    for (int k = 0; k < W; ++k) seed(seedv[k], k);
    local_residual_cell(P, G, tu, tc, tcp, tz, H);
    for (int r = 0; r < 14; ++r)
        for (int k = 0; k < W; ++k) out[r * W + k] = H[r].d[k];
}

// which: 0 = u (16), 1 = c (14), 2 = z (4), 3 = projected pressure gradient (12).
template <int W>
void global_jacobian_cell(const orc_j2_problem* P, int which, const CellGeom& G, const double u[16],
                          const double c[14], const double pi[12], const double z[4], double* out) {
    using T = Fad<W>;
    T tu[16], tc[14], tpi[12], tz[4], R[16];
    for (int k = 0; k < 16; ++k) tu[k] = u[k];
    for (int k = 0; k < 14; ++k) tc[k] = c[k];
    for (int k = 0; k < 12; ++k) tpi[k] = pi[k];
    for (int k = 0; k < 4; ++k) tz[k] = z[k];
    T* seedv = which == 0 ? tu : which == 1 ? tc : which == 2 ? tz : tpi;
    for (int k = 0; k < W; ++k) seed(seedv[k], k);
    global_residual_cell(P, G, tu, tc, tpi, tz, R);
    for (int r = 0; r < 16; ++r)
        for (int k = 0; k < W; ++k) out[r * W + k] = R[r].d[k];
}

// KokkosBatched SerialLU (no pivoting) + two SerialTrsm on the identity,
// src/BLAS3.hpp:154-183.
template <int M>
void lu_inverse(const double* A_in, double* Ainv) {
    double A[M][M];
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) A[i][j] = A_in[i * M + j];
    for (int k = 0; k < M; ++k)
        for (int i = k + 1; i < M; ++i) {
            A[i][k] /= A[k][k];
            for (int j = k + 1; j < M; ++j) A[i][j] -= A[i][k] * A[k][j];
        }
    for (int col = 0; col < M; ++col) {
        double x[M];
        for (int i = 0; i < M; ++i) x[i] = (i == col) ? 1.0 : 0.0;
        for (int i = 0; i < M; ++i)  // L (unit) forward
            for (int j = 0; j < i; ++j) x[i] -= A[i][j] * x[j];
        for (int i = M - 1; i >= 0; --i) {  // U backward
            for (int j = i + 1; j < M; ++j) x[i] -= A[i][j] * x[j];
            x[i] /= A[i][i];
        }
        for (int i = 0; i < M; ++i) Ainv[i * M + col] = x[i];
    }
}

inline long entry_ordinal(const int* rowmap, const int* colidx, int row, int col) {
    for (int e = rowmap[row]; e < rowmap[row + 1]; ++e)
        if (colidx[e] == col) return e;
    return -1;
}

inline void atomic_add(double* dst, double x) {
#ifdef _OPENMP
#pragma omp atomic
#endif
    *dst += x;
}

}  // namespace

extern "C" {

// LocalVectorFunctionInc::updateLocalState. u: D N, c_prev / c: C*14, z: N.
void orc_j2_update_local_state(const orc_j2_problem* P, const double* u, const double* cp, const double* z,
                               double* c) {
    for_cells(P->num_cells, [&](int e) {
        CellGeom G = cell_geom(P, e);
        double ue[16], ze[4];
        gather_u(P, u, e, ue);
        gather_z(P, z, e, ze);
        update_cell(P, G, ue, cp + (size_t)e * 14, ze, c + (size_t)e * 14);
    });
}

// LocalVectorFunctionInc::valueWorkSet: H per cell (C*14).
void orc_j2_local_value(const orc_j2_problem* P, const double* u, const double* c, const double* cp,
                        const double* z, double* H) {
    for_cells(P->num_cells, [&](int e) {
        CellGeom G = cell_geom(P, e);
        double ue[16], ze[4];
        gather_u(P, u, e, ue);
        gather_z(P, z, e, ze);
        local_residual_cell<double>(P, G, ue, c + (size_t)e * 14, cp + (size_t)e * 14, ze, H + (size_t)e * 14);
    });
}

// LocalVectorFunctionInc::gradient_{c,u,z,cp} workset: C x 14 x W.
void orc_j2_local_gradient(const orc_j2_problem* P, int which, const double* u, const double* c, const double* cp,
                           const double* z, double* out) {
    const int W = which == 1 ? 16 : which == 2 ? 4 : 14;
    for_cells(P->num_cells, [&](int e) {
        CellGeom G = cell_geom(P, e);
        double ue[16], ze[4];
        gather_u(P, u, e, ue);
        gather_z(P, z, e, ze);
        double* o = out + (size_t)e * 14 * W;
        const double* ce = c + (size_t)e * 14;
        const double* cpe = cp + (size_t)e * 14;
        if (W == 16) local_jacobian_cell<16>(P, which, G, ue, ce, cpe, ze, o);
        else if (W == 4) local_jacobian_cell<4>(P, which, G, ue, ce, cpe, ze, o);
        else local_jacobian_cell<14>(P, which, G, ue, ce, cpe, ze, o);
    });
}

// Global residual cell workset (C x 16) with the projected pressure gradient
// given as a cell workset (C x 12), the layout the reference's evaluators take.
void orc_j2_global_value_ws(const orc_j2_problem* P, const double* u, const double* c, const double* pi_ws,
                            const double* z, double* Rws) {
    for_cells(P->num_cells, [&](int e) {
        CellGeom G = cell_geom(P, e);
        double ue[16], ze[4];
        gather_u(P, u, e, ue);
        gather_z(P, z, e, ze);
        global_residual_cell<double>(P, G, ue, c + (size_t)e * 14, pi_ws + (size_t)e * 12, ze, Rws + (size_t)e * 16);
    });
}

// GlobalVectorFunctionInc::gradient_{u,c,z} and the projection partial,
// cell workset C x 16 x W.
void orc_j2_global_gradient_ws(const orc_j2_problem* P, int which, const double* u, const double* c,
                               const double* pi_ws, const double* z, double* out) {
    const int W = which == 0 ? 16 : which == 1 ? 14 : which == 2 ? 4 : 12;
    for_cells(P->num_cells, [&](int e) {
        CellGeom G = cell_geom(P, e);
        double ue[16], ze[4];
        gather_u(P, u, e, ue);
        gather_z(P, z, e, ze);
        double* o = out + (size_t)e * 16 * W;
        const double* ce = c + (size_t)e * 14;
        const double* pe = pi_ws + (size_t)e * 12;
        if (W == 16) global_jacobian_cell<16>(P, which, G, ue, ce, pe, ze, o);
        else if (W == 14) global_jacobian_cell<14>(P, which, G, ue, ce, pe, ze, o);
        else if (W == 4) global_jacobian_cell<4>(P, which, G, ue, ce, pe, ze, o);
        else global_jacobian_cell<12>(P, which, G, ue, ce, pe, ze, o);
    });
}

// Batched inverse of n x n matrices (n = 2 or 14), src/BLAS3.hpp:154-183.
void orc_batched_inverse(int count, int n, const double* A, double* Ainv) {
#pragma omp parallel for schedule(static)
    for (int e = 0; e < count; ++e) {
        if (n == 14) lu_inverse<14>(A + (size_t)e * 196, Ainv + (size_t)e * 196);
        else if (n == 2) lu_inverse<2>(A + (size_t)e * 4, Ainv + (size_t)e * 4);
    }
}

// One Newton linearization of the condensed system
// (src/NewtonRaphsonSolver.hpp:204-330):
//   K_t = dR/du - dR/dc (dH/dc)^-1 dH/du   -> block CRS, D x D blocks
//   R_t = R - dR/dc (dH/dc)^-1 H           -> D N
// (D = 4 in the mixed formulation, 3 in the displacement one).
// pi: nodal projected pressure gradient (N*3).  Scatter by row search and
// atomic add (src/Assembly.hpp:1222-1521).  The caller's values/R must be zero.
void orc_j2_condensed_system(const orc_j2_problem* P, const double* u, const double* c, const double* cp,
                             const double* pi, const double* z, const int* rowmap, const int* colidx,
                             double* values, double* R) {
    for_cells(P->num_cells, [&](int e) {
        CellGeom G = cell_geom(P, e);
        double ue[16], ze[4], pe[12];
        gather_u(P, u, e, ue);
        gather_z(P, z, e, ze);
        for (int n = 0; n < 4; ++n)
            for (int d = 0; d < 3; ++d) pe[n * 3 + d] = pi[(size_t)P->conn[(size_t)e * 4 + n] * 3 + d];
        const double* ce = c + (size_t)e * 14;
        const double* cpe = cp + (size_t)e * 14;
        double dHdc[196], inv[196], dHdu[14 * 16], dRdc[16 * 14], dRdu[256], H[14], Re[16];
        local_jacobian_cell<14>(P, 0, G, ue, ce, cpe, ze, dHdc);
        local_jacobian_cell<16>(P, 1, G, ue, ce, cpe, ze, dHdu);
        global_jacobian_cell<14>(P, 1, G, ue, ce, pe, ze, dRdc);
        global_jacobian_cell<16>(P, 0, G, ue, ce, pe, ze, dRdu);
        local_residual_cell<double>(P, G, ue, ce, cpe, ze, H);
        global_residual_cell<double>(P, G, ue, ce, pe, ze, Re);
        lu_inverse<14>(dHdc, inv);
        double X[14 * 16], y[14];
        for (int i = 0; i < 14; ++i) {
            for (int j = 0; j < 16; ++j) {
                double s = 0.0;
                for (int k = 0; k < 14; ++k) s += inv[i * 14 + k] * dHdu[k * 16 + j];
                X[i * 16 + j] = s;
            }
            double s = 0.0;
            for (int k = 0; k < 14; ++k) s += inv[i * 14 + k] * H[k];
            y[i] = s;
        }
        double Kt[256], Rt[16];
        for (int i = 0; i < 16; ++i) {
            for (int j = 0; j < 16; ++j) {
                double s = 0.0;
                for (int k = 0; k < 14; ++k) s += dRdc[i * 14 + k] * X[k * 16 + j];
                Kt[i * 16 + j] = dRdu[i * 16 + j] - s;
            }
            double s = 0.0;
            for (int k = 0; k < 14; ++k) s += dRdc[i * 14 + k] * y[k];
            Rt[i] = Re[i] - s;
        }
        const int D = dofs(P);
        for (int a = 0; a < 4; ++a) {
            const int row = P->conn[(size_t)e * 4 + a];
            for (int i = 0; i < D; ++i) atomic_add(&R[(size_t)row * D + i], Rt[a * 4 + i]);
            for (int b = 0; b < 4; ++b) {
                const long ent = entry_ordinal(rowmap, colidx, row, P->conn[(size_t)e * 4 + b]);
                for (int i = 0; i < D; ++i)
                    for (int j = 0; j < D; ++j)
                        atomic_add(&values[ent * D * D + i * D + j], Kt[(a * 4 + i) * 16 + b * 4 + j]);
            }
        }
    });
}

// Assembled global residual R(u, c, pi, z) (D N), without condensation.
void orc_j2_global_value(const orc_j2_problem* P, const double* u, const double* c, const double* pi,
                         const double* z, double* R) {
    for_cells(P->num_cells, [&](int e) {
        CellGeom G = cell_geom(P, e);
        double ue[16], ze[4], pe[12], Re[16];
        gather_u(P, u, e, ue);
        gather_z(P, z, e, ze);
        for (int n = 0; n < 4; ++n)
            for (int d = 0; d < 3; ++d) pe[n * 3 + d] = pi[(size_t)P->conn[(size_t)e * 4 + n] * 3 + d];
        global_residual_cell<double>(P, G, ue, c + (size_t)e * 14, pe, ze, Re);
        const int D = dofs(P);
        for (int a = 0; a < 4; ++a)
            for (int i = 0; i < D; ++i) atomic_add(&R[(size_t)P->conn[(size_t)e * 4 + a] * D + i], Re[a * 4 + i]);
    });
}

// Projected pressure gradient of the next step (src/PlasticityProblem.hpp:872-891,
// src/PressureGradientProjectionResidual.hpp:111-162, Solve::RowSummed
// src/Plato_Solve.hpp:22-45): with Pi = 0 the residual is
//   r_a = - sum_cells 1/4 * V * w(rho) * ps * grad p,
// the Jacobian block row sum is sum_cells 1/4 * V * w(rho) (identity blocks),
// and Pi = -r / rowsum.  u is the 4-dof global state (pressure at offset 3).
void orc_j2_projection(const orc_j2_problem* P, const double* u, const double* z, double* pi) {
    const int N = P->num_nodes;
    std::vector<double> r((size_t)N * 3, 0.0), rowsum((size_t)N, 0.0);
    const double ps = P->pressure_scaling;
    for_cells(P->num_cells, [&](int e) {
        CellGeom G = cell_geom(P, e);
        double ue[16], ze[4];
        gather_u(P, u, e, ue);
        gather_z(P, z, e, ze);
        double w = 1.0;
        if (P->projection_penalty)
            w = msimp(cell_density(ze), P->projection_penalty_exponent, P->projection_min_ersatz);
        double pg[3];
        for (int d = 0; d < 3; ++d) {
            pg[d] = 0.0;
            for (int n = 0; n < 4; ++n) pg[d] += ps * ue[n * 4 + 3] * G.g[n][d];
            pg[d] *= w;
        }
        for (int a = 0; a < 4; ++a) {
            const int node = P->conn[(size_t)e * 4 + a];
            for (int d = 0; d < 3; ++d) atomic_add(&r[(size_t)node * 3 + d], -1.0 * 0.25 * pg[d] * G.vol);
            // mass entries 1/4 * 1/4 * V * w summed over the cell's four columns
            for (int b = 0; b < 4; ++b) atomic_add(&rowsum[node], 0.25 * 0.25 * w * G.vol);
        }
    });
    for (int i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) pi[(size_t)i * 3 + d] = (-r[(size_t)i * 3 + d]) / rowsum[i];
}

// Basis gradients (C x 4 x 3) and volumes times the cubature weight (C) of
// every cell, for the projection partials assembled on the Python side.
void orc_j2_cell_geometry(const orc_j2_problem* P, double* grads, double* vols) {
    for_cells(P->num_cells, [&](int e) {
        CellGeom G = cell_geom(P, e);
        for (int n = 0; n < 4; ++n)
            for (int d = 0; d < 3; ++d) grads[(size_t)e * 12 + n * 3 + d] = G.g[n][d];
        vols[e] = G.vol;
    });
}

// Plastic work of one step, summed over cells.
double orc_j2_plastic_work(const orc_j2_problem* P, const double* u, const double* up, const double* c,
                           const double* cp, const double* z) {
    double total = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : total)
    for (int e = 0; e < P->num_cells; ++e) {
        CellGeom G = cell_geom(P, e);
        double ue[16], upe[16], ze[4];
        gather_u(P, u, e, ue);
        gather_u(P, up, e, upe);
        gather_z(P, z, e, ze);
        total += plastic_work_cell<double>(P, G, ue, upe, c + (size_t)e * 14, cp + (size_t)e * 14, ze);
    }
    return total;
}

// Partials of one step's plastic work (dual numbers), each accumulated into
// the caller's zeroed arrays: dW/du (D N), dW/du_prev (D N), dW/dc (C*14),
// dW/dc_prev (C*14), dW/dz (N).
void orc_j2_plastic_work_partials(const orc_j2_problem* P, const double* u, const double* up, const double* c,
                                  const double* cp, const double* z, double* gu, double* gup, double* gc,
                                  double* gcp, double* gz) {
    for_cells(P->num_cells, [&](int e) {
        CellGeom G = cell_geom(P, e);
        double ue[16], upe[16], ze[4];
        gather_u(P, u, e, ue);
        gather_u(P, up, e, upe);
        gather_z(P, z, e, ze);
        const double* ce = c + (size_t)e * 14;
        const double* cpe = cp + (size_t)e * 14;
        using T = Fad<64>;
        T tu[16], tup[16], tc[14], tcp[14], tz[4];
        for (int k = 0; k < 16; ++k) { tu[k] = ue[k]; seed(tu[k], k); tup[k] = upe[k]; seed(tup[k], 16 + k); }
        for (int k = 0; k < 14; ++k) { tc[k] = ce[k]; seed(tc[k], 32 + k); tcp[k] = cpe[k]; seed(tcp[k], 46 + k); }
        for (int k = 0; k < 4; ++k) { tz[k] = ze[k]; seed(tz[k], 60 + k); }
        T w = plastic_work_cell(P, G, tu, tup, tc, tcp, tz);
        const int D = dofs(P);
        for (int a = 0; a < 4; ++a) {
            const int node = P->conn[(size_t)e * 4 + a];
            for (int d = 0; d < D; ++d) {
                atomic_add(&gu[(size_t)node * D + d], w.d[a * 4 + d]);
                atomic_add(&gup[(size_t)node * D + d], w.d[16 + a * 4 + d]);
            }
            atomic_add(&gz[node], w.d[60 + a]);
        }
        for (int k = 0; k < 14; ++k) {
            gc[(size_t)e * 14 + k] = w.d[32 + k];
            gcp[(size_t)e * 14 + k] = w.d[46 + k];
        }
    });
}

}  // extern "C"
