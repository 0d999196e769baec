// Internal declarations shared by the J2 plasticity kernels (pa_plasticity.hip)
// and the plasticity problem driver (pa_plasticity_problem.hip).
#pragma once

#include "pa_internal.h"

struct pa_plasticity {
    const pa_mesh* mesh = nullptr;
    pa_j2_params prm{};
    // per-cell workspaces
    pa::DevBuf<double> kcell;   // C*256: condensed cell matrices, block layout [a][b][i][j]
    pa::DevBuf<double> vcell;   // C*16: cell vectors (residuals, adjoint right-hand sides)
    pa::DevBuf<double> part;    // reduction partials
    pa::DevBuf<double> scal;    // device scalars
    pa::DevBuf<double> rowsum;  // N: lumped projection mass
    pa::DevBuf<double> tmp4;    // N*4: projection accumulators
    pa::DevBuf<double> zero4;   // N*4 zeros
    int dpn = 4;                // global dofs per node of the formulation
};

namespace pa {
namespace j2 {

// Kernel-side material constants (unpenalized moduli and the penalty models).
struct KP {
    double mu, kappa;                  // shear and bulk modulus
    double hiso, hkin, sy;             // hardening moduli, initial yield stress
    double pe_e, pe_m, pp_e, pp_m;     // local elastic / plastic properties penalty (exponent, minimum)
    double pg_e, pg_m;                 // global residual penalty
    double ps;                         // pressure scaling
    int proj;                          // projection penalty on/off
    double pj_e, pj_m;                 // projection penalty
    double pw_e, pw_m;                 // plastic work penalty
    int mixed;                         // 1: stabilized mixed u-p; 0: pure displacement
    int dpn;                           // global dofs per node: 4 (mixed) or 3
};

KP constants(const pa_j2_params& p);

// Host wrappers of the kernels used by the problem driver (all asynchronous on st).
void update_local_state(pa_plasticity* p, const double* u, const double* cp, const double* z, double* c,
                        hipStream_t st);
void condensed_system(pa_plasticity* p, const double* u, const double* c, const double* cp, const double* pi,
                      const double* z, double* values, double* R, bool transpose, hipStream_t st);
void projection(pa_plasticity* p, const double* u, const double* z, double* pi, hipStream_t st);
// one step's plastic work summed over cells into scal_dst (device)
void plastic_work(pa_plasticity* p, const double* u, const double* up, const double* c, const double* cp,
                  const double* z, double* scal_dst, hipStream_t st);

// Adjoint pieces of one step (see pa_plasticity_problem.hip for the recursion).
struct AdjointWork {
    double* wu;      // C*16: d f_k / d u_k cell vectors (equal to d f_k / d u_{k-1})
    double* gc;      // C*14: d f_k / d c_k
    double* gcp;     // C*14: d f_k / d c_{k-1}
    double* zc;      // C: 1/4 d f_k / d rho (per node of the cell)
};
// criterion partials of step k (scaled by mult)
void work_partials(pa_plasticity* p, const double* u, const double* up, const double* c, const double* cp,
                   const double* z, double mult, const AdjointWork& w, hipStream_t st);
// local adjoint right-hand side and the global adjoint right-hand side cell vectors:
//   b_local = gc_k + [gcp_{k+1} - mask(mu_{k+1})],  t = (dH/dc)^-T b_local,
//   vcell = wu_k + [wu_{k+1} + DpDn^T gamma_{k+1}] - (dH/du)^T t
// (bracketed terms only when the next-step arrays are given).
void adjoint_local_rhs(pa_plasticity* p, const double* u, const double* c, const double* cp, const double* z,
                       const double* gc, const double* gcp_next, const double* mu_next, const double* wu,
                       const double* wu_next, const double* gamma_next, double* blocal, hipStream_t st);
// mu_k = -(dH/dc)^-T (b_local + (dR/dc)^T lambda_k); per-cell control contributions
// zc_out = 1/4 (dR/drho . lambda + dH/drho . mu) + zc_k; projection adjoint source in vcell[e*16 + 0..2]
void adjoint_local_update(pa_plasticity* p, const double* u, const double* c, const double* cp, const double* pi,
                          const double* z, const double* lambda, const double* blocal, const double* zc,
                          double* mu, double* zc_out, hipStream_t st);
// node gathers of cell-uniform values: out[node*W + k] = scale * sum over incident cells of ws[cell*W + k]
// (accumulate: out += ...)
void gather_uniform(pa_plasticity* p, int W, const double* ws, double scale, bool accumulate, double* out,
                    hipStream_t st);
// the same with a cell stride other than W
void gather_uniform(pa_plasticity* p, int W, const double* ws, int stride, double scale, bool accumulate,
                    double* out, hipStream_t st);
// node gather of per-local-node cell vectors (vcell, layout node*4 + d): out[node*dpn + d] = scale * sum, d < dpn
void gather_vcell(pa_plasticity* p, double scale, double* out, hipStream_t st);
// gamma = v / rowsum (node-wise, 3 components), rowsum of the projection's lumped mass
void divide_rowsum(pa_plasticity* p, double* gamma, const double* z, hipStream_t st);
// projection-penalty control term of step k: zc_out += 1/16 V dw (Pi_gp - ps grad p_k) . sum_a gamma_a
void projection_control_term(pa_plasticity* p, const double* u, const double* pi, const double* gamma,
                             const double* z, double* zc_out, hipStream_t st);

}  // namespace j2
}  // namespace pa
