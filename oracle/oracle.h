/* Oracle C ABI (test infrastructure only; loaded by tests/ through ctypes).
 * The struct layouts match include/bsim.h so one set of Python structures
 * drives both the oracle and the HIP library. */
#ifndef BSIM_ORACLE_H
#define BSIM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int solver, pc;
    double atol, rtol;
    int max_iter, warm_start, rebuild_interval;
    double reuse_growth;
    int amg_max_levels, amg_coarse_size, amg_pre_sweeps, amg_post_sweeps;
    double amg_omega, amg_strength, amg_correction;
    int precision;           /* ignored by the oracle (always FP64) */
    double inner_rtol;       /* ignored by the oracle */
} orc_solver_config;

typedef struct {
    int status, iterations;
    double initial, final_residual, relative;
} orc_solve_result;

typedef struct {
    double rho, mu;
    int m_interp, m_grad, m_conv, m_diff, m_coef_interp, m_coef_grad;
    int p_interp, p_grad, p_diff, p_coef_interp, p_coef_grad;
    double alpha_u, alpha_p;
    int non_orth_corr;
    double tol_mass, tol_u, tol_mom, tol_p;
    orc_solver_config mom, pre;
} orc_simple_config;

typedef struct {
    double rU, dU, dP, mass, linP;
    int mom_iters[3], p_iters, mom_status, p_status, converged, referenced;
    double d_turb, r_turb, bounded_omega;
    int k_iters, w_iters, turb_linear, turb_converged;
    int numerical_failure;
} orc_iter_report;

typedef struct {
    double beta_star, beta, gamma, sigma_k, sigma_omega, k_min, omega_min, viscosity_ratio_max, relaxation,
        tolerance;
    int k_interp, k_grad, k_conv, k_diff, k_coef_interp, k_coef_grad;
    int w_interp, w_grad, w_conv, w_diff, w_coef_interp, w_coef_grad;
    orc_solver_config k_solver, w_solver;
} orc_komega_config;

const char* orc_last_error(void);
/* OpenMP threads the oracle's parallel loops use. */
int orc_threads(void);
void* orc_mesh_new(int nv, const double* vertices, int nf, const int* fpo, const int* fpi, const int* owner,
                   const int* neighbour, const int* patch, int npatch);
void orc_mesh_free(void* mesh);
void orc_mesh_sizes(void* mesh, int* nc, int* nf, int* orthogonal, int* nfp);
void orc_mesh_cells(void* mesh, double* centres, double* volumes);
void orc_mesh_faces(void* mesh, double* cf, double* sf, double* af, double* nrm, double* w, double* orth,
                    double* kf, double* skew);
void orc_mesh_rings(void* mesh, int* fpo, int* fpi);

int orc_gradient(void* mesh, int ncomp, const double* x, const int* bct, const double* bcv, const double* flux_ctx,
                 int method, double* out);
int orc_gradient_traces(void* mesh, const double* x, const int* bct, const double* bcv, const double* grad,
                        double* traces);
int orc_interpolate(void* mesh, int ncomp, const double* x, const int* bct, const double* bcv,
                    const double* flux_ctx, int method, int gmethod, double* out);
int orc_flux(void* mesh, const double* U, const int* bct, const double* bcv, const double* flux_ctx, int method,
             int gmethod, double* out);
int orc_normal_gradient(void* mesh, const double* x, const int* bct, const double* bcv, int gmethod, int dmethod,
                        double* out);
/* k_f times integratedNormalGradient with the field's gradient (operators.cpp:1004-1024). */
int orc_diffusion_flux(void* mesh, const double* x, const int* bct, const double* bcv, const double* flux_ctx,
                       const double* kf, int gmethod, int dmethod, double* out);
int orc_divergence_face(void* mesh, const double* phi, double* out);
int orc_explicit_laplacian(void* mesh, const double* x, const int* bct, const double* bcv, const double* kf,
                           double k, int gmethod, int dmethod, double* out);
int orc_explicit_convection(void* mesh, int ncomp, const double* x, const int* bct, const double* bcv,
                            const double* phi, int conv, int interp, int gmethod, double* out);
int orc_assemble_convection(void* mesh, int ncomp, const double* x, const int* bct, const double* bcv,
                            const double* phi, int conv, int interp, int gmethod, double scale, double* diag,
                            double* upper, double* lower, double* source);
int orc_assemble_diffusion(void* mesh, int ncomp, const double* x, const int* bct, const double* bcv,
                           const double* flux_ctx, const double* kf, int gmethod, int dmethod, double* diag,
                           double* upper, double* lower, double* source);

int orc_csr_solve(int n, const int* ptr, const int* col, const double* diag, const double* val, const double* b,
                  double* x, const orc_solver_config* cfg, orc_solve_result* res, double* hist, int hist_cap,
                  int* hist_len);
/* One persistent system on a mesh, fed LDU arrays (ComputeBackend::solve,
 * petsc_backend.cpp:123-198): nrhs 1 or 3 (sources and unknowns interleaved
 * per cell); x holds the guess (warm start) and receives the solution. */
void* orc_system_new(void* mesh);
void orc_system_free(void* sys);
int orc_system_solve_ldu(void* sys, const double* diag, const double* upper, const double* lower, int nrhs,
                         const double* source, double* x, const orc_solver_config* cfg, orc_solve_result* results);
/* matrix updates, rhs-only solves, PC setups, PC reuses, retries, solves,
 * Krylov iterations, true-residual checks; last status, initial, final and
 * relative residual. */
int orc_system_counters(void* sys, uint64_t* out8, int* last_status, double* last3);
/* Block Jacobi (ILU(0) in 64-row blocks) of a CSR matrix: the factors (U
 * diagonal fd[n], factor at each entry fv[nnz]) and z = (LU)^-1 r. */
int orc_block_ilu(int n, const int* ptr, const int* col, const double* diag, const double* val, double* fd,
                  double* fv);
/* IC(0) (chol != 0) or ILU(0) of the whole matrix: the factor pattern
 * (fptr[n + 1], fcol[nf], columns ascending, the diagonal included) and the
 * factors f (L strictly below the diagonal, U on and above it). */
int orc_whole_ilu(int n, const int* ptr, const int* col, const double* diag, const double* val, int chol,
                  int* fptr, int* fcol, double* f, int cap, int* nf, int* failed);
int orc_block_ilu_apply(int n, const int* ptr, const int* col, const double* diag, const double* fd,
                        const double* fv, const double* r, double* z);
int orc_amg_hierarchy(int n, const int* ptr, const int* col, const double* diag, const double* val,
                      const orc_solver_config* cfg, int* nlevels, int* sizes, int* agg0, double* coarse_diag1);

void* orc_simple_new(void* mesh, const orc_simple_config* cfg, const int* ubct, const double* ubcv,
                     const int* pbct, const double* pbcv);
void orc_simple_free(void* s);
int orc_simple_set_fields(void* s, const double* U, const double* p);
int orc_simple_start(void* s);
int orc_simple_iterate(void* s, int n, orc_iter_report* reports);
int orc_simple_get_fields(void* s, double* U, double* p, double* phi);
/* Counters of the momentum (0) or pressure (1) system, as orc_system_counters. */
int orc_simple_counters(void* s, int which, uint64_t* out8);
int orc_piso_counters(void* s, int which, uint64_t* out8);
int orc_simple_history(void* s, int which, double* out, int cap, int* len);
/* Wilcox k-omega for the SIMPLE driver: k and omega boundary conditions and
 * initial values; before start(). */
void* orc_simple_new_komega(void* mesh, const orc_simple_config* cfg, const orc_komega_config* kw, const int* ubct,
                            const double* ubcv, const int* pbct, const double* pbcv, const int* kbct,
                            const double* kbcv, const int* wbct, const double* wbcv);
int orc_simple_set_turbulence(void* s, const double* k, const double* omega);
int orc_simple_get_turbulence(void* s, double* k, double* omega, double* mut);
/* One solveTransport of a fresh model on given U, phi, k, omega (the set-up
 * of tests/rans_equations_test.cpp:17-48); out4 = rOmega, rK, dOmega, dK. */
int orc_komega_transport(void* mesh, const orc_komega_config* kw, double rho, double mu, int m_grad,
                         const double* U, const int* ubct, const double* ubcv, const double* phi,
                         const double* k0, const int* kbct, const double* kbcv, const double* w0, const int* wbct,
                         const double* wbcv, double* k, double* omega, double* mut, double* out4);
/* Operator-level pieces of the model, for known answers. */
int orc_strain_squared(const double* g9, double* out);
int orc_stress_divergence(void* mesh, const double* U, const int* bct, const double* bcv, const double* ctx,
                          const double* coef, const double* coef_trace, int interp, int gmethod, double* out);

typedef struct {
    double rho, mu;
    int m_interp, m_grad, m_conv, m_diff, m_coef_interp, m_coef_grad;
    int p_interp, p_grad, p_diff, p_coef_interp, p_coef_grad;
    int time_method;                       /* 1 Euler, 2 BDF2 */
    int max_iterations, n_correctors, max_correctors, non_orth_corr;
    double alpha_u, coupling_tol, tol_mass, tol_u, tol_mom, tol_p;
    orc_solver_config mom, pre;
    int initialize_potential_flow, initial_non_orth_corr;
} orc_piso_config;

typedef struct {
    double initial_mass, initial_u_max;
    int initial_pressure_solves, p_iters, done, status;
} orc_piso_initial;

typedef struct {
    double dt, co_max, co_mean, mass, dU, rU, dP, r_coupling_first, r_coupling, lin_u, lin_p;
    int passes, correctors, mom_iters[3], p_iters, accepted, converged, linear_ok, status;
    double d_turb, r_turb;
    int k_iters, w_iters, turb_linear, turb_converged, turb_corrections;
} orc_step_report;

/* turbulenceCoupling (0 iterated, 1 segregated), turbulenceCorrectors (<= 0:
 * the reference's default) and turbulenceMaxIterations. */
typedef struct {
    int coupling, correctors, max_iterations;
} orc_piso_turbulence;

void* orc_piso_new(void* mesh, const orc_piso_config* cfg, const int* ubct, const double* ubcv, const int* pbct,
                   const double* pbcv);
void orc_piso_free(void* s);
int orc_piso_set_fields(void* s, const double* U, const double* p);
int orc_piso_enable_komega(void* s, const orc_komega_config* kw, const int* kbct, const double* kbcv,
                           const int* wbct, const double* wbcv, const orc_piso_turbulence* t);
int orc_piso_set_turbulence(void* s, const double* k, const double* omega);
int orc_piso_get_turbulence(void* s, double* k, double* omega, double* mut);
int orc_piso_start(void* s);
int orc_piso_initial_report(void* s, orc_piso_initial* out);
int orc_piso_step(void* s, double dt, orc_step_report* report);
int orc_piso_get_fields(void* s, double* U, double* p, double* phi);
/* rho * temporal flux correction of the last step (nf values). */
int orc_piso_temporal_flux(void* s, double* out);
/* which: 0..2 momentum components, 3 pressure (last solve). */
int orc_piso_history(void* s, int which, double* out, int cap, int* len);
/* Time-term weights a0, a1, a2 and ddtPhiCorr weights (previous, older). */
int orc_time_weights(int method, int levels, double dt, double previous_dt, double* out5);

#ifdef __cplusplus
}
#endif
#endif
