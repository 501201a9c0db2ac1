/* bsim.h - C ABI of the MI355X-native BabelSim hot path (libbsim.so).
 *
 * Plain pointers and sizes only; every entry point returns 0 on success and a
 * non-zero code on failure, with the message available from bsim_last_error()
 * (thread-local).  Host arrays are in the caller's numbering: cells and faces
 * are renumbered internally and permuted back on every upload/download.
 *
 * Each declaration names the reference interface it replaces.  The reference is
 * BabelSim (C++17): src/internal/compute_backend.h (the linear-system seam) and
 * include/babelsim/{operators,equ,math}.h (the face-operator seam).
 */
#ifndef BSIM_H
#define BSIM_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
const char* bsim_last_error(void);
int bsim_version(void);

/* ------------------------------------------------------- host mesh lists
 * A polyhedral face list exactly as Mesh::polyhedral consumes it
 * (include/babelsim/mesh.h:116-121 PolyhedralFaceSpec; src/core/mesh.cpp:321). */
typedef struct bsim_mesh bsim_mesh;

/* Replaces the reference's external mesh writers (tools/polyhedral_mesh.py) for
 * synthetic cavities: hex lattice, jitter, seeded prism splits. */
int bsim_mesh_generate_hex(int nx, int ny, int nz, const double* lo, const double* hi, const double* vertices,
                           double jitter, double split_fraction, uint64_t seed, bsim_mesh** out);
/* Wraps caller arrays (Mesh::polyhedral inputs, mesh.h:131-145). */
int bsim_mesh_from_arrays(int nv, const double* vertices, int nf, const int* face_point_offsets,
                          const int* face_point_ids, const int* owner, const int* neighbour, const int* patch,
                          int npatch, bsim_mesh** out);
/* Cell renumbering (reverse Cuthill-McKee) and face ordering used in HBM; no
 * reference counterpart (the reference keeps input order). */
int bsim_mesh_prepare(const bsim_mesh* in, int renumber, bsim_mesh** out, int* cell_new_to_old,
                      int* face_new_to_old);
void bsim_mesh_sizes(const bsim_mesh* m, int* nv, int* nf, int* nfp, int* npatch, int* ncells);
void bsim_mesh_arrays(const bsim_mesh* m, double* vertices, int* face_point_offsets, int* face_point_ids,
                      int* owner, int* neighbour, int* patch);
void bsim_mesh_free(bsim_mesh* m);

/* ------------------------------------------------------ configuration */
enum { BSIM_FIXED_VALUE = 0, BSIM_FIXED_GRADIENT = 1, BSIM_ZERO_GRADIENT = 2, BSIM_INLET_OUTLET = 3,
       BSIM_SYMMETRY = 4 };                                   /* field.h:24-33 BoundaryType */
enum { BSIM_LINEAR = 0, BSIM_CORRECTED = 1 };                 /* methods.h:10-13 */
enum { BSIM_GREEN_GAUSS = 0, BSIM_LEAST_SQUARES = 1 };        /* methods.h:15-18 */
enum { BSIM_UPWIND = 0, BSIM_LINEAR_UPWIND = 1, BSIM_CENTRAL = 2 };  /* methods.h:20-24 */
enum { BSIM_ORTHOGONAL = 0, BSIM_DIFF_CORRECTED = 1, BSIM_LIMITED_CORRECTED = 2 };  /* methods.h:26-30 */
enum { BSIM_CG = 0, BSIM_BICGSTAB = 1 };                      /* solver_control.h:14-19 */
enum { BSIM_PC_NONE = 0, BSIM_PC_JACOBI = 1, BSIM_PC_AMG = 2 };/* solver_control.h:21-29 (hypre/gamg -> AMG) */
enum { BSIM_FP64 = 0, BSIM_MIXED = 1 };
enum { BSIM_CONVERGED = 0, BSIM_MAX_ITERATIONS = 1, BSIM_NUMERICAL_FAILURE = 2 };  /* solver_control.h:30-34 */

/* LinearSolverConfig (solver_control.h:165-187) plus the aggregation-AMG and
 * precision controls of this backend. */
typedef struct {
    int solver, pc;
    double atol, rtol;
    int max_iter, warm_start, rebuild_interval;
    double reuse_growth;
    int amg_max_levels, amg_coarse_size, amg_pre_sweeps, amg_post_sweeps;
    double amg_omega, amg_strength, amg_correction;
    int precision;      /* BSIM_FP64 or BSIM_MIXED (FP32 AMG / FP32 inner Krylov) */
    double inner_rtol;  /* mixed BiCGSTAB: relative reduction per FP32 refinement sweep */
} bsim_solver_config;

/* SolveResult (solver_control.h:211-249). */
typedef struct {
    int status, iterations;
    double initial, final_residual, relative;
} bsim_solve_result;

#ifdef __cplusplus
}
#endif
#endif
