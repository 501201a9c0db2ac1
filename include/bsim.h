/* bsim.h - C ABI of the MI355X-native BabelSim hot path (libbsim.so).
 *
 * Plain pointers and sizes only; every entry point returns 0 on success, 2 for
 * an invalid argument (std::invalid_argument in the reference) and 1 for any
 * other failure (std::runtime_error / std::logic_error), with the message
 * available from bsim_last_error() (thread-local).  Host arrays are in the caller's numbering: cells and faces
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
/* No counterpart in the reference: whether AMG V-cycles replay a captured HIP
 * graph (1, the default) or launch their kernels one by one from the stream
 * (0).  Both give the same bits; stream launches are for tools that trace
 * single kernels.  Applies to every context of the process. */
int bsim_set_graph_replay(int enable);

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

/* ------------------------------------------------------------ file formats
 * BABELSIM_MESH 3 text: readMeshFile (src/io/mesh_reader.cpp:44-111), with its
 * topology checks; the writer emits the same format.  Patch roles are the
 * file's kinds (include/babelsim/mesh.h:79-86; "mirror" reads as symmetry). */
enum { BSIM_PATCH_GENERIC = 0, BSIM_PATCH_WALL = 1, BSIM_PATCH_INLET = 2, BSIM_PATCH_OUTLET = 3,
       BSIM_PATCH_SYMMETRY = 4, BSIM_PATCH_PROCESSOR = 5 };
int bsim_mesh_read(const char* path, bsim_mesh** out);
int bsim_mesh_write(const bsim_mesh* m, const char* path);
/* Name (NUL-terminated, truncated to cap) and role of a patch; meshes not read
 * from a file report "patch<k>" and BSIM_PATCH_GENERIC. */
int bsim_mesh_patch(const bsim_mesh* m, int patch, char* name, int cap, int* kind);
/* Result output of one rank (src/parallel/parallel_writer.cpp:14-75):
 * <time_dir>/rank-NNNN/<name>.csv with the header global_id,x,y,z,value0..
 * and 17 significant digits.  global_ids may be NULL (ids 0..n-1, the
 * caller's cell order); centres are 3 per cell; ncomp is 1, 3 or 9. */
int bsim_write_field_csv(const char* time_dir, int rank, const char* name, int ncells, const int* global_ids,
                         const double* centres, int ncomp, const double* values);
/* metadata.bs and mesh.geometry of one rank (parallel_writer.cpp:116-169);
 * field types are the reference's words (scalar, vector, tensor) and
 * locations 0 cell, 1 face, 2 vertex. */
int bsim_write_result_metadata(const char* time_dir, int rank, int ranks, const bsim_mesh* m, const int* global_ids,
                               long long global_cells, const char* time_name, int nfields, const char* const* names,
                               const char* const* types, const int* locations);
/* Nonuniform cell values for a restart (src/io/field_reader.cpp:143-241):
 * a directory of rank-NNNN subdirectories holding <name>.csv, one .csv file, or a text file of
 * "id value..." rows with # comments.  Every id in [0, ncells) exactly once;
 * out holds ncomp values per cell in id order. */
int bsim_read_field_values(const char* path, const char* name, int ncells, int ncomp, double* out);

/* ------------------------------------------------------ configuration */
enum { BSIM_FIXED_VALUE = 0, BSIM_FIXED_GRADIENT = 1, BSIM_ZERO_GRADIENT = 2, BSIM_INLET_OUTLET = 3,
       BSIM_SYMMETRY = 4 };                                   /* field.h:24-33 BoundaryType */
enum { BSIM_LINEAR = 0, BSIM_CORRECTED = 1 };                 /* methods.h:10-13 */
enum { BSIM_GREEN_GAUSS = 0, BSIM_LEAST_SQUARES = 1 };        /* methods.h:15-18 */
enum { BSIM_UPWIND = 0, BSIM_LINEAR_UPWIND = 1, BSIM_CENTRAL = 2 };  /* methods.h:20-24 */
enum { BSIM_ORTHOGONAL = 0, BSIM_DIFF_CORRECTED = 1, BSIM_LIMITED_CORRECTED = 2 };  /* methods.h:26-30 */
/* solver_control.h:13-18.  GMRES and FGMRES restart every 30 iterations
 * (PETSc's default) and precondition from the right. */
enum { BSIM_CG = 0, BSIM_BICGSTAB = 1, BSIM_GMRES = 2, BSIM_FGMRES = 3 };
/* solver_control.h:21-29: hypre/gamg -> AMG; bjacobi -> ILU(0) in blocks of
 * 64 consecutive rows (PCBJACOBI with ceil(n/64) blocks); icc -> IC(0) of the
 * whole matrix from its upper triangle (PCICC); asm -> ILU(0) of the whole
 * matrix (PCASM on one rank: one subdomain, sub-PC ILU(0)).  The last three
 * are FP64 only; icc and asm run level-scheduled substitutions. */
enum { BSIM_PC_NONE = 0, BSIM_PC_JACOBI = 1, BSIM_PC_AMG = 2, BSIM_PC_BJACOBI = 3, BSIM_PC_ICC = 4,
       BSIM_PC_ASM = 5 };
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

/* ------------------------------------------------------------- contexts
 * One device, one stream, the mesh resident in HBM.  Replaces
 * detail::makeComputeBackend(const Mesh&, ParallelContext)
 * (src/internal/compute_backend.h:39-41) together with the geometry of
 * Mesh::polyhedral (src/core/mesh.cpp:321-578), computed on the GPU. */
typedef struct bsim_ctx bsim_ctx;
int bsim_ctx_create(const bsim_mesh* mesh, int renumber, int device, bsim_ctx** out);
void bsim_ctx_destroy(bsim_ctx* ctx);
int bsim_ctx_sizes(bsim_ctx* ctx, int* ncells, int* nfaces, int* ninterior, int* nnz_off, int* orthogonal);
/* Mesh::Storage arrays (include/babelsim/mesh.h:138-177) in caller numbering. */
int bsim_ctx_geometry(bsim_ctx* ctx, double* centres, double* volumes, double* face_centres, double* sf,
                      double* areas, double* normals, double* owner_weights, double* orth, double* non_orth,
                      double* skew);

/* ------------------------------------------------------------ operators
 * Host-array forms of the face operators (include/babelsim/operators.h:9-186),
 * one field at a time: values (ncomp per cell), per-patch boundary type and
 * value (BoundaryCondition, field.h:34-55) and an optional face flux context
 * for inletOutlet (Field::setBoundaryFlux, field.h:277-282). */
int bsim_op_gradient(bsim_ctx* ctx, int ncomp, const double* x, const int* bc_type, const double* bc_value,
                     const double* flux_ctx, int method, double* out);                 /* operators.h gradient */
int bsim_op_interpolate(bsim_ctx* ctx, int ncomp, const double* x, const int* bc_type, const double* bc_value,
                        const double* flux_ctx, int method, int gmethod, double* out);  /* operators.h interpolate */
int bsim_op_flux(bsim_ctx* ctx, const double* U, const int* bc_type, const double* bc_value, const double* flux_ctx,
                 int method, int gmethod, double* out);                                 /* operators.h flux */
int bsim_op_normal_gradient(bsim_ctx* ctx, const double* x, const int* bc_type, const double* bc_value,
                            int gmethod, int dmethod, double* out);  /* fvm_execution.cpp:522-543 */
int bsim_op_divergence_face(bsim_ctx* ctx, const double* phi, double* out);             /* operators.h divergence */
int bsim_op_explicit_laplacian(bsim_ctx* ctx, const double* x, const int* bc_type, const double* bc_value,
                               const double* kf, double k, int gmethod, int dmethod, double* out); /* laplacian */
int bsim_op_explicit_convection(bsim_ctx* ctx, int ncomp, const double* x, const int* bc_type, const double* bc_value,
                                const double* phi, int conv, int interp, int gmethod, double* out); /* convection */
/* Implicit terms into an LDU system (operators.h addConvection/addDiffusion);
 * the outputs are accumulated into the caller's diag/upper/lower/source. */
int bsim_op_assemble_convection(bsim_ctx* ctx, int ncomp, const double* x, const int* bc_type, const double* bc_value,
                                const double* phi, int conv, int interp, int gmethod, double scale, double* diag,
                                double* upper, double* lower, double* source);
int bsim_op_assemble_diffusion(bsim_ctx* ctx, int ncomp, const double* x, const int* bc_type, const double* bc_value,
                               const double* flux_ctx, const double* kf, int gmethod, int dmethod, double* diag,
                               double* upper, double* lower, double* source);

/* ------------------------------------------------------- linear systems
 * One persistent system per equation identity: ComputeBackend::solve for a
 * scalar (nrhs = 1) or vector (nrhs = 3, sources and unknowns interleaved per
 * cell like Vec3) DiscreteEquation (compute_backend.h:30-31,
 * src/backend/petsc_backend.cpp:123-198). */
typedef struct bsim_system bsim_system;
int bsim_system_create(bsim_ctx* ctx, bsim_system** out);
int bsim_system_solve_ldu(bsim_system* sys, const double* diag, const double* upper, const double* lower, int nrhs,
                          const double* source, double* x, const bsim_solver_config* cfg, bsim_solve_result* results);
/* Residual norm history of the last right-hand side solved. */
int bsim_system_history(bsim_system* sys, double* out, int cap, int* len);
/* PerformanceCounters subset (solver_control.h:36-163): matrix updates,
 * rhs-only solves, PC setups, PC reuses, retries, solves, Krylov iterations,
 * true-residual checks, SpMVs, PC applications. */
int bsim_system_counters(bsim_system* sys, uint64_t* out10);
/* EquationPerformanceCounters (include/babelsim/solver_control.h:36-60) of one
 * persistent system, counted as src/backend/petsc_backend.cpp:236-286 counts
 * them (a preconditioner reuse per right-hand side solved with a lagged
 * preconditioner; a true-residual check per Krylov restart or accepted initial
 * guess; solve seconds inside the Krylov calls only), plus the system's SpMVs
 * and preconditioner applications.  ksp_type, pc_type and last_status use the
 * reference's words ("cg", "bcgs"; "none", "jacobi", "amg" for the aggregation
 * AMG; "converged", "maxIterations", "numericalFailure", empty before the first
 * solve).  Seconds are host wall time around device work that ends in a
 * stream synchronisation.  The identity is set by the solver drivers below
 * ("vector:momentum:U", "scalar:pressureCorrection:pPrime", ...) and empty
 * for a stand-alone system, whose identity the caller holds. */
typedef struct {
    char identity[64], ksp_type[16], pc_type[16], last_status[24];
    uint64_t matrix_pattern_builds, matrix_value_updates, rhs_only_solves, linear_solves, krylov_iterations,
        preconditioner_setups, preconditioner_reuses, preconditioner_retries, true_residual_checks, sparse_matvecs,
        preconditioner_applications;
    double last_initial_residual, last_final_residual, last_relative_residual;
    double pattern_build_seconds, matrix_update_seconds, preconditioner_seconds, rhs_seconds, solve_seconds,
        residual_check_seconds;
} bsim_equation_counters;
int bsim_system_performance(bsim_system* sys, bsim_equation_counters* out);

/* PerformanceCounters (include/babelsim/solver_control.h:62-107) of a solver
 * run: the 31 values of Case::performance (src/io/case.cpp:72-104) in that
 * order, summed over the run's systems.  On one device there are no halo
 * exchanges, output writes or MPI reductions: global_reductions counts the
 * device reductions the host reads (the convergence measures and residual
 * norms the reference reduces over ranks), assembly_seconds the face-loop
 * assembly phases, and the SpMV and preconditioner-application seconds are not
 * measured apart from the solves (0).  reported_iteration_intervals and
 * iteration_compute_seconds follow RunTime::recordIteration
 * (src/runtime/runtime.cpp:117-130): the window opens at the end of the first
 * outer iteration or time step. */
typedef struct {
    uint64_t linear_solves, krylov_iterations, sparse_matvecs, halo_exchanges, halo_bytes, global_reductions,
        equation_assemblies, preconditioner_setups, preconditioner_applications, output_writes;
    double elapsed_seconds, assembly_seconds, preconditioner_seconds, preconditioner_apply_seconds,
        linear_solve_seconds, sparse_matvec_seconds, halo_seconds, global_reduction_seconds, output_seconds;
    uint64_t matrix_pattern_builds, matrix_value_updates, rhs_only_solves, true_residual_checks;
    double matrix_update_seconds, rhs_seconds, residual_check_seconds, pattern_build_seconds;
    uint64_t preconditioner_reuses, preconditioner_retries, reported_iteration_intervals;
    double iteration_compute_seconds;
} bsim_performance;
void bsim_system_destroy(bsim_system* sys);

/* Stand-alone CSR solve / SpMV / AMG hierarchy (diagonal stored apart from the
 * off-diagonal entries); the minimum slice of the seam above. */
int bsim_csr_solve(int device, int n, const int* ptr, const int* col, const double* diag, const double* val,
                   const double* b, double* x, const bsim_solver_config* cfg, bsim_solve_result* res, double* hist,
                   int cap, int* len);
int bsim_csr_spmv(int device, int n, const int* ptr, const int* col, const double* diag, const double* val,
                  const double* x, double* y);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code between HIP graph captures, tight serving loops, harnesses that
 * forbid hipMalloc): bsim_csr_solve on caller-owned DEVICE arrays (ptr n+1,
 * col and val nnz, diag, b and x n), enqueued on `stream` (a hipStream_t; NULL
 * is the null stream), with the same results, history and status.  It
 * allocates nothing: its Krylov vectors, reduction partials and scalars live in
 * a fixed device array of the library (49.5 MiB, held on each device it runs
 * on until the process ends).  Limits: n <= BSIM_NOALLOC_MAX_ROWS, FP64 CG or
 * BiCGSTAB with BSIM_PC_NONE or BSIM_PC_JACOBI (larger systems, AMG and mixed
 * precision return 2).  Calls on one device share that array; the library
 * serialises them and each returns after its stream has drained, so calls
 * that overlap on one device wait for each other rather than race. */
#define BSIM_NOALLOC_MAX_ROWS 524288
int bsim_csr_solve_noalloc(int device, void* stream, int n, const int* ptr, const int* col, const double* diag,
                           const double* val, const double* b, double* x, const bsim_solver_config* cfg,
                           bsim_solve_result* res, double* hist, int cap, int* len);
int bsim_amg_hierarchy(int device, int n, const int* ptr, const int* col, const double* diag, const double* val,
                       const bsim_solver_config* cfg, int* nlevels, int* sizes, int* agg0, double* coarse_diag1);

/* ------------------------------------------------ fields and equations
 * The procedural equation seam resident on a context, so that a solver
 * written against equ:: and math:: (include/babelsim/equ.h:16-91,
 * include/babelsim/math.h:259-267, src/internal/fvm_execution.h:38-49) keeps
 * its fields and matrices in HBM: assembly writes the equation's own matrix
 * and a solve hands it to the equation's persistent linear system.
 *
 * bsim_field: Field<T> (include/babelsim/field.h:57-374).  location 0 is a
 * cell field of ncomp 1 or 3 values per cell with a boundary condition per
 * patch (zero gradient until set), location 1 a face field of one value per
 * face.  Values cross the ABI in the caller's numbering. */
typedef struct bsim_field bsim_field;
enum { BSIM_CELL = 0, BSIM_FACE = 1 };
enum { BSIM_STEADY = 0 }; /* with BSIM_EULER and BSIM_BDF2 below: TimeMethod */
int bsim_field_create(bsim_ctx* ctx, int location, int ncomp, const double* values, bsim_field** out);
void bsim_field_destroy(bsim_field* f);
int bsim_field_set(bsim_field* f, const double* values);
int bsim_field_get(const bsim_field* f, double* values);
/* Field::setBoundary per patch: npatch types and ncomp values each. */
int bsim_field_set_boundary(bsim_field* f, const int* bc_type, const double* bc_value);
/* Field::setBoundaryFlux (field.h:277-282): the inletOutlet direction context. */
int bsim_field_set_flux(bsim_field* f, const bsim_field* face_flux);

/* bsim_equ: equ::Equation<T> bound to a cell field, with its operator options
 * (methods.h) and linear solver configuration (EquationControl). */
typedef struct bsim_equ bsim_equ;
typedef struct {
    int interp, grad, conv, diff, coef_interp, coef_grad;
} bsim_equ_methods;
int bsim_equ_create(bsim_field* unknown, const bsim_equ_methods* methods, const bsim_solver_config* linear,
                    bsim_equ** out);
void bsim_equ_destroy(bsim_equ* a);
int bsim_equ_reset(bsim_equ* a);                                         /* Equation::reset */
int bsim_equ_copy(const bsim_equ* a, bsim_equ** out);                    /* Equation::copy */
int bsim_equ_add(bsim_equ* a, const bsim_equ* b, double factor);         /* equ::add */
int bsim_equ_scale(bsim_equ* a, double factor);                          /* equ::scale */
int bsim_equ_div(bsim_equ* a, const bsim_field* flux, double factor);    /* equ::div */
/* equ::laplacian: coefficient NULL for the constant k, else a cell or face field. */
int bsim_equ_laplacian(bsim_equ* a, const bsim_field* coefficient, double k, double multiplier);
/* equ::reaction: coefficient NULL for the constant value, else factor * field. */
int bsim_equ_reaction(bsim_equ* a, const bsim_field* coefficient, double value, double factor);
/* equ::source: value NULL for the constant (ncomp values), else factor * field. */
int bsim_equ_source(bsim_equ* a, const bsim_field* value, const double* constant, double factor);
/* equ::ddt: capacity NULL for the constant; method BSIM_STEADY, BSIM_EULER or
 * BSIM_BDF2 (older required; previous_dt 0 means dt). */
int bsim_equ_ddt(bsim_equ* a, const bsim_field* capacity, double cap, const bsim_field* previous, double dt, int method,
                 const bsim_field* older, double previous_dt);
int bsim_equ_relax(bsim_equ* a, const bsim_field* previous, double alpha);        /* equ::relax */
int bsim_equ_reference(bsim_equ* a, int cell, double value);                       /* Equation::reference */
int bsim_equ_reference_if_unanchored(bsim_equ* a, double value, int* applied);     /* referenceIfUnanchored */
int bsim_equ_diagonal(const bsim_equ* a, bsim_field* out);                         /* Equation::diagonal */
int bsim_equ_rhs(const bsim_equ* a, bsim_field* out);                              /* Equation::rhs */
int bsim_equ_apply(const bsim_equ* a, const bsim_field* x, bsim_field* out);       /* equ::apply */
int bsim_equ_residual(const bsim_equ* a, const bsim_field* x, bsim_field* out);    /* equ::residual, b - A x */
int bsim_equ_face_flux(const bsim_equ* a, const bsim_field* solution, bsim_field* out);  /* equ::faceFlux */
int bsim_equ_relative_residual(const bsim_equ* a, const bsim_field* x, double* out);    /* diagnostics:: */
/* equ::solve: the unknown is solved in place; cfg NULL uses the equation's
 * configuration; results receives ncomp entries. */
int bsim_equ_solve(bsim_equ* a, const bsim_solver_config* cfg, bsim_solve_result* results);
int bsim_equ_performance(const bsim_equ* a, bsim_equation_counters* out);

/* math::normsL2 (field_math.cpp:30-62) of n fields. */
int bsim_math_norms(const bsim_field* const* fields, int n, double* out);
/* FvmExecution::relativeChange / relativeMagnitude (fvm_execution.cpp:91-146). */
int bsim_math_relative_change(const bsim_field* current, const bsim_field* previous, double* out);
int bsim_math_relative_magnitude(const bsim_field* value, const bsim_field* reference, double* out);
/* fluxBalance (fvm_execution.cpp:148-175): l1, l2, maximum, relative. */
int bsim_math_flux_balance(const bsim_field* flux, double* out4);
/* courantNumber (fvm_execution.cpp:177-199): maximum, mean. */
int bsim_math_courant(const bsim_field* flux, double dt, double* out2);
/* math::add / subtract (field_math.cpp:65-82): target += factor * increment on
 * all faces (region 0) or interior faces (region 1). */
int bsim_math_face_add(const bsim_field* increment, bsim_field* target, double factor, int region);

/* ----------------------------------------------------------------- SIMPLE
 * Steady SIMPLE (src/physics/simple/main.cpp:11-131) resident on the GPU. */
typedef struct {
    double rho, mu;
    int m_interp, m_grad, m_conv, m_diff, m_coef_interp, m_coef_grad;
    int p_interp, p_grad, p_diff, p_coef_interp, p_coef_grad;
    double alpha_u, alpha_p;
    int non_orth_corr;
    double tol_mass, tol_u, tol_mom, tol_p;
    bsim_solver_config mom, pre;
} bsim_simple_config;

/* One SIMPLE outer iteration (main.cpp:114-125 report): the turbulence
 * entries are dTurb, rTurb (max over omega and k), the omega cells the last
 * bounding raised, Krylov iterations of the k and omega solves, and whether
 * they converged and the model is settled (zero / one when laminar). */
typedef struct {
    double rU, dU, dP, mass, linP;
    int mom_iters[3], p_iters, mom_status, p_status, converged, referenced;
    double d_turb, r_turb, bounded_omega;
    int k_iters, w_iters, turb_linear, turb_converged;
    /* runSimple's numericalFailure (main.cpp:61-62, 67, 88, 106, 115-116): a
     * non-finite measure or an unhealthy velocity, pressure or turbulence
     * solve ends the iteration at that point and the run. */
    int numerical_failure;
} bsim_iter_report;

/* Wilcox (1988) k-omega (src/physics/RANS/k_omega.cpp:25-157): constants,
 * bounds, relaxation and tolerance as the constructor reads them (33-48), and
 * the kTransport / omegaTransport equations' methods and linear solvers. */
typedef struct {
    double beta_star, beta, gamma, sigma_k, sigma_omega, k_min, omega_min, viscosity_ratio_max, relaxation,
        tolerance;
    int k_interp, k_grad, k_conv, k_diff, k_coef_interp, k_coef_grad;
    int w_interp, w_grad, w_conv, w_diff, w_coef_interp, w_coef_grad;
    bsim_solver_config k_solver, w_solver;
} bsim_komega_config;

typedef struct bsim_simple bsim_simple;
int bsim_simple_create(bsim_ctx* ctx, const bsim_simple_config* cfg, const int* u_bc_type, const double* u_bc_value,
                       const int* p_bc_type, const double* p_bc_value, bsim_simple** out);
int bsim_simple_set_fields(bsim_simple* s, const double* U, const double* p);
/* turbulenceModel kOmega for this SIMPLE run (model.cpp:35-40), with the k and
 * omega boundary conditions; call before bsim_simple_start.  The momentum
 * diffusion then uses muEff = mu + mut and gains the explicit stress
 * remainder (main.cpp:53-59); each iteration ends with the omega and k
 * solves (main.cpp:100-110). */
int bsim_simple_enable_komega(bsim_simple* s, const bsim_komega_config* cfg, const int* k_bc_type,
                              const double* k_bc_value, const int* w_bc_type, const double* w_bc_value);
int bsim_simple_set_turbulence(bsim_simple* s, const double* k, const double* omega);
int bsim_simple_get_turbulence(bsim_simple* s, double* k, double* omega, double* mut);
int bsim_simple_start(bsim_simple* s);
/* Runs up to n iterations; it stops after one that ends in numerical failure,
 * and the reports after it are zero with numerical_failure = 1. */
int bsim_simple_iterate(bsim_simple* s, int n, bsim_iter_report* reports);
int bsim_simple_get_fields(bsim_simple* s, double* U, double* p, double* phi);
/* which: 0..2 momentum components, 3 pressure correction (last solve), 4 k and
 * 5 omega (last transport solves, k-omega only). */
int bsim_simple_history(bsim_simple* s, int which, double* out, int cap, int* len);
/* Seconds per phase: assembly, momentum solve, Rhie-Chow, pressure assembly,
 * pressure solve, correction, convergence, iterations, turbulence transport. */
int bsim_simple_timing(bsim_simple* s, double* out9);
/* The run's PerformanceCounters and its systems' counters (momentum,
 * pressure correction, then k and omega with k-omega); count receives the
 * number of systems, at most cap are written. */
int bsim_simple_performance(bsim_simple* s, bsim_performance* total, bsim_equation_counters* systems, int cap,
                            int* count);
/* Times `reps` launches of the solver's SpMV on the current momentum (which
 * 0) or pressure (1) matrix with HIP events on the context stream; returns
 * mean seconds and the matrix size.  cold != 0: every launch follows a
 * read-only pass over a 512 MiB buffer (twice the 256 MB Infinity Cache), so its operands
 * are read from HBM; cold == 0: back-to-back launches, which may find part of
 * a matrix up to that size in the cache. */
int bsim_simple_time_spmv(bsim_simple* s, int which, int reps, int cold, double* seconds, int* n, int* nnz_off);
/* Diagnostics with no counterpart in the reference: HIP-event timing of the
 * fine-level AMG smoother sweeps of the pressure solve.  start() creates a
 * pool of `capacity` event pairs and times every later sweep launch (up to
 * the pool size) by events recorded on the context stream around it; stop()
 * waits for them and returns the launch count, the summed kernel seconds and
 * the kernel's algorithmic bytes per launch. */
int bsim_simple_probe_start(bsim_simple* s, int capacity);
int bsim_simple_probe_stop(bsim_simple* s, int* launches, double* seconds, double* bytes_per_launch);
void bsim_simple_destroy(bsim_simple* s);

/* ------------------------------------------------------------------- PISO
 * Transient PISO (src/physics/piso/main.cpp:66-302, laminar) resident on the
 * GPU.  The controls are runPiso's solution dictionary (main.cpp:83-105):
 * maxIterations, nCorrectors, maxCorrectors (0 means nCorrectors),
 * nonOrthogonalCorrections, velocityRelaxation, couplingTolerance and the
 * continuity / velocity / momentum / pressure tolerances; time_method is the
 * methods dictionary's time scheme.  The run loop and its TimeStepper
 * (src/io/case.cpp:388-403) stay with the caller: each bsim_piso_step call is
 * one `time.advance()` of length dt followed by the body of the while loop
 * (main.cpp:161-299).  Field writing (main.cpp:298) stays with the caller too. */
enum { BSIM_EULER = 1, BSIM_BDF2 = 2 };
/* Step outcomes, SolverResult's completed / notConverged / numericalFailure
 * (include/babelsim/solver.h).  A step that is not completed ends the run in
 * the reference; the fields then hold the rejected step's values. */
enum { BSIM_STEP_COMPLETED = 0, BSIM_STEP_NOT_CONVERGED = 1, BSIM_STEP_NUMERICAL_FAILURE = 2 };
typedef struct {
    double rho, mu;
    int m_interp, m_grad, m_conv, m_diff, m_coef_interp, m_coef_grad;
    int p_interp, p_grad, p_diff, p_coef_interp, p_coef_grad;
    int time_method;
    int max_iterations, n_correctors, max_correctors, non_orth_corr;
    double alpha_u, coupling_tol, tol_mass, tol_u, tol_mom, tol_p;
    bsim_solver_config mom, pre;
    /* initializePotentialFlow (0 or 1) and initialNonOrthogonalCorrections
     * (negative: non_orth_corr), main.cpp:104-105: the start projects the
     * initial velocity onto a conservative face flux (main.cpp:127-150). */
    int initialize_potential_flow, initial_non_orth_corr;
} bsim_piso_config;

/* What the initial projection recorded (main.cpp:147-149): initialMass,
 * initialUmax, initialPressureSolves, plus its Krylov iterations, whether it
 * ran, and its outcome (BSIM_STEP_COMPLETED, or BSIM_STEP_NUMERICAL_FAILURE
 * when a solve did not converge; every later step then reports that). */
typedef struct {
    double initial_mass, initial_u_max;
    int initial_pressure_solves, p_iters, done, status;
} bsim_piso_initial;

/* The per-step record runPiso reports (main.cpp:282-289): CoMax, CoMean,
 * mass, dU, rU, dP, rCouplingFirst, rCoupling, linU, linP, the outer passes
 * and corrector passes used, Krylov iterations, accepted / converged /
 * linear-ok flags and the step outcome. */
typedef struct {
    double dt, co_max, co_mean, mass, dU, rU, dP, r_coupling_first, r_coupling, lin_u, lin_p;
    int passes, correctors, mom_iters[3], p_iters, accepted, converged, linear_ok, status;
    /* With k-omega (main.cpp:251-266, 290-293): dTurb and rTurb of the last
     * transport correction, the k and omega Krylov iterations over all of
     * them, the linear and settled flags and the corrections run. */
    double d_turb, r_turb;
    int k_iters, w_iters, turb_linear, turb_converged, turb_corrections;
} bsim_step_report;

/* turbulenceCoupling (0 iterated, 1 segregated), turbulenceCorrectors (<= 0:
 * the reference's default) and turbulenceMaxIterations (main.cpp:98-120). */
typedef struct {
    int coupling, correctors, max_iterations;
} bsim_piso_turbulence;

typedef struct bsim_piso bsim_piso;
int bsim_piso_create(bsim_ctx* ctx, const bsim_piso_config* cfg, const int* u_bc_type, const double* u_bc_value,
                     const int* p_bc_type, const double* p_bc_value, bsim_piso** out);
int bsim_piso_set_fields(bsim_piso* s, const double* U, const double* p);
/* Wilcox k-omega in transient form, before the start: k and omega histories
 * with their ddt terms (k_omega.cpp:59-62, 82, 96), the eddy viscosity and
 * stress remainder in the momentum predictor, and the transport corrections
 * after the pressure correctors (main.cpp:98-121, 165, 176-182, 251-266). */
int bsim_piso_enable_komega(bsim_piso* s, const bsim_komega_config* k, const int* k_bc_type, const double* k_bc_value,
                            const int* w_bc_type, const double* w_bc_value, const bsim_piso_turbulence* t);
int bsim_piso_set_turbulence(bsim_piso* s, const double* k, const double* omega);
int bsim_piso_get_turbulence(bsim_piso* s, double* k, double* omega, double* mut);
/* phi = flux(U) and the time histories from the current fields (main.cpp:123, 157-158). */
int bsim_piso_start(bsim_piso* s);
/* The initial projection's record (zeros with done = 0 when it is not configured). */
int bsim_piso_initial_report(bsim_piso* s, bsim_piso_initial* out);
int bsim_piso_step(bsim_piso* s, double dt, bsim_step_report* report);
int bsim_piso_get_fields(bsim_piso* s, double* U, double* p, double* phi);
/* rho * temporalFluxCorrection of the last step (main.cpp:19-53, 166-167), one value per face. */
int bsim_piso_temporal_flux(bsim_piso* s, double* out);
/* which: 0..2 momentum components, 3 pressure (last solve). */
int bsim_piso_history(bsim_piso* s, int which, double* out, int cap, int* len);
/* Seconds per phase: history + ddtPhiCorr, momentum assembly, momentum solve,
 * corrector (HbyA, fluxes, corrections, coupling), pressure assembly,
 * pressure solve, diagnostics, steps. */
int bsim_piso_timing(bsim_piso* s, double* out8);
/* As bsim_simple_performance: momentum and pressure systems. */
int bsim_piso_performance(bsim_piso* s, bsim_performance* total, bsim_equation_counters* systems, int cap,
                          int* count);
/* Diagnostics with no counterpart in the reference: HIP-event timing of the
 * pressure solve's fine-level AMG smoother sweeps, as bsim_simple_probe_*. */
int bsim_piso_probe_start(bsim_piso* s, int capacity);
int bsim_piso_probe_stop(bsim_piso* s, int* launches, double* seconds, double* bytes_per_launch);
void bsim_piso_destroy(bsim_piso* s);


/* ---------------------------------------------------------------- serving
 * No counterpart in the reference, where one RunTime is active per thread
 * (src/runtime/runtime.cpp:15, 48-51) and a Case owns one mesh
 * (src/io/case.cpp:280-298).  A server owns `slots` worker threads on one
 * device.  Meshes are registered once as cases; a request names a case and
 * is a steady SIMPLE run to convergence (bsim_simple_* semantics, capped at
 * max_iterations outer iterations) or a transient PISO run of max_iterations
 * steps of length dt (bsim_piso_* semantics).  Requests become eligible at
 * their arrival time on the server clock (seconds since create; earlier
 * times mean at once) and are admitted in arrival order while their device
 * memory estimate fits the budget next to what is running; each runs on its
 * own context and stream, so admitted requests overlap on the device.  A
 * case's contexts (mesh and geometry in HBM) stay resident after a request
 * and are reused by the next request on that case, until their memory is
 * needed (least recently used first).  Results equal those of a solo run of
 * the same request. */
enum { BSIM_REQ_STEADY = 0, BSIM_REQ_TRANSIENT = 1 };
/* Request outcomes: converged (steady) or every step completed (transient);
 * not converged within max_iterations, or a step not converged; numerical
 * failure; rejected (larger than the whole budget); error (message in the
 * result); cancelled (still queued when the server was destroyed). */
enum { BSIM_SERVE_CONVERGED = 0, BSIM_SERVE_NOT_CONVERGED = 1, BSIM_SERVE_NUMERICAL_FAILURE = 2,
       BSIM_SERVE_REJECTED = 3, BSIM_SERVE_ERROR = 4, BSIM_SERVE_CANCELLED = 5 };
typedef struct {
    int kind, case_id;
    const bsim_simple_config* simple;  /* BSIM_REQ_STEADY */
    const bsim_piso_config* piso;      /* BSIM_REQ_TRANSIENT */
    const int* u_bc_type;
    const double* u_bc_value;
    const int* p_bc_type;
    const double* p_bc_value;
    const double* U0;  /* optional initial fields, caller numbering (NULL: zero) */
    const double* p0;
    int max_iterations;
    double dt, arrival;
} bsim_request;
typedef struct {
    int status, iterations, slot, reused_context;
    /* server-clock seconds: eligible, admitted, context ready, finished */
    double released, admitted, started, finished;
    double rU, dU, dP, mass;              /* last iteration or step */
    double reserved_bytes, measured_bytes; /* admission reservation; solver allocations measured */
    char message[256];
} bsim_request_result;
typedef struct bsim_server bsim_server;
int bsim_server_create(int device, int slots, double memory_budget_bytes, bsim_server** out);
/* Registers a mesh (copied); case ids count from 0. */
int bsim_server_add_case(bsim_server* s, const bsim_mesh* mesh, int* case_id);
/* Copies the request's configuration, conditions and fields; returns its id. */
int bsim_server_submit(bsim_server* s, const bsim_request* req, int* id);
/* Blocks until the request is done. */
int bsim_server_wait(bsim_server* s, int id, bsim_request_result* out);
int bsim_server_done(bsim_server* s, int id, int* done);
/* U (3 per cell), p (per cell) and phi (per face) of a finished request, caller numbering. */
int bsim_server_fields(bsim_server* s, int id, double* U, double* p, double* phi);
/* Drops a finished request's record and fields. */
int bsim_server_release(bsim_server* s, int id);
/* budget, peak reserved, reserved, peak running, running, completed,
 * rejected, contexts created, context reuses, evictions, resident idle contexts */
int bsim_server_stats(bsim_server* s, double* out11);
double bsim_server_clock(bsim_server* s);
/* The admission estimates: device bytes of one context (mesh and geometry)
 * and of one request's solver on a mesh of the given sizes.  A case's context
 * is reserved at its measured size once one has been built. */
int bsim_server_estimate(int kind, int ncells, int nfaces, double* context_bytes, double* solver_bytes);
/* Cancels queued requests, waits for running ones, frees every context. */
void bsim_server_destroy(bsim_server* s);

/* The library's device block cache (no counterpart in the reference).
 * Buffers that a served request frees stay cached per size, stream-ordered
 * on the request's context stream, for the next request on the same mesh;
 * at most a quarter of the device's memory is held.  Stats: cached bytes,
 * live bytes, cache hits, runtime allocations, blocks returned to the
 * runtime.  Trim returns every cached block to the runtime. */
int bsim_device_pool_stats(int device, double* out5);
int bsim_device_pool_trim(int device);

#ifdef __cplusplus
}
#endif
#endif
