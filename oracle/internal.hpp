// Oracle: FP64 CPU restatement of BabelSim's SIMPLE hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (babelsim_amd/) links,
// loads or calls this code; only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it, and only as the checker.
//
// Every function names the reference file:line it restates.  The reference is
// BabelSim (C++17, CPU/MPI, FP64).  Parity status per function is listed in
// oracle/README.md.  Parity unpinned (no reference output or known answer fixes
// them exactly): Krylov residual histories against PETSc, the linear-system
// counter sequences, relativeResidual and the convergence measures, and the
// exact first-iteration SIMPLE state on non-orthogonal meshes (the SIMPLE loop
// as a whole is pinned by the Ghia cavity).
#pragma once

#include <cmath>
#include <cstdint>
#include <exception>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

// Runs body(i) for i in [0, n) on OpenMP threads.  Every iteration writes only
// its own entries, so the result does not depend on the thread count; the
// first exception an iteration throws is rethrown after the loop.
template <class F> void parallelFor(std::int32_t n, F&& body) {
    std::exception_ptr err;
#pragma omp parallel for schedule(static)
    for (std::int32_t i = 0; i < n; ++i) {
        try {
            body(i);
        } catch (...) {
#pragma omp critical(orc_parallel_for)
            if (!err) err = std::current_exception();
        }
    }
    if (err) std::rethrow_exception(err);
}

using Index = std::int32_t;
constexpr Index kNone = -1;

struct V3 {
    double x = 0.0, y = 0.0, z = 0.0;
    double& operator[](int i) { return (&x)[i]; }
    double operator[](int i) const { return (&x)[i]; }
};
inline V3 operator+(V3 a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(const V3& a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(double s, const V3& a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(const V3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(const V3& a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3& operator+=(V3& a, const V3& b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline V3& operator-=(V3& a, const V3& b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm2(const V3& a) { return dot(a, a); }
inline double norm(const V3& a) { return std::sqrt(norm2(a)); }

struct T3 {  // rows[c] = gradient of component c
    V3 r[3];
};

// Scalar/vector traits so the operators are written once.
inline double comp(double v, int) { return v; }
inline double comp(const V3& v, int c) { return v[c]; }
inline void setComp(double& v, int, double x) { v = x; }
inline void setComp(V3& v, int c, double x) { v[c] = x; }
template <class T> constexpr int ncomp();
template <> constexpr int ncomp<double>() { return 1; }
template <> constexpr int ncomp<V3>() { return 3; }

// Gradient type: double -> V3, V3 -> T3.
template <class T> struct GradOf;
template <> struct GradOf<double> { using type = V3; };
template <> struct GradOf<V3> { using type = T3; };
inline const V3& gradRow(const V3& g, int) { return g; }
inline const V3& gradRow(const T3& g, int c) { return g.r[c]; }
inline V3& gradRow(V3& g, int) { return g; }
inline V3& gradRow(T3& g, int c) { return g.r[c]; }

enum BcType : int { FixedValue = 0, FixedGradient = 1, ZeroGradient = 2, InletOutlet = 3, Symmetry = 4 };
enum Interp : int { Linear = 0, Corrected = 1 };
enum GradM : int { GreenGauss = 0, LeastSquares = 1 };
enum ConvM : int { Upwind = 0, LinearUpwind = 1, Central = 2 };
enum DiffM : int { Orthogonal = 0, CorrectedDiff = 1, LimitedCorrected = 2 };

// Polyhedral mesh and its FVM factors (restates include/babelsim/mesh.h:138-177
// and src/core/mesh.cpp:321-578).
struct Mesh {
    Index nc = 0, nf = 0, nv = 0, np = 0;
    std::vector<V3> vert;
    std::vector<Index> fpo, fpi;          // oriented face rings (CSR)
    std::vector<Index> own, nei, patch;
    std::vector<V3> Cf, Sf, nrm, kf, skew;
    std::vector<double> Af, w, orth;
    std::vector<V3> C;
    std::vector<double> V, invV;
    std::vector<Index> cfo, cfi, cfs;     // cell -> faces CSR (input face order) and signs
    std::vector<std::vector<Index>> patchFaces;
    // Off-diagonal pattern of a cell operator: rows sorted by (column, face),
    // each entry's face and whether the row is the face's owner.
    std::vector<Index> csrPtr, csrCol, csrFace;
    std::vector<char> csrOwn;
    bool orthogonal = true;
    bool boundary(Index f) const { return nei[f] == kNone; }
};

Mesh buildMesh(const std::vector<V3>& vertices, const std::vector<Index>& fpo,
               const std::vector<Index>& fpi, const std::vector<Index>& owner,
               const std::vector<Index>& neighbour, const std::vector<Index>& patch,
               Index npatch);

template <class T> struct BC {
    int type = ZeroGradient;
    T value{};
};

// A cell field with patch constraints, or a calculated field with per-face
// traces (include/babelsim/field.h:57-374).
template <class T> struct Field {
    const Mesh* m = nullptr;
    std::vector<T> v;
    std::vector<BC<T>> bc;            // per patch
    std::vector<double> fluxCtx;      // inletOutlet direction context (size nf or empty)
    bool calculated = false;
    std::vector<T> trace;             // size nf when calculated
    Field() = default;
    explicit Field(const Mesh& mesh, Index n = -1) : m(&mesh) {
        v.assign(n < 0 ? mesh.nc : n, T{});
        bc.assign(mesh.np, BC<T>{});
    }
};
using SField = Field<double>;
using VField = Field<V3>;

// Face-condition resolution (field.h:285-301).
template <class T> BC<T> faceCondition(const Field<T>& f, Index face,
                                       double flux = std::numeric_limits<double>::quiet_NaN());
// Boundary trace (field.h:302-313).
template <class T> T boundaryTrace(const Field<T>& f, Index face,
                                   double flux = std::numeric_limits<double>::quiet_NaN());

// LDU system (include/babelsim/discrete_equation.h:14-50).
template <class T> struct Ldu {
    std::vector<double> diag, upper, lower;
    std::vector<T> source;
    void init(const Mesh& m) {
        diag.assign(m.nc, 0.0); upper.assign(m.nf, 0.0); lower.assign(m.nf, 0.0);
        source.assign(m.nc, T{});
    }
};

// ---- operators (src/discretization/operators.cpp) ----
template <class T> void gradient(const Field<T>& f, std::vector<typename GradOf<T>::type>& out, int method);
template <class T> void interpolate(const Field<T>& f, std::vector<T>& out, int method, int gmethod);
template <class T> T correctedFaceValue(const Field<T>& f, const typename GradOf<T>::type* g, Index face,
                                        double flux = std::numeric_limits<double>::quiet_NaN());
template <class T> T integratedNormalGradient(const Field<T>& f, const std::vector<typename GradOf<T>::type>& g,
                                              Index face, int dmethod);
void flux(VField& U, std::vector<double>& out, int interp, int gmethod);
void divergenceFace(const Mesh& m, const std::vector<double>& phi, std::vector<double>& out);
void divergenceVector(const VField& U, std::vector<double>& out, int interp, int gmethod);
void normalGradient(const SField& p, const std::vector<V3>* g, int gmethod, int dmethod, std::vector<double>& out);
void diffusionFlux(const std::vector<double>& kf, const SField& x, const std::vector<V3>& g, int dmethod,
                   std::vector<double>& out);
template <class T> void addConvection(Ldu<T>& eq, const std::vector<double>& phi, Field<T>& x, int conv,
                                      int interp, int gmethod, double scale);
template <class T> void addDiffusion(Ldu<T>& eq, const std::vector<double>& kf, const Field<T>& x, int gmethod,
                                     int dmethod, const std::vector<typename GradOf<T>::type>* g);
template <class T> void explicitConvection(const std::vector<double>& phi, const Field<T>& x, std::vector<T>& out,
                                           int conv, int interp, int gmethod);
void explicitLaplacian(const SField& x, const std::vector<double>* kf, double k, int gmethod, int dmethod,
                       std::vector<double>& out);
// Computed-gradient boundary traces (src/discretization/fvm_execution.cpp:221-235).
void gradientTraces(const SField& p, const std::vector<V3>& g, std::vector<V3>& traces);

// ---- linear algebra ----
struct Csr {  // diagonal stored separately; off-diagonal entries one per interior half-face
    Index n = 0;
    std::vector<Index> ptr, col;
    std::vector<double> diag, val;
};

struct SolverConfig {
    int solver = 1;          // 0 CG, 1 BiCGSTAB, 2 GMRES, 3 FGMRES (restart 30, right preconditioning)
    int pc = 2;              // 0 none, 1 Jacobi, 2 AMG, 3 block Jacobi (ILU(0) per 64 rows), 4 IC(0), 5 ILU(0)
    double atol = 1e-12, rtol = 1e-8;
    int maxIter = 1000;
    int warmStart = 0;
    int rebuildInterval = 8;
    double reuseGrowth = 2.0;
    int amgMaxLevels = 12;
    int amgCoarseSize = 48;
    int amgPreSweeps = 2, amgPostSweeps = 2;
    double amgOmega = 0.8;
    double amgStrength = 0.25;
    double amgCorrection = 1.0;   // coarse-correction scaling
};

enum Status : int { Converged = 0, MaxIterations = 1, NumericalFailure = 2 };
struct SolveResult {
    int status = NumericalFailure;
    int iterations = 0;
    double initial = 0, final_ = 0, relative = 0;
};

struct Amg;
struct LinearSystem {  // one persistent system per equation identity (petsc_linear_system.cpp)
    Csr A;
    std::vector<double> prevDiag, prevVal;
    bool ready = false, configured = false, setupPending = false, forceRebuild = false;
    SolverConfig cfg;
    std::uint64_t matrixUpdates = 0, pcMatrixUpdate = 0, pcRhsSolves = 0, pcSetups = 0, rhsOnly = 0, retries = 0;
    // Accounting of src/backend/petsc_backend.cpp:236-286: a reuse per
    // right-hand side solved with a lagged PC, a true-residual check per Krylov
    // restart or accepted initial guess.
    std::uint64_t solves = 0, krylovIterations = 0, pcReuses = 0, trueResidualChecks = 0;
    int lastStatus = -1;
    double lastInitial = 0, lastFinal = 0, lastRelative = 0;
    int freshIterations[3] = {0, 0, 0};
    Amg* amg = nullptr;
    std::vector<double> hist;  // residual history of the last solve (every KSP iteration, all restarts)
    ~LinearSystem();
    void update(const std::vector<double>& diag, const std::vector<double>& val);
    void prepare(const SolverConfig& c);
    void setup(bool reuse);
    SolveResult solve(const std::vector<double>& b, std::vector<double>& x, int rhsIndex);
};

void lduToCsr(const Mesh& m, const std::vector<double>& diag, const std::vector<double>& upper,
              const std::vector<double>& lower, Csr& A);
Csr csrPattern(const Mesh& m);
void spmv(const Csr& A, const double* x, double* y);

// Aggregation AMG (the build's own algorithm, restated here).
struct AmgLevel {
    Csr A;
    std::vector<Index> agg;          // fine -> coarse aggregate (size n), empty on coarsest
    Index nc = 0;
    std::vector<Index> memPtr, memIdx;  // coarse -> fine members, ascending (the restriction's gather)
};
// IC(0) or ILU(0) of the whole matrix with no fill (pc 4 icc, pc 5 asm on
// one rank; petsc_linear_system.cpp:204-208, 224-226), in the rows' natural
// order.  Factor pattern row-major with columns ascending, the diagonal at
// dpos[i]; IC(0) takes the upper triangle and its transpose.
struct WholeIlu {
    bool chol = false, analysed = false, fail = false;
    Index n = 0;
    std::vector<Index> ptr, col, dpos, srcPtr, src, uptr, udst, usrc;
    std::vector<double> f;
    void analyse(const Csr& A, bool cholesky);
    void factor(const Csr& A);
    void apply(const double* r, double* z) const;
};

struct Amg {
    std::vector<AmgLevel> levels;
    std::vector<double> coarseLU;    // dense LU of coarsest, row-major
    std::vector<double> coarseInv;   // its explicit inverse, row-major
    std::vector<Index> coarsePiv;
    SolverConfig cfg;
    std::vector<std::vector<double>> wx, wb, wr;
    bool aggregated = false;  // aggregates are built at the first setup and kept
    std::vector<double> iluDiag, iluVal;  // block Jacobi factors (pc 3), at the matrix's entries
    WholeIlu whole;                       // icc / asm factors (pc 4, 5)
    void setup(const Csr& A, const SolverConfig& c);
    void apply(const double* r, double* z);
    void vcycle(int l);
};
void aggregate(const Csr& A, double theta, std::vector<Index>& agg, Index& nc);
// Block Jacobi: ILU(0) inside blocks of 64 consecutive rows (PCBJACOBI with
// ceil(n/64) blocks and ILU(0) sub-preconditioners).
void blockIluFactor(const Csr& A, std::vector<double>& fd, std::vector<double>& fv);
void blockIluApply(const Csr& A, const std::vector<double>& fd, const std::vector<double>& fv, const double* r,
                   double* z);
void galerkin(const Csr& A, const std::vector<Index>& agg, Index nc, Csr& Ac);

}  // namespace orc
