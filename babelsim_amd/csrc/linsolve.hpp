// Linear systems behind BabelSim's ComputeBackend seam
// (src/internal/compute_backend.h:30-31): CSR matrices in HBM, CG and
// BiCGSTAB following PETSc's KSPSolve_CG / KSPSolve_BCGS recurrences, Jacobi
// and aggregation-AMG preconditioners, and the true-residual / preconditioner
// reuse contract of src/backend/petsc_linear_system.cpp:257-481.
#pragma once

#include "amg_host.hpp"
#include "device_common.hpp"

#include <map>
#include <memory>

namespace bsim {

// Restart length of GMRES and FGMRES (PETSc's default for KSPGMRES/KSPFGMRES).
constexpr int kGmRestart = 30;

struct SolverCfg {
    // solver: 0 CG, 1 BiCGSTAB, 2 GMRES, 3 FGMRES (LinearSolverType order);
    // pc: 0 none, 1 Jacobi, 2 aggregation AMG, 3 block Jacobi (ILU(0) per 64 rows),
    // 4 IC(0) of the whole matrix (icc), 5 ILU(0) of the whole matrix (asm on one rank)
    int solver = 1, pc = 2;
    double atol = 1e-12, rtol = 1e-8;
    int maxIter = 1000, warmStart = 0, rebuildInterval = 8;
    double reuseGrowth = 2.0;
    int amgMaxLevels = 12, amgCoarseSize = 48, amgPre = 2, amgPost = 2;
    double amgOmega = 0.8, amgStrength = 0.25, amgCorrection = 1.0;
    int precision = 0;
    double innerRtol = 1e-3;
    bool operator==(const SolverCfg& o) const;
};

enum SolveStatus : int { kConverged = 0, kMaxIterations = 1, kNumericalFailure = 2 };
struct SolveRes {
    int status = kNumericalFailure, iterations = 0;
    double initial = 0, final_ = 0, relative = 0;
};

// Matrix with the diagonal stored apart from the off-diagonal entries, which
// are either row-major CSR (ptr) or a sliced layout (sliceOff, K; see RowRange
// in device_common.hpp; padding entries carry column -1).  In the sliced
// layout, storage slot q holds row perm[q] (perm == nullptr: row q).
template <class T> struct CsrView {
    int n;
    const int* ptr;
    const int* col;
    const T* diag;
    const T* val;
    const int* sliceOff;
    int K;
    const int* perm;
};

// Storage pattern of a linear system's off-diagonal values.
struct Pattern {
    int n = 0, nnzStore = 0, K = 0;
    const int* ptr = nullptr;       // device CSR row pointers, or nullptr when sliced
    const int* col = nullptr;       // device columns in storage order
    const int* sliceOff = nullptr;  // device slice offsets when sliced
    std::vector<int> hptr, hcol;    // host row-major CSR
    std::vector<int> csrToStore;    // row-major entry -> storage position (empty: identity)
    std::vector<int> order;         // each row's position in the caller's numbering (empty: identity);
                                    // incomplete factorisations eliminate in that order
};

class AmgDevice;
class LevelIlu;

// HIP-event timing of one kernel's launches: each timed launch is bracketed
// by a pair of events recorded on the launching stream.  The events are
// created up front so that recording allocates nothing.
class KernelProbe {
public:
    ~KernelProbe();
    void start(int capacity);        // creates the event pool, clears the count
    void begin(hipStream_t s);       // before a launch (no-op when stopped or full)
    void end(hipStream_t s);         // after it
    double stop(int& launches);      // waits, stops, returns the summed seconds
    bool on = false;
    double bytesPerLaunch = 0.0;     // algorithmic bytes of the probed kernel
private:
    std::vector<hipEvent_t> ev;
    int used = 0, pending = 0;
};

// EquationPerformanceCounters (include/babelsim/solver_control.h:36-60) of one
// persistent system, counted as src/backend/petsc_backend.cpp:236-286 does:
// a preconditioner reuse per right-hand side solved with a lagged PC, a true
// residual check per Krylov restart or accepted initial guess (the warm-start
// initial residual itself is not one), solve seconds inside the Krylov calls
// only.  Device time is measured on the host around work that ends in a
// stream synchronisation.
struct SystemCounters {
    std::uint64_t matrixUpdates = 0, rhsOnly = 0, pcSetups = 0, pcReuses = 0, retries = 0, solves = 0,
                  krylovIterations = 0, trueResidualChecks = 0, spmvs = 0, pcApplications = 0;
    double setupSeconds = 0, solveSeconds = 0, matrixUpdateSeconds = 0, rhsSeconds = 0, residualSeconds = 0,
           patternSeconds = 0;
    int lastStatus = -1;  // -1 before the first solve
    double lastInitial = 0, lastFinal = 0, lastRelative = 0;
    void last(int status, double initial, double fin, double relative) {
        lastStatus = status; lastInitial = initial; lastFinal = fin; lastRelative = relative;
    }
};

// Process-wide switch between graph replay and stream launches of AMG cycles.
void setGraphReplay(bool on);
bool graphReplay();

// Device memory a system can run in without allocating: the caller's matrix
// (diagonal and off-diagonal values in the pattern's storage order) and a
// scratch region of at least workspaceDoubles(n) doubles, 256-byte aligned,
// for the Krylov vectors, reduction partials and device scalars.  A system
// built on it runs FP64 CG or BiCGSTAB with no preconditioner or Jacobi;
// anything that would allocate (AMG, mixed precision, the lockstep vector
// solve, SpMV timing) throws instead.
struct SystemWorkspace {
    const double* diag;
    const double* val;
    double* scratch;
    std::size_t doubles;
};

class DeviceSystem {
public:
    DeviceSystem(hipStream_t stream, Pattern pattern);
    DeviceSystem(hipStream_t stream, Pattern pattern, const SystemWorkspace& ws);
    ~DeviceSystem();
    static std::size_t workspaceDoubles(int n);
    int n, nnz;  // rows, stored off-diagonal entries
    // Assembly writes the candidate matrix here; update() compares it with the
    // effective matrix and swaps it in when any value differs.
    double* candDiag() { return diagBuf[cand].p; }
    double* candVal() { return valBuf[cand].p; }
    const double* diag() const { return diagBuf[1 - cand].p; }
    const double* val() const { return valBuf[1 - cand].p; }
    bool update();                          // petsc_linear_system.cpp:257-282
    void prepare(const SolverCfg& cfg);     // petsc_linear_system.cpp:284-301
    SolveRes solve(const double* b, double* x, int rhsIndex);  // petsc_linear_system.cpp:355-481
    // The three right-hand sides of a vector equation (petsc_backend.cpp:156-198):
    // the results, histories and counters of solve() on components 0, 1, 2 in
    // order, with BiCGSTAB (Jacobi or no preconditioner, FP64) run in lockstep.
    void solve3(const double* const b[3], double* const x[3], SolveRes res[3], std::vector<double> hists[3]);
    void spmv(const double* x, double* y);   // y = A x with the effective matrix
    // Mean seconds of the solver's SpMV kernel (HIP events).  cold: each launch
    // follows a read-only pass over a buffer twice the size of the 256 MB
    // Infinity Cache, so the matrix and vectors come from HBM rather than from
    // that cache.
    double timeSpmv(int reps, bool cold = false);
    std::vector<double> hist;               // unpreconditioned residual norm per Krylov iteration
    SystemCounters counters;
    SolverCfg cfg;
    CsrView<double> view() const { return {n, pat.ptr, pat.col, diag(), val(), pat.sliceOff, pat.K, nullptr}; }
    hipStream_t stream;
    Pattern pat;

private:
    void setup(bool reuse);
    void applyPC(const double* r, double* z);
    // ||b - A x||; counted as a true-residual check unless `check` is false.
    double trueResidual(const double* b, const double* x, bool check = true);
    void cg(const double* b, double* x, bool guessZero, int maxIt, double target, int& its, int& reason);
    void bcgs(const double* b, double* x, bool guessZero, int maxIt, double target, int& its, int& reason);
    void gmres(const double* b, double* x, bool guessZero, int maxIt, double target, int& its, int& reason,
               bool flexible);
    double readScalar(int idx, int* flags = nullptr);
    // The prelude and first Krylov call of one component, already executed.
    struct FirstRun {
        double rhsNorm = 0, initial = 0, fin = 0;
        int its = 0, reason = 0;
        std::vector<double> hist;
        SystemCounters delta;
    };
    SolveRes solveFrom(const double* b, double* x, int rhsIndex, const FirstRun* first);
    // Work vectors of the lockstep solver, one per component.
    template <class T> struct Work3 {
        DBuf<T> r[3], p[3], v[3], t[3], s[3], rp[3], tmp[3], guess[3];
        void ensure(int n);
    };
    template <class T>
    void bcgs3(Work3<T>& w, const CsrView<T>& A, const T* jv, const T* const b[3], T* const x[3], const bool (&run)[3],
               bool guessZero, int maxIt, const double (&target)[3], int (&its)[3], int (&reason)[3],
               std::vector<double> (&h)[3], SystemCounters (&dl)[3]);
    void solve3Mixed(const double* const b[3], double* const x[3], SolveRes res[3], std::vector<double> hists[3],
                     int ncomp);
    Work3<double> w64;
    Work3<float> w32;
    DBuf<float> d32[3], r32[3], jinv32;
    std::uint64_t f32Stamp = ~0ull;  // matrix/PC state the FP32 copies were made from
    void syncScalars3();
    void uploadAct(const bool (&m)[3]);
    DBuf<double> saved3[3];     // initial guesses of the three components
    DBuf<double> sc3;           // 3 x S_COUNT device scalars
    DBuf<int> fl3, act3;
    double* hsc3 = nullptr;     // pinned: 3 x S_COUNT scalars then 3 flags
    static constexpr int kActSlots = 16;
    int* hact3 = nullptr;       // pinned ring of active masks (kActSlots x 4 ints)
    unsigned actSlot = 0;
    DBuf<double> diagBuf[2], valBuf[2];
    int cand = 0;
    bool ready = false, configured = false, setupPending = false, forceRebuild = false;
    std::uint64_t pcMatrixUpdate = 0, pcRhsSolves = 0;
    int freshIterations[3] = {0, 0, 0};
    DBuf<double> jinv;          // Jacobi: 1/diag of the matrix at PC setup
    DBuf<double> iluDiag, iluVal;  // block Jacobi: ILU(0) factors of the 64-row blocks at PC setup
    std::unique_ptr<LevelIlu> lilu;  // icc / asm: level-scheduled factors of the whole matrix
    std::unique_ptr<AmgDevice> amg;
public:
    KernelProbe probe;  // times the fine-level AMG smoother sweeps
private:
    DBuf<double> r, z, p, w, rp, v, t, s, tmp, guess, saved;
    DBuf<double> gmV, gmZ, gmPart, gmH;  // GMRES basis, FGMRES preconditioned basis, partials, inner products
    double* gmHost = nullptr;            // pinned mirror of gmH
    DBuf<float> diag32, val32;  // FP32 copy of the effective matrix (mixed precision)
    DBuf<double> sc;            // device scalars
    DBuf<int> flags;
    DBuf<double> partials;      // per-block reduction partials
    double* hostScalars = nullptr;  // pinned mirror (or hostLocal on a workspace)
    double hostLocal[32] = {};
    bool borrowed = false;      // runs in a SystemWorkspace
};

// Aggregation AMG V-cycle in FP64 or FP32 with an FP64 dense coarsest solve.
class AmgDevice {
public:
    AmgDevice(hipStream_t s) : stream(s) {}
    ~AmgDevice();
    AmgDevice(const AmgDevice&) = delete;
    AmgDevice& operator=(const AmgDevice&) = delete;
    void setup(const CsrView<double>& A, const Pattern& pattern, const SolverCfg& cfg);
    // z = M^{-1} r (FP64 in/out).  With zrPartials, the FP64 cycle's last
    // fine-level sweep also stores the per-block partials of z . r there and
    // apply() returns true.
    bool apply(const double* r, double* z, double* zrPartials = nullptr);
    int levels() const { return static_cast<int>(L.size()); }
    std::vector<int> sizes() const;
    std::vector<int> aggregates0() const { return hostAgg0; }
    std::vector<double> coarseDiag(int level) const;
    double setupSeconds = 0.0;

    KernelProbe* probe = nullptr;
    double* dotPartials = nullptr;  // set by apply() for the fused last sweep

private:
    struct Level {
        int n = 0, nnz = 0, nc = 0, K = 0;
        const int* ptr = nullptr;
        const int* col = nullptr;
        const int* sliceOff = nullptr;
        const int* perm = nullptr;
        DBuf<int> ptrOwn, colOwn, sliceOwn, permOwn, agg, memPtr, memIdx, gPtr, gSrc;
        DBuf<double> diag, val;
        DBuf<float> diag32, val32;
        DBuf<double> x, b, r, t;
        DBuf<float> x32, b32, r32, t32;
    };
    // Levels from tailStart down run in one single-workgroup kernel (they are
    // small enough that launches, not bytes, would set their time).
    template <class T> void cycle(int l, const T* b, T* xOut, bool firstDone);
    template <class T> void tail();
    template <class T> void levelArrays(const Level& lv, const T*& diag, const T*& val, T*& x, T*& b, T*& r,
                                        T*& t) const;
    int tailStart = 0;
    // One V-cycle's launches run from the stream directly or, when no probe
    // is timing kernels, as a captured graph replayed per application.
    void run(const double* r, double* z, double* zrPartials);
    struct GraphKey {
        const double* r;
        const double* z;
        const double* partials;
        bool operator<(const GraphKey& o) const {
            return r != o.r ? r < o.r : z != o.z ? z < o.z : partials < o.partials;
        }
    };
    struct Captured {
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::map<GraphKey, Captured> graphs;
    void dropGraphs();
    hipStream_t stream;
    std::vector<std::unique_ptr<Level>> L;
    DBuf<double> lu, inv;
    DBuf<int> piv, luFlag;
    bool aggregated = false;
    bool mixed = false;
    SolverCfg cfg;
    std::vector<int> hostAgg0;
};

}  // namespace bsim
