// Case context: device selection, stream, mesh upload, numbering maps.
#include "context.hpp"

#include "../../include/bsim.h"

namespace bsim {

Context::Context(const MeshDesc& caller, bool renumber, int dev) : device(dev) {
    int count = 0;
    BSIM_HIP(hipGetDeviceCount(&count));
    if (dev < 0 || dev >= count) throw std::invalid_argument("device index is out of range");
    BSIM_HIP(hipSetDevice(dev));
    {
        std::lock_guard<std::recursive_mutex> rl(pool::runtimeLock());
        BSIM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    }
    mesh = std::make_unique<DeviceMesh>(caller, renumber, stream);
    err.alloc(1);
    err.zero(stream);
    cellOldToNew.assign(mesh->nc, 0);
    for (int i = 0; i < mesh->nc; ++i) cellOldToNew[mesh->cellNewToOld[i]] = i;
    BSIM_HIP(hipStreamSynchronize(stream));
}

Context::~Context() {
    // Work still queued on the stream finishes first: a later stream may get
    // the same handle, and the block cache treats same-handle reuse as
    // stream-ordered.
    if (stream) (void)hipStreamSynchronize(stream);
    mesh.reset();
    err.release();
    if (stream) {
        std::lock_guard<std::recursive_mutex> rl(pool::runtimeLock());
        (void)hipStreamDestroy(stream);
    }
}

void Context::checkErrors() {
    const int e = err.toHost(stream, 1)[0];
    if (!e) return;
    err.zero(stream);
    BSIM_HIP(hipStreamSynchronize(stream));
    if (e & kErrNoFluxContext) throw std::invalid_argument("inletOutlet requires a boundary flux context");
    if (e & kErrSingularLs) throw std::runtime_error("least-squares gradient matrix is singular");
    if (e & kErrNegativeDiffusivity) throw std::invalid_argument("face diffusivity must be non-negative and finite");
    throw std::invalid_argument("reference requires a positive diagonal");
}

template <class T> std::vector<T> Context::cellsIn(const T* a, int ncomp) const {
    std::vector<T> v(static_cast<std::size_t>(mesh->nc) * ncomp);
    for (int i = 0; i < mesh->nc; ++i)
        for (int k = 0; k < ncomp; ++k) v[static_cast<std::size_t>(i) * ncomp + k] = a[static_cast<std::size_t>(mesh->cellNewToOld[i]) * ncomp + k];
    return v;
}
template <class T> void Context::cellsOut(const std::vector<T>& v, T* a, int ncomp) const {
    for (int i = 0; i < mesh->nc; ++i)
        for (int k = 0; k < ncomp; ++k) a[static_cast<std::size_t>(mesh->cellNewToOld[i]) * ncomp + k] = v[static_cast<std::size_t>(i) * ncomp + k];
}
template <class T> std::vector<T> Context::facesIn(const T* a, int ncomp) const {
    std::vector<T> v(static_cast<std::size_t>(mesh->nf) * ncomp);
    for (int f = 0; f < mesh->nf; ++f)
        for (int k = 0; k < ncomp; ++k) v[static_cast<std::size_t>(f) * ncomp + k] = a[static_cast<std::size_t>(mesh->faceNewToOld[f]) * ncomp + k];
    return v;
}
template <class T> void Context::facesOut(const std::vector<T>& v, T* a, int ncomp) const {
    for (int f = 0; f < mesh->nf; ++f)
        for (int k = 0; k < ncomp; ++k) a[static_cast<std::size_t>(mesh->faceNewToOld[f]) * ncomp + k] = v[static_cast<std::size_t>(f) * ncomp + k];
}
template std::vector<double> Context::cellsIn(const double*, int) const;
template void Context::cellsOut(const std::vector<double>&, double*, int) const;
template std::vector<double> Context::facesIn(const double*, int) const;
template void Context::facesOut(const std::vector<double>&, double*, int) const;

Pattern meshPattern(const DeviceMesh& M) {
    Pattern P;
    P.n = M.nc;
    P.nnzStore = M.nnzSliced;
    P.K = M.K;
    P.ptr = nullptr;
    P.col = M.mcol.p;
    P.sliceOff = M.sliceOff.p;
    P.hptr = M.hmptr;
    P.hcol = M.hmcol;
    P.csrToStore = M.csrToSliced;
    P.order = M.cellNewToOld;
    return P;
}

SolverCfg toSolverCfg(const void* p) {
    const auto* c = static_cast<const bsim_solver_config*>(p);
    SolverCfg s;
    s.solver = c->solver; s.pc = c->pc; s.atol = c->atol; s.rtol = c->rtol; s.maxIter = c->max_iter;
    s.warmStart = c->warm_start; s.rebuildInterval = c->rebuild_interval; s.reuseGrowth = c->reuse_growth;
    s.amgMaxLevels = c->amg_max_levels; s.amgCoarseSize = c->amg_coarse_size; s.amgPre = c->amg_pre_sweeps;
    s.amgPost = c->amg_post_sweeps; s.amgOmega = c->amg_omega; s.amgStrength = c->amg_strength;
    s.amgCorrection = c->amg_correction; s.precision = c->precision; s.innerRtol = c->inner_rtol;
    // LinearSolverConfig::validate (src/backend/solver_control.cpp:8-45)
    if (!(s.atol > 0.0) || !std::isfinite(s.atol) || !(s.rtol > 0.0) || !std::isfinite(s.rtol) || s.maxIter <= 0)
        throw std::invalid_argument("linear solver tolerances or iteration limit are invalid");
    if (s.solver < 0 || s.solver > 3) throw std::invalid_argument("unsupported Krylov solver type");
    if (s.pc < 0 || s.pc > 5) throw std::invalid_argument("unsupported preconditioner type");
    if (s.pc >= 3 && s.precision != 0)
        throw std::invalid_argument("the block Jacobi, icc and asm preconditioners run in FP64");
    if (s.pc == 2 && (s.amgMaxLevels <= 0 || s.amgCoarseSize <= 0 || s.amgPre <= 0 || s.amgPost < 0))
        throw std::invalid_argument("AMG hierarchy controls must be positive");
    if (s.rebuildInterval < 1 || !std::isfinite(s.reuseGrowth) || s.reuseGrowth < 1.0)
        throw std::invalid_argument("preconditioner rebuild interval must be positive and iteration growth must be finite and >= 1");
    if (s.precision != 0 && s.precision != 1) throw std::invalid_argument("unsupported precision mode");
    return s;
}

}  // namespace bsim
