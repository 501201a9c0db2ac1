// Steady SIMPLE on the GPU (src/physics/simple/main.cpp:47-127).  Each step
// names the reference line it follows; fused kernels keep BabelSim's
// per-element arithmetic (relax-then-scale rows, V/aP, Rhie-Chow face flux,
// frozen pressure flux) while gathering by row instead of scattering by face.
#include "flow_kernels.hpp"
#include "simple.hpp"

#include <chrono>
#include <cstring>

namespace bsim {
namespace {

using namespace dev;
using namespace flow;
using Clock = std::chrono::steady_clock;
double since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

constexpr int kRedK = 6;

// relax (procedural_equation.cpp:298-305), scale (240-247) and rAU = V/aP (main.cpp:62-66).
__global__ void __launch_bounds__(kBlock) kRelaxScale(MeshView m, double* diag, double* off, double* src, const double* Uprev, double alpha,
                            double* rAU) {
    const int i = globalThread();
    if (i >= m.nc) return;
    const double extra = diag[i] * (1 / alpha - 1);
    double d = diag[i] + extra;
    for (int k = 0; k < 3; ++k) {
        const long long j = 3 * static_cast<long long>(i) + k;
        const double s = src[j] + extra * Uprev[j];
        src[j] = alpha * s;
    }
    d *= alpha;
    diag[i] = d;
    for (int k_ = 0, e = 0; m.row(i).valid(k_) && (e = m.row(i).at(k_), true); ++k_) off[e] *= alpha;
    rAU[i] = m.V[i] / d;
}

// X = rAU * gradP with traces rAU_trace * gradP_trace.
__global__ void __launch_bounds__(kBlock) kScaleVector(MeshView m, const double* s, const double* g, double* X, const double* str,
                             const double* gtr, double* Xtr) {
    const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    if (i < m.nc)
        for (int k = 0; k < 3; ++k) X[3 * i + k] = s[i] * g[3 * i + k];
    const long long f = m.nfi + i;
    if (f < m.nf)
        for (int k = 0; k < 3; ++k) Xtr[3 * f + k] = str[f] * gtr[3 * f + k];
}

// Rhie-Chow on interior faces (main.cpp:71-79): phiH += (rAU gradP)_f . Sf,
// phiH += -1 * ((rAU_f * snGrad p) * |Sf|).
__global__ void __launch_bounds__(kBlock) kRhieChow(MeshView m, double* phiH, const double* Xf, const double* rAUf, const double* sn) {
    const int f = globalThread();
    if (f >= m.nfi) return;
    double v = phiH[f] + dot(ld3(Xf, f), ld3(m.Sf, f));
    v += -1.0 * ((rAUf[f] * sn[f]) * m.Af[f]);
    phiH[f] = v;
}

// p += alphaP p'; U -= rAU grad(p')  (main.cpp:95-97)
__global__ void __launch_bounds__(kBlock) kCorrectUP(MeshView m, double* p, const double* pp, double alphaP, double* U, const double* rAU,
                           const double* gpp) {
    const int i = globalThread();
    if (i >= m.nc) return;
    p[i] += 1.0 * (alphaP * pp[i]);
    for (int k = 0; k < 3; ++k) {
        const long long j = 3 * static_cast<long long>(i) + k;
        U[j] += -1.0 * (rAU[i] * gpp[j]);
    }
}

// relativeChange, relativeMagnitude and fluxBalance sums (fvm_execution.cpp:91-177).
__global__ void __launch_bounds__(kBlock) kConvergence(MeshView m, const double* U, const double* Up, const double* pp, const double* p,
                             const double* phi, double* part) {
    __shared__ double smem[kRedK * kBlock / 64];
    double a[kRedK] = {0, 0, 0, 0, 0, 0};
    for (int c = globalThread(); c < m.nc; c += gridDim.x * blockDim.x) {
        const D3 u = ld3(U, c), up = ld3(Up, c);
        const D3 d = u - up;
        a[0] += dot(d, d);
        a[1] += dot(u, u);
        a[2] += pp[c] * pp[c];
        a[3] += p[c] * p[c];
        double imb = 0.0, sc = 0.0;
        for (int k_ = 0, e = 0; m.row(c).valid(k_) && (e = m.row(c).at(k_), true); ++k_) {
            const int f = m.mface[e];
            if (f < 0) continue;
            const double out = m.own[f] == c ? phi[f] : -phi[f];
            imb += out;
            sc += fabs(out);
        }
        for (int e = m.bptr[c]; e < m.bptr[c + 1]; ++e) {
            const double out = phi[m.bface[e]];
            imb += out;
            sc += fabs(out);
        }
        a[4] += fabs(imb);
        a[5] += sc;
    }
    blockPartial<kRedK>(a, part, smem);
}

}  // namespace

SimpleSolver::SimpleSolver(Context& ctx, const SimpleCfg& cfg, const int* ubct, const double* ubcv, const int* pbct,
                           const double* pbcv)
    : C(ctx), c(cfg) {
    const DeviceMesh& M = *C.mesh;
    hipStream_t s = C.stream;
    const int nc = M.nc, nf = M.nf, np = M.np;
    if (!(c.alphaU > 0 && c.alphaU <= 1) || !(c.alphaP > 0 && c.alphaP <= 1))
        throw std::invalid_argument("relaxation must be in (0,1]");
    if (!(c.rho > 0) || !(c.mu > 0)) throw std::invalid_argument("density and viscosity must be positive");
    std::vector<int> ppt(np), zg(np, kZeroGradient);
    std::vector<double> ppv(np, 0.0), zgv(3 * np, 0.0);
    for (int k = 0; k < np; ++k) {
        if (ubct[k] < 0 || ubct[k] > 4 || pbct[k] < 0 || pbct[k] > 4) throw std::invalid_argument("invalid boundary type");
        // field.h:382-409: homogeneous correction boundaries
        ppt[k] = pbct[k] == kFixedValue ? kFixedValue : pbct[k] == kSymmetry ? kSymmetry : kZeroGradient;
    }
    uType.upload(ubct, np, s); uVal.upload(ubcv, 3 * np, s);
    pType.upload(pbct, np, s); pVal.upload(pbcv, np, s);
    ppType.upload(ppt, s); ppVal.upload(ppv, s);
    zgType.upload(zg, s); zgVal.upload(zgv, s);
    for (DBuf<double>* b : {&p, &rAU, &divPhiH, &pPrime, &psrc, &mu}) b->alloc(nc);
    for (int c = 0; c < 3; ++c) { bcomp[c].alloc(nc); xcomp[c].alloc(nc); }
    for (DBuf<double>* b : {&U, &Uprev, &gradP, &gradPP, &X, &src}) b->alloc(3 * static_cast<std::size_t>(nc));
    for (DBuf<double>* b : {&gradU, &work9, &gradX}) b->alloc(9 * static_cast<std::size_t>(nc));
    for (DBuf<double>* b : {&phi, &phiH, &rAUtr, &rAUf, &sn, &kfMom, &kfP, &corr, &ctxSeed}) b->alloc(nf);
    for (DBuf<double>* b : {&gtr, &Xtr, &Xf}) b->alloc(3 * static_cast<std::size_t>(nf));
    faceWork.alloc(6 * static_cast<std::size_t>(std::max(M.nfi, 1)));
    red.alloc(kRedK * kMaxReduceBlocks + 16);
    U.zero(s); p.zero(s); phi.zero(s); corr.zero(s);
    const MeshView mv = M.view();
    momSys = std::make_unique<DeviceSystem>(s, meshPattern(M));
    pSys = std::make_unique<DeviceSystem>(s, meshPattern(M));
    refCell = C.cellOldToNew.at(0);
    // muEffective: constant cell field with default (zeroGradient) patches
    // (src/physics/RANS/model.cpp:58-63), interpolated with the momentum
    // diffusion coefficient scheme once: it never changes for laminar flow.
    kFillD<<<blocksFor(nc), kBlock, 0, s>>>(nc, mu.p, c.mu);
    FieldArg muF{mu.p, zgType.p, zgVal.p, nullptr, nullptr};
    if (c.mCoefInterp == kCorrected) {
        ops::gradient(mv, M.orthogonal, 1, muF, c.mCoefGrad, gradP.p, work9.p, C.err.p, s);
        ops::interpolate(mv, 1, muF, kCorrected, gradP.p, kfMom.p, C.err.p, s);
    } else {
        ops::interpolate(mv, 1, muF, kLinear, nullptr, kfMom.p, C.err.p, s);
    }
    C.checkErrors();
}

void SimpleSolver::enableKOmega(const KOmegaCfg& kw, const int* kbct, const double* kbcv, const int* wbct,
                                const double* wbcv) {
    turb = std::make_unique<KOmegaDevice>(C, kw, c.rho, c.mu, kbct, kbcv, wbct, wbcv);
    started = false;
}

FieldArg SimpleSolver::fieldU(const double* ctx) const { return {U.p, uType.p, uVal.p, ctx, nullptr}; }
FieldArg SimpleSolver::fieldP() const { return {p.p, pType.p, pVal.p, nullptr, nullptr}; }
FieldArg SimpleSolver::fieldPP() const { return {pPrime.p, ppType.p, ppVal.p, nullptr, nullptr}; }
FieldArg SimpleSolver::fieldRAU() const { return {rAU.p, zgType.p, zgVal.p, nullptr, rAUtr.p}; }

void SimpleSolver::setFields(const double* Uc, const double* pc) {
    hipStream_t s = C.stream;
    if (Uc) { auto v = C.cellsIn(Uc, 3); U.upload(v, s); }
    if (pc) { auto v = C.cellsIn(pc, 1); p.upload(v, s); }
    BSIM_HIP(hipStreamSynchronize(s));
    started = false;
}

void SimpleSolver::getFields(double* Uc, double* pc, double* phic) {
    hipStream_t s = C.stream;
    if (Uc) C.cellsOut(U.toHost(s), Uc, 3);
    if (pc) C.cellsOut(p.toHost(s), pc, 1);
    if (phic) C.facesOut(phi.toHost(s), phic, 1);
}

double SimpleSolver::sync() {
    BSIM_HIP(hipStreamSynchronize(C.stream));
    return 0.0;
}

void SimpleSolver::start() {
    // main.cpp:41: phi = flux(U); the first use seeds inletOutlet directions
    // from owner velocities (operators.cpp:977-987).
    const DeviceMesh& M = *C.mesh;
    const MeshView mv = M.view();
    hipStream_t s = C.stream;
    if (turb) turb->start();  // the model exists before phi (main.cpp:33, 41)
    ops::seedFluxContext(mv, U.p, ctxSeed.p, s);
    const double* g = nullptr;
    if (c.mInterp == kCorrected) { ops::gradient(mv, M.orthogonal, 3, fieldU(ctxSeed.p), c.mGrad, gradU.p, work9.p, C.err.p, s); g = gradU.p; }
    ops::flux(mv, fieldU(ctxSeed.p), g, phi.p, C.err.p, s);
    C.checkErrors();
    started = true;
}

IterRep SimpleSolver::iterate() {
    if (!started) start();
    struct Record {  // RunTime::recordIteration at every way out of the iteration
        RunClock& r;
        ~Record() { r.record(); }
    } record{run};
    const DeviceMesh& M = *C.mesh;
    const MeshView mv = M.view();
    hipStream_t s = C.stream;
    const int nc = M.nc, nf = M.nf;
    const int nb = blocksFor(nc);
    double* part = red.p + 16;
    IterRep rep;
    auto t0 = Clock::now();
    // U_previous = U (main.cpp:48)
    kCopyD<<<blocksFor(3LL * nc), kBlock, 0, s>>>(3LL * nc, U.p, Uprev.p);
    // Momentum assembly (main.cpp:51-57): convection with U's flux context = phi.
    double* diag = momSys->candDiag();
    double* off = momSys->candVal();
    kFillD<<<nb, kBlock, 0, s>>>(nc, diag, 0.0);
    kFillD<<<blocksFor(std::max(M.nnzSliced, 1)), kBlock, 0, s>>>(M.nnzSliced, off, 0.0);
    kFillD<<<blocksFor(3LL * nc), kBlock, 0, s>>>(3LL * nc, src.p, 0.0);
    const FieldArg Uf = fieldU(phi.p);
    const bool needGradU = c.mConv == kLinearUpwind || (c.mConv == kCentral && c.mInterp == kCorrected) ||
                           c.mDiff != kOrthogonal || turb;
    if (needGradU) ops::gradient(mv, M.orthogonal, 3, Uf, c.mGrad, gradU.p, work9.p, C.err.p, s);
    if (turb) {
        // laplacian(muEff) with the current eddy viscosity (main.cpp:53)
        const FieldArg mf{turb->muEff(), zgType.p, zgVal.p, nullptr, turb->muEffTrace()};
        const double* gm = nullptr;
        if (c.mCoefInterp == kCorrected) { ops::gradient(mv, M.orthogonal, 1, mf, c.mCoefGrad, gradPP.p, work9.p, C.err.p, s); gm = gradPP.p; }
        ops::interpolate(mv, 1, mf, c.mCoefInterp, gm, kfMom.p, C.err.p, s);
    }
    // div(phi, U) then laplacian(mu, U) (main.cpp:51-57) in one face-ordered pass.
    ops::assembleTransport(mv, 3, Uf, phi.p, c.mConv, c.mInterp, c.rho, kfMom.p, c.mDiff, -(-1.0), gradU.p, diag, off,
                           src.p, faceWork.p, C.err.p, s);
    ops::gradient(mv, M.orthogonal, 1, fieldP(), c.pGrad, gradP.p, work9.p, C.err.p, s);
    kMomentumPressureSource<<<nb, kBlock, 0, s>>>(mv, gradP.p, src.p);
    if (turb) turb->addStressSource(Uf, gradU.p, c.mInterp, c.mGrad, src.p);  // main.cpp:57-59
    // rU before relaxation (main.cpp:60)
    kMomentumResidual<<<reduceBlocks(nc), kBlock, 0, s>>>(mv, diag, off, src.p, U.p, part);
    kReduceTail<3><<<1, kTail, 0, s>>>(reduceBlocks(nc), part, StoreSums<3>{red.p});
    kRelaxScale<<<nb, kBlock, 0, s>>>(mv, diag, off, src.p, Uprev.p, c.alphaU, rAU.p);
    if (nf > M.nfi) kOwnerTrace<<<blocksFor(nf - M.nfi), kBlock, 0, s>>>(mv, rAU.p, rAUtr.p);
    BSIM_LAUNCH_CHECK();
    double hr[3];
    BSIM_HIP(hipMemcpyAsync(hr, red.p, sizeof(hr), hipMemcpyDeviceToHost, s));
    ++run.reductions;
    C.checkErrors();
    rep.rU = std::sqrt(hr[0]) / std::max(std::sqrt(hr[1]) + std::sqrt(hr[2]), 1e-30);
    timing.assembly += since(t0);
    if (!std::isfinite(rep.rU)) {  // main.cpp:61
        rep.failed = 1;
        return rep;
    }
    t0 = Clock::now();
    // Momentum solve: one matrix, three right-hand sides (petsc_backend.cpp:156-198).
    momSys->update();
    momSys->prepare(c.mom);
    rep.momStatus = kConverged;
    kSplit<<<nb, kBlock, 0, s>>>(nc, src.p, bcomp[0].p, bcomp[1].p, bcomp[2].p);
    kSplit<<<nb, kBlock, 0, s>>>(nc, U.p, xcomp[0].p, xcomp[1].p, xcomp[2].p);
    {
        const double* const b3[3] = {bcomp[0].p, bcomp[1].p, bcomp[2].p};
        double* const x3[3] = {xcomp[0].p, xcomp[1].p, xcomp[2].p};
        SolveRes r3[3];
        momSys->solve3(b3, x3, r3, momHist);
        for (int comp = 0; comp < 3; ++comp) {
            const SolveRes& r = r3[comp];
            rep.momIters[comp] = r.iterations;
            if (r.status == kNumericalFailure) rep.momStatus = kNumericalFailure;
            else if (r.status != kConverged && rep.momStatus != kNumericalFailure) rep.momStatus = kMaxIterations;
        }
    }
    kMerge<<<nb, kBlock, 0, s>>>(nc, xcomp[0].p, xcomp[1].p, xcomp[2].p, U.p);
    BSIM_LAUNCH_CHECK();
    timing.momentumSolve += since(t0);
    if (rep.momStatus == kNumericalFailure) {  // main.cpp:67
        rep.failed = 1;
        return rep;
    }
    t0 = Clock::now();
    // Rhie-Chow (main.cpp:71-79).  grad(p) is unchanged since assembly.
    const double* gU = nullptr;
    if (c.mInterp == kCorrected) { ops::gradient(mv, M.orthogonal, 3, Uf, c.mGrad, gradU.p, work9.p, C.err.p, s); gU = gradU.p; }
    ops::flux(mv, Uf, gU, phiH.p, C.err.p, s);
    ops::gradientTraces(mv, fieldP(), gradP.p, gtr.p, s);
    kScaleVector<<<blocksFor(std::max(nc, nf - M.nfi)), kBlock, 0, s>>>(mv, rAU.p, gradP.p, X.p, rAUtr.p, gtr.p, Xtr.p);
    {
        const FieldArg Xa{X.p, zgType.p, zgVal.p, nullptr, Xtr.p};
        const double* gX = nullptr;
        if (c.pInterp == kCorrected) { ops::gradient(mv, M.orthogonal, 3, Xa, c.pGrad, gradX.p, work9.p, C.err.p, s); gX = gradX.p; }
        ops::interpolate(mv, 3, Xa, c.pInterp, gX, Xf.p, C.err.p, s);
    }
    {
        const double* gR = nullptr;
        if (c.pCoefInterp == kCorrected) { ops::gradient(mv, M.orthogonal, 1, fieldRAU(), c.pCoefGrad, gradPP.p, work9.p, C.err.p, s); gR = gradPP.p; }
        ops::interpolate(mv, 1, fieldRAU(), c.pCoefInterp, gR, rAUf.p, C.err.p, s);
    }
    ops::normalGradient(mv, fieldP(), gradP.p, c.pDiff, sn.p, C.err.p, s);
    kRhieChow<<<blocksFor(std::max(M.nfi, 1)), kBlock, 0, s>>>(mv, phiH.p, Xf.p, rAUf.p, sn.p);
    ops::divergenceFace(mv, phiH.p, divPhiH.p, s);
    BSIM_LAUNCH_CHECK();
    timing.rhieChow += since(t0);
    // Pressure correction (main.cpp:81-93).
    kFillD<<<nb, kBlock, 0, s>>>(nc, pPrime.p, 0.0);
    const int solves = c.pDiff == kOrthogonal ? 1 : c.nonOrthCorr + 1;
    bool pConverged = true;
    rep.pStatus = kConverged;
    // Coefficient interpolation of rAU (procedural_equation.cpp:149-176).
    {
        const double* gR = nullptr;
        if (c.pCoefInterp == kCorrected) { ops::gradient(mv, M.orthogonal, 1, fieldRAU(), c.pCoefGrad, gradPP.p, work9.p, C.err.p, s); gR = gradPP.p; }
        ops::interpolate(mv, 1, fieldRAU(), c.pCoefInterp, gR, kfP.p, C.err.p, s);
    }
    DBuf<int>& refFlag = C.err;  // reuse: bit layout below uses a dedicated word
    (void)refFlag;
    for (int k = 0; k < solves; ++k) {
        t0 = Clock::now();
        double* pd = pSys->candDiag();
        double* po = pSys->candVal();
        kFillD<<<nb, kBlock, 0, s>>>(nc, pd, 0.0);
        kFillD<<<blocksFor(std::max(M.nnzSliced, 1)), kBlock, 0, s>>>(M.nnzSliced, po, 0.0);
        kFillD<<<nb, kBlock, 0, s>>>(nc, psrc.p, 0.0);
        if (c.pDiff != kOrthogonal) {
            ops::gradient(mv, M.orthogonal, 1, fieldPP(), c.pGrad, gradPP.p, work9.p, C.err.p, s);
            // The assembly and diffusionCorrection (procedural_equation.cpp:77-87) from one face pass.
            ops::assembleDiffusionCorrected(mv, fieldPP(), kfP.p, gradPP.p, c.pDiff, -(-1.0), pd, po, psrc.p, corr.p,
                                            faceWork.p, C.err.p, s);
        } else {
            ops::assembleDiffusion(mv, 1, fieldPP(), kfP.p, nullptr, kOrthogonal, -(-1.0), pd, po, psrc.p, C.err.p, s);
            kFillD<<<blocksFor(nf), kBlock, 0, s>>>(nf, corr.p, 0.0);
        }
        kScalarSource<<<nb, kBlock, 0, s>>>(mv, divPhiH.p, psrc.p);
        int* notConst = reinterpret_cast<int*>(red.p + 8);
        BSIM_HIP(hipMemsetAsync(notConst, 0, sizeof(int), s));
        kReferenceTest<<<nb, kBlock, 0, s>>>(mv, pd, po, notConst);
        kReferenceApply<<<1, 64, 0, s>>>(pd, psrc.p, refCell, 0.0, notConst, C.err.p);
        BSIM_LAUNCH_CHECK();
        int nconst = 0;
        BSIM_HIP(hipMemcpyAsync(&nconst, notConst, sizeof(int), hipMemcpyDeviceToHost, s));
        ++run.reductions;
        C.checkErrors();
        rep.referenced = nconst == 0;
        timing.pressureAssembly += since(t0);
        t0 = Clock::now();
        pSys->update();
        pSys->prepare(c.pre);
        const SolveRes r = pSys->solve(psrc.p, pPrime.p, 0);
        pHist = pSys->hist;
        rep.pIters += r.iterations;
        pConverged = pConverged && r.status == kConverged;
        rep.linP = r.relative;
        timing.pressureSolve += since(t0);
        if (r.status == kNumericalFailure) {  // main.cpp:88
            rep.pStatus = kNumericalFailure;
            rep.failed = 1;
            return rep;
        }
    }
    if (!pConverged && rep.pStatus != kNumericalFailure) rep.pStatus = kMaxIterations;
    t0 = Clock::now();
    // Corrections (main.cpp:95-98).
    kPressureFluxPhi<<<blocksFor(nf), kBlock, 0, s>>>(mv, fieldPP(), kfP.p, corr.p, phiH.p, phi.p, C.err.p);
    ops::gradient(mv, M.orthogonal, 1, fieldPP(), c.pGrad, gradPP.p, work9.p, C.err.p, s);
    kCorrectUP<<<nb, kBlock, 0, s>>>(mv, p.p, pPrime.p, c.alphaP, U.p, rAU.p, gradPP.p);
    BSIM_LAUNCH_CHECK();
    timing.correction += since(t0);
    // Turbulence transport (main.cpp:100-110).
    bool turbulenceConverged = true;
    if (turb) {
        t0 = Clock::now();
        const TransportRep t = turb->solveTransport(fieldU(phi.p), phi.p, c.mGrad);
        rep.kIters = t.itK;
        rep.wIters = t.itOmega;
        rep.boundedOmega = t.boundedOmega;
        rep.dTurb = std::max(t.dOmega, t.dK);
        rep.rTurb = std::max(t.rOmega, t.rK);
        if (!t.healthy || !std::isfinite(rep.dTurb) || !std::isfinite(rep.rTurb)) {  // main.cpp:106
            timing.turbulence += since(t0);
            rep.failed = 1;
            return rep;
        }
        rep.turbLinear = t.statusK == kConverged && t.statusOmega == kConverged;
        turbulenceConverged = rep.turbLinear && rep.dTurb <= turb->tolerance() && rep.rTurb <= turb->tolerance();
        rep.turbConverged = turbulenceConverged;
        timing.turbulence += since(t0);
    }
    t0 = Clock::now();
    // Convergence measures (main.cpp:112-126).
    kConvergence<<<reduceBlocks(nc), kBlock, 0, s>>>(mv, U.p, Uprev.p, pPrime.p, p.p, phi.p, part);
    kReduceTail<kRedK><<<1, kTail, 0, s>>>(reduceBlocks(nc), part, StoreSums<kRedK>{red.p});
    BSIM_LAUNCH_CHECK();
    double h[kRedK];
    BSIM_HIP(hipMemcpyAsync(h, red.p, sizeof(h), hipMemcpyDeviceToHost, s));
    ++run.reductions;
    C.checkErrors();
    rep.dU = std::sqrt(h[0]) / std::max(std::sqrt(h[1]), 1e-30);
    rep.dP = std::sqrt(h[2]) / std::max(std::sqrt(h[3]), 1e-30);
    rep.mass = h[4] / std::max(h[5], 1e-30);
    const bool linearConverged = rep.momStatus == kConverged && pConverged;
    rep.converged = linearConverged && turbulenceConverged && rep.rU <= c.tolMom && rep.dU <= c.tolU &&
                    rep.dP <= c.tolP && rep.mass <= c.tolMass;
    // main.cpp:115-116
    if (!std::isfinite(rep.dU) || !std::isfinite(rep.dP) || !std::isfinite(rep.mass)) rep.failed = 1;
    timing.convergence += since(t0);
    ++timing.iterations;
    return rep;
}

}  // namespace bsim
