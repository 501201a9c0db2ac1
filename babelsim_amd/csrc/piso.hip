// Transient PISO on the GPU (src/physics/piso/main.cpp:19-53, 161-300).
// One physical time step: history save, the ddtPhiCorr flux, the momentum
// predictor with the BDF2/Euler time term, nCorrectors..maxCorrectors
// corrector passes that rebuild H(U)/aP from the current velocity and solve
// for the total pressure, then the step diagnostics.  Kernels gather by row
// (cell) or run per face; reductions are the deterministic two-stage ones.
#include "flow_kernels.hpp"
#include "piso.hpp"

#include <algorithm>
#include <chrono>

namespace bsim {
namespace {

using namespace dev;
using namespace flow;
using Clock = std::chrono::steady_clock;
double since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

// Layout of the small result buffer `red` (doubles).
constexpr int kSlotRU = 0;          // 3 sums of the momentum residual
constexpr int kSlotDiag = 8;        // 6 sums + 1 maximum of the step diagnostics
constexpr int kSlotFlag = 16;       // int flag of the reference test
constexpr int kSlotCoupling = 32;   // 8 per corrector pass (6 used)
constexpr int kMaxCorrectorPasses = 100;
constexpr int kPartOffset = kSlotCoupling + 8 * kMaxCorrectorPasses + 64;
constexpr int kDiagK = 6;

// equ::ddt(momentum, rho, U_old) (procedural_equation.cpp:213-222): w = rho V,
// diag += w a0, source += w a1 U_prev (+ w a2 U_older).
__global__ void __launch_bounds__(kBlock) kTimeTerm(MeshView m, double rho, double a0, double a1, double a2, const double* prev,
                          const double* older, double* diag, double* src) {
    const int i = globalThread();
    if (i >= m.nc) return;
    const double w = rho * m.V[i];
    diag[i] += w * a0;
    for (int k = 0; k < 3; ++k) {
        const long long j = 3 * static_cast<long long>(i) + k;
        src[j] += w * a1 * prev[j];
        if (a2 != 0) src[j] += w * a2 * older[j];
    }
}

// rho * temporalFluxCorrection (piso/main.cpp:19-53): the old-time flux
// mismatch weighted by the time scheme and by the flux-normalised coupling,
// on interior faces only.
__global__ void __launch_bounds__(kBlock) kTemporalFlux(int nf, int nfi, const double* phiP, const double* fluxP, const double* phiO,
                              const double* fluxO, double wp, double wo, double rho, double* tf) {
    const int f = globalThread();
    if (f >= nf) return;
    if (f >= nfi) { tf[f] = rho * 0.0; return; }
    const double e = phiP[f] - fluxP[f];
    double corr = wp * e;
    if (wo != 0.0) corr += wo * (phiO[f] - fluxO[f]);
    const double fl = phiP[f];
    const double coupling = e == 0.0 ? 1.0 : fl == 0.0 ? 0.0 : fmax(0.0, 1.0 - fabs(e) / fabs(fl));
    tf[f] = rho * (0.0 + coupling * corr);
}

// relax without rescaling (main.cpp:186; procedural_equation.cpp:298-305),
// rAU = V/aP, 1/aP (main.cpp:187-189) and the non-pressure right-hand side
// b + V grad(p) (main.cpp:193).
__global__ void __launch_bounds__(kBlock) kPisoRelax(MeshView m, double* diag, double* src, const double* Uiter, double alpha,
                           const double* gradP, double* rAU, double* invAP, double* npr) {
    const int i = globalThread();
    if (i >= m.nc) return;
    const double extra = diag[i] * (1 / alpha - 1);
    const double d = diag[i] + extra;
    diag[i] = d;
    const double v = m.V[i];
    for (int k = 0; k < 3; ++k) {
        const long long j = 3 * static_cast<long long>(i) + k;
        const double s = src[j] + extra * Uiter[j];
        src[j] = s;
        npr[j] = s + v * gradP[j];
    }
    rAU[i] = v / d;
    invAP[i] = 1.0 / d;
}

// A U over the momentum rows (equ::apply, procedural_equation.cpp:281-292).
__device__ inline void applyRow(const MeshView& m, const double* diag, const double* off, const double* U, int i,
                                double (&au)[3]) {
    for (int k = 0; k < 3; ++k) au[k] = diag[i] * U[3 * static_cast<long long>(i) + k];
    const RowRange r = m.row(i);
    for (int k_ = 0; r.valid(k_); ++k_) {
        const int e = r.at(k_);
        const int nb = m.mcol[e];
        if (nb < 0) continue;
        const double a = off[e];
        for (int k = 0; k < 3; ++k) au[k] += a * U[3 * static_cast<long long>(nb) + k];
    }
}

// HbyA = (1/aP) ((b0 - A U) + aP U) (main.cpp:206-207).
__global__ void __launch_bounds__(kBlock) kHbyA(MeshView m, const double* diag, const double* off, const double* U, const double* npr,
                      const double* invAP, double* H) {
    const int i = globalThread();
    if (i >= m.nc) return;
    double au[3];
    applyRow(m, diag, off, U, i, au);
    const double d = diag[i], ia = invAP[i];
    for (int k = 0; k < 3; ++k) {
        const long long j = 3 * static_cast<long long>(i) + k;
        H[j] = ia * ((npr[j] - au[k]) + d * U[j]);
    }
}

// phiH += rAUf * ddtPhiCorr on interior faces (main.cpp:210).
__global__ void __launch_bounds__(kBlock) kAddTemporal(int nfi, const double* rAUf, const double* tf, double* phiH) {
    const int f = globalThread();
    if (f < nfi) phiH[f] += rAUf[f] * tf[f];
}

// U = HbyA - rAU grad(p) (main.cpp:230-231).
__global__ void __launch_bounds__(kBlock) kPisoCorrectU(int nc, const double* H, const double* rAU, const double* g, double* U) {
    const int i = globalThread();
    if (i >= nc) return;
    for (int k = 0; k < 3; ++k) {
        const long long j = 3 * static_cast<long long>(i) + k;
        double u = H[j];
        u += -1.0 * (rAU[i] * g[j]);
        U[j] = u;
    }
}

// Squared L2 norms of b0, A U, V grad(p) and the defect b0 - A U - V grad(p)
// (main.cpp:234-239), and of p - p_before and p (relativeMagnitude,
// fvm_execution.cpp:130-147).
__global__ void __launch_bounds__(kBlock) kCoupling(MeshView m, const double* diag, const double* off, const double* U, const double* npr,
                          const double* gradP, const double* p, const double* pOld, double* part) {
    __shared__ double smem[6 * kBlock / 64];
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int i = globalThread(); i < m.nc; i += gridDim.x * blockDim.x) {
        double au[3];
        applyRow(m, diag, off, U, i, au);
        const double v = m.V[i];
        for (int k = 0; k < 3; ++k) {
            const long long j = 3 * static_cast<long long>(i) + k;
            const double b = npr[j], ps = v * gradP[j];
            const double d = (b - au[k]) - ps;
            a[0] += b * b;
            a[1] += au[k] * au[k];
            a[2] += ps * ps;
            a[3] += d * d;
        }
        const double dp = p[i] - pOld[i];
        a[4] += dp * dp;
        a[5] += p[i] * p[i];
    }
    blockPartial<6>(a, part, smem);
}

// relativeChange(U, U_previous), fluxBalance(phi) and courantNumber(phi, dt)
// (fvm_execution.cpp:91-205): six sums per block and the block maximum of
// the cell Courant number.
__global__ void __launch_bounds__(kBlock) kStepDiagnostics(MeshView m, const double* U, const double* Uiter, const double* phi, double dt,
                                 double* part) {
    __shared__ double smem[kDiagK * kBlock / 64];
    double a[kDiagK] = {0, 0, 0, 0, 0, 0};
    double mx = 0.0;
    for (int c = globalThread(); c < m.nc; c += gridDim.x * blockDim.x) {
        const D3 u = ld3(U, c), d = u - ld3(Uiter, c);
        a[0] += dot(d, d);
        a[1] += dot(u, u);
        double imb = 0.0, sc = 0.0, sum = 0.0;
        const RowRange r = m.row(c);
        for (int k_ = 0; r.valid(k_); ++k_) {
            const int f = m.mface[r.at(k_)];
            if (f < 0) continue;
            const double out = m.own[f] == c ? phi[f] : -phi[f];
            imb += out;
            sc += fabs(out);
            sum += fabs(phi[f]);
        }
        for (int e = m.bptr[c]; e < m.bptr[c + 1]; ++e) {
            const double out = phi[m.bface[e]];
            imb += out;
            sc += fabs(out);
            sum += fabs(out);
        }
        a[2] += fabs(imb);
        a[3] += sc;
        const double v = m.V[c];
        const double co = 0.5 * dt * sum / v;
        mx = fmax(mx, co);
        a[4] += co * v;
        a[5] += v;
    }
    blockPartial<kDiagK>(a, part, smem);
    mx = blockMaxValue(mx, smem);
    if (threadIdx.x == 0) part[kDiagK * kMaxReduceBlocks + xcdBlock()] = mx;
}

// Second stage of kStepDiagnostics: the six sums in block order and the
// maximum, stored to out[0..6].
__global__ void __launch_bounds__(kTail) kDiagnosticsTail(int nb, const double* part, double* out) {
    __shared__ double smem[(kDiagK + 1) * kTail / 64];
    double acc[kDiagK];
    for (int k = 0; k < kDiagK; ++k) acc[k] = 0.0;
    double mx = 0.0;
    for (int b = threadIdx.x; b < nb; b += kTail) {
        for (int k = 0; k < kDiagK; ++k) acc[k] += part[k * kMaxReduceBlocks + b];
        mx = fmax(mx, part[kDiagK * kMaxReduceBlocks + b]);
    }
    blockSum<kDiagK, kTail>(acc, smem);
    mx = blockMaxValue<kTail>(mx, smem);
    if (threadIdx.x == 0) {
        for (int k = 0; k < kDiagK; ++k) out[k] = acc[k];
        out[kDiagK] = mx;
    }
}

// U -= g (Field operator-=, component by component).
__global__ void __launch_bounds__(kBlock) kSubtractVec(int nc, const double* g, double* U) {
    const int i = globalThread();
    if (i >= nc) return;
    for (int k = 0; k < 3; ++k) {
        const long long j = 3 * static_cast<long long>(i) + k;
        U[j] = U[j] - g[j];
    }
}

// Per-block maxima of |U| (math::max(map(U, norm))), then their maximum.
__global__ void __launch_bounds__(kBlock) kSpeedMax(int nc, const double* U, double* part) {
    __shared__ double smem[kBlock / 64];
    double mx = 0.0;
    for (int c = globalThread(); c < nc; c += gridDim.x * blockDim.x) {
        const D3 u = ld3(U, c);
        mx = fmax(mx, sqrt(dot(u, u)));
    }
    mx = blockMaxValue(mx, smem);
    if (threadIdx.x == 0) part[xcdBlock()] = mx;
}
__global__ void __launch_bounds__(kTail) kMaxTail(int nb, const double* part, double* out) {
    __shared__ double smem[kTail / 64];
    double mx = 0.0;
    for (int b = threadIdx.x; b < nb; b += kTail) mx = fmax(mx, part[b]);
    mx = blockMaxValue<kTail>(mx, smem);
    if (threadIdx.x == 0) out[0] = mx;
}

}  // namespace

PisoSolver::PisoSolver(Context& ctx, const PisoCfg& cfg, const int* ubct, const double* ubcv, const int* pbct,
                       const double* pbcv)
    : C(ctx), c(cfg) {
    const DeviceMesh& M = *C.mesh;
    hipStream_t s = C.stream;
    const int nc = M.nc, nf = M.nf, np = M.np;
    // main.cpp:83-105 and procedural_equation.cpp:298-299
    if (!(c.alphaU > 0 && c.alphaU <= 1)) throw std::invalid_argument("relaxation must be in (0,1]");
    if (c.timeMethod != kEuler && c.timeMethod != kBdf2) throw std::invalid_argument("PISO requires a transient time scheme");
    if (c.maxIterations < 1 || c.nCorrectors < 1 || c.nCorrectors > kMaxCorrectorPasses || c.nonOrthCorr < 0 ||
        c.nonOrthCorr > 20)
        throw std::invalid_argument("PISO iteration counts out of range");
    if (c.maxCorrectors == 0) c.maxCorrectors = c.nCorrectors;
    if (c.maxCorrectors < c.nCorrectors || c.maxCorrectors > kMaxCorrectorPasses)
        throw std::invalid_argument("maxCorrectors must lie in [nCorrectors, 100]");
    if (!(c.couplingTol > 0)) throw std::invalid_argument("couplingTolerance must be positive");
    if (!(c.rho > 0) || !(c.mu > 0)) throw std::invalid_argument("density and viscosity must be positive");
    if (c.potentialInit != 0 && c.potentialInit != 1) throw std::invalid_argument("initializePotentialFlow must be 0 or 1");
    if (c.initialNonOrthCorr < 0) c.initialNonOrthCorr = c.nonOrthCorr;
    if (c.initialNonOrthCorr > 20) throw std::invalid_argument("PISO iteration counts out of range");
    std::vector<int> zg(np, kZeroGradient), ppt(np);
    std::vector<double> zgv(3 * np, 0.0), ppv(np, 0.0);
    for (int k = 0; k < np; ++k) {
        if (ubct[k] < 0 || ubct[k] > 4 || pbct[k] < 0 || pbct[k] > 4) throw std::invalid_argument("invalid boundary type");
        inletOutlet = inletOutlet || ubct[k] == kInletOutlet;
        // homogeneousLike(p) (field.h:382-409): the potential's boundary conditions
        ppt[k] = pbct[k] == kFixedValue ? kFixedValue : pbct[k] == kSymmetry ? kSymmetry : kZeroGradient;
    }
    uType.upload(ubct, np, s); uVal.upload(ubcv, 3 * np, s);
    pType.upload(pbct, np, s); pVal.upload(pbcv, np, s);
    zgType.upload(zg, s); zgVal.upload(zgv, s);
    ppType.upload(ppt, s); ppVal.upload(ppv, s);
    for (DBuf<double>* b : {&p, &rAU, &invAP, &imbalance, &pOld, &psrc, &mu}) b->alloc(nc);
    for (int k = 0; k < 3; ++k) { bcomp[k].alloc(nc); xcomp[k].alloc(nc); }
    for (DBuf<double>* b : {&U, &Uiter, &HbyA, &gradP, &npr, &src, &Uh[0], &Uh[1]})
        b->alloc(3 * static_cast<std::size_t>(nc));
    for (DBuf<double>* b : {&gradU, &work9}) b->alloc(9 * static_cast<std::size_t>(nc));
    for (DBuf<double>* b : {&phi, &rAUtr, &rAUf, &kfMom, &corr, &full, &orthF, &tf, &fluxTmp, &phih[0], &phih[1],
                            &fluxh[0], &fluxh[1]})
        b->alloc(nf);
    faceWork.alloc(6 * static_cast<std::size_t>(std::max(M.nfi, 1)));
    red.alloc(kPartOffset + (kDiagK + 1) * static_cast<std::size_t>(kMaxReduceBlocks));
    U.zero(s); p.zero(s); phi.zero(s); corr.zero(s); tf.zero(s);
    const MeshView mv = M.view();
    momSys = std::make_unique<DeviceSystem>(s, meshPattern(M));
    pSys = std::make_unique<DeviceSystem>(s, meshPattern(M));
    refCell = C.cellOldToNew.at(0);
    // muEffective: a constant cell field with zeroGradient patches
    // (src/physics/RANS/model.cpp:58-63), interpolated once with the momentum
    // diffusion coefficient scheme.
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

void PisoSolver::setFields(const double* Uc, const double* pc) {
    hipStream_t s = C.stream;
    if (Uc) { auto v = C.cellsIn(Uc, 3); U.upload(v, s); }
    if (pc) { auto v = C.cellsIn(pc, 1); p.upload(v, s); }
    BSIM_HIP(hipStreamSynchronize(s));
    started = false;
}

void PisoSolver::getFields(double* Uc, double* pc, double* phic) {
    hipStream_t s = C.stream;
    if (Uc) C.cellsOut(U.toHost(s), Uc, 3);
    if (pc) C.cellsOut(p.toHost(s), pc, 1);
    if (phic) C.facesOut(phi.toHost(s), phic, 1);
}

void PisoSolver::temporalFlux(double* out) {
    if (!stepped) throw std::invalid_argument("no time step has been taken");
    C.facesOut(tf.toHost(C.stream), out, 1);
}

void PisoSolver::enableKOmega(const KOmegaCfg& kw, const int* kbct, const double* kbcv, const int* wbct,
                              const double* wbcv, const PisoTurbCfg& t) {
    if (started) throw std::logic_error("the turbulence model is enabled before the run starts");
    // main.cpp:99-120
    if (t.coupling != 0 && t.coupling != 1) throw std::invalid_argument("turbulenceCoupling must be iterated or segregated");
    const bool segregated = t.coupling == 1;
    if (segregated && c.maxIterations != 1)
        throw std::invalid_argument("segregated turbulence requires one PISO momentum pass");
    if (segregated && kw.relaxation != 1.0)
        throw std::invalid_argument("single-pass transient turbulence requires turbulenceRelaxation 1");
    if (t.maxIterations < 1) throw std::invalid_argument("turbulenceMaxIterations must be positive");
    PisoTurbCfg v = t;
    if (v.correctors <= 0) v.correctors = c.maxIterations == 1 && !segregated ? t.maxIterations : 1;
    if (v.correctors > v.maxIterations) throw std::invalid_argument("turbulenceCorrectors out of range");
    if (segregated && v.correctors != 1) throw std::invalid_argument("segregated turbulence requires one transport correction");
    auto model = std::make_unique<KOmegaDevice>(C, kw, c.rho, c.mu, kbct, kbcv, wbct, wbcv);
    model->setTransient(c.timeMethod);
    turb = std::move(model);
    tc = v;
}

void PisoSolver::start() {
    // main.cpp:123: phi = flux(U, pressure options); the first flux seeds
    // inletOutlet directions from owner velocities (operators.cpp:977-987).
    // The histories start from the current fields (main.cpp:157-158).
    const DeviceMesh& M = *C.mesh;
    const MeshView mv = M.view();
    hipStream_t s = C.stream;
    if (turb) turb->start();  // the model exists before phi (main.cpp:71, 98)
    ops::seedFluxContext(mv, U.p, fluxTmp.p, s);
    const double* g = nullptr;
    if (c.pInterp == kCorrected) { ops::gradient(mv, M.orthogonal, 3, fieldU(fluxTmp.p), c.pGrad, gradU.p, work9.p, C.err.p, s); g = gradU.p; }
    ops::flux(mv, fieldU(fluxTmp.p), g, phi.p, C.err.p, s);
    initial = InitRep{};
    if (c.potentialInit) initializePotentialFlow();
    const long long n3 = 3LL * M.nc;
    for (int k = 0; k < 2; ++k) {
        kCopyD<<<blocksFor(n3), kBlock, 0, s>>>(n3, U.p, Uh[k].p);
        kCopyD<<<blocksFor(M.nf), kBlock, 0, s>>>(M.nf, phi.p, phih[k].p);
        kCopyD<<<blocksFor(M.nf), kBlock, 0, s>>>(M.nf, phi.p, fluxh[k].p);
    }
    BSIM_LAUNCH_CHECK();
    C.checkErrors();
    cur = 0;
    levels = 0;
    dtCur = dtPrev = 0;
    started = true;
    stepped = false;
}

// initializePotentialFlow (main.cpp:127-150): a conservative initial face flux
// before the first time history.  The potential equation is the total-pressure
// equation of the correctors with diffusivity 1 on every face and the
// divergence of the initial flux as its imbalance; it has its own persistent
// linear system (another numerical identity, like potentialEquation), which
// lives only for the projection.  Then U -= grad(potential) and phi +=
// faceFlux(potential equation, potential); U's flux context is phi as in
// every step.  A solve that does not converge ends the run as a numerical
// failure (main.cpp:139), which step() then reports.
void PisoSolver::initializePotentialFlow() {
    const DeviceMesh& M = *C.mesh;
    const MeshView mv = M.view();
    hipStream_t s = C.stream;
    const int nc = M.nc, nf = M.nf, nb = blocksFor(nc);
    const int solves = c.pDiff == kOrthogonal ? 1 : c.initialNonOrthCorr + 1;
    pPrime.ensure(nc);
    kFillD<<<nb, kBlock, 0, s>>>(nc, pPrime.p, 0.0);
    kFillD<<<blocksFor(nf), kBlock, 0, s>>>(nf, full.p, 1.0);  // the constant diffusivity, every face
    const double* ones = full.p;
    ops::divergenceFace(mv, phi.p, imbalance.p, s);
    DeviceSystem pot(s, meshPattern(M));
    initial.solves = solves;
    for (int k = 0; k < solves; ++k) {
        double* pd = pot.candDiag();
        double* po = pot.candVal();
        kFillD<<<nb, kBlock, 0, s>>>(nc, pd, 0.0);
        kFillD<<<blocksFor(std::max(M.nnzSliced, 1)), kBlock, 0, s>>>(M.nnzSliced, po, 0.0);
        kFillD<<<nb, kBlock, 0, s>>>(nc, psrc.p, 0.0);
        if (c.pDiff != kOrthogonal) {
            ops::gradient(mv, M.orthogonal, 1, fieldPP(), c.pGrad, gradP.p, work9.p, C.err.p, s);
            ops::assembleDiffusionCorrected(mv, fieldPP(), ones, gradP.p, c.pDiff, -(-1.0), pd, po, psrc.p, corr.p,
                                            faceWork.p, C.err.p, s);
        } else {
            ops::assembleDiffusion(mv, 1, fieldPP(), ones, nullptr, kOrthogonal, -(-1.0), pd, po, psrc.p, C.err.p, s);
            kFillD<<<blocksFor(nf), kBlock, 0, s>>>(nf, corr.p, 0.0);
        }
        kScalarSource<<<nb, kBlock, 0, s>>>(mv, imbalance.p, psrc.p);
        int* notConst = reinterpret_cast<int*>(red.p + kSlotFlag);
        kFillD<<<1, 64, 0, s>>>(1, red.p + kSlotFlag, 0.0);
        kReferenceTest<<<nb, kBlock, 0, s>>>(mv, pd, po, notConst);
        kReferenceApply<<<1, 64, 0, s>>>(pd, psrc.p, refCell, 0.0, notConst, C.err.p);
        BSIM_LAUNCH_CHECK();
        pot.update();
        pot.prepare(c.pre);
        const SolveRes r = pot.solve(psrc.p, pPrime.p, 0);
        C.checkErrors();
        initial.pIters += r.iterations;
        if (r.status != kConverged) {
            initial.status = kStepNumericalFailure;
            initial.done = 1;
            return;
        }
    }
    ops::gradient(mv, M.orthogonal, 1, fieldPP(), c.pGrad, gradP.p, work9.p, C.err.p, s);
    kSubtractVec<<<nb, kBlock, 0, s>>>(nc, gradP.p, U.p);
    kPressureFluxPhi<<<blocksFor(nf), kBlock, 0, s>>>(mv, fieldPP(), ones, corr.p, phi.p, phi.p, C.err.p);
    // initialMass (fluxBalance(phi).relative) and initialUmax.
    double* part = red.p + kPartOffset;
    kStepDiagnostics<<<reduceBlocks(nc), kBlock, 0, s>>>(mv, U.p, U.p, phi.p, 0.0, part);
    kDiagnosticsTail<<<1, kTail, 0, s>>>(reduceBlocks(nc), part, red.p + kSlotDiag);
    kSpeedMax<<<reduceBlocks(nc), kBlock, 0, s>>>(nc, U.p, part);
    kMaxTail<<<1, kTail, 0, s>>>(reduceBlocks(nc), part, red.p + kSlotDiag + kDiagK + 1);
    BSIM_LAUNCH_CHECK();
    double h[kDiagK + 2];
    BSIM_HIP(hipMemcpyAsync(h, red.p + kSlotDiag, sizeof(h), hipMemcpyDeviceToHost, s));
    ++run.reductions;
    C.checkErrors();
    initial.mass = h[2] / std::max(h[3], 1e-30);
    initial.uMax = h[kDiagK + 1];
    initial.done = 1;
}

// temporalFluxCorrection (main.cpp:19-53).  flux(U_prev) also becomes the
// saved velocity's flux context (fvm_execution.cpp:277-279), which the next
// step's flux(U_older) resolves inletOutlet faces with.
void PisoSolver::computeTemporalFlux() {
    const DeviceMesh& M = *C.mesh;
    const MeshView mv = M.view();
    hipStream_t s = C.stream;
    const int old = 1 - cur;
    double wp = 1.0 / dtCur, wo = 0.0;
    if (c.timeMethod == kBdf2 && levels >= 2) {
        const double ratio = dtCur / dtPrev;
        wp = (1.0 + ratio) / dtCur;
        wo = -ratio * ratio / ((1.0 + ratio) * dtCur);
    }
    // The previous velocity carries phi_prev as its flux context (U's context
    // is phi whenever it is saved).
    const FieldArg Up{Uh[cur].p, uType.p, uVal.p, phih[cur].p, nullptr};
    const double* g = nullptr;
    if (c.pInterp == kCorrected) { ops::gradient(mv, M.orthogonal, 3, Up, c.pGrad, gradU.p, work9.p, C.err.p, s); g = gradU.p; }
    ops::flux(mv, Up, g, fluxTmp.p, C.err.p, s);
    const double* fluxO = fluxh[old].p;
    if (wo != 0.0 && inletOutlet) {
        // flux(U_older) with the context it was given last step; without
        // inletOutlet patches it equals the stored flux bit for bit.
        const FieldArg Uo{Uh[old].p, uType.p, uVal.p, fluxh[old].p, nullptr};
        const double* go = nullptr;
        if (c.pInterp == kCorrected) { ops::gradient(mv, M.orthogonal, 3, Uo, c.pGrad, gradU.p, work9.p, C.err.p, s); go = gradU.p; }
        ops::flux(mv, Uo, go, orthF.p, C.err.p, s);
        fluxO = orthF.p;
    }
    kTemporalFlux<<<blocksFor(M.nf), kBlock, 0, s>>>(M.nf, M.nfi, phih[cur].p, fluxTmp.p, phih[old].p, fluxO, wp, wo,
                                                     c.rho, tf.p);
    kCopyD<<<blocksFor(M.nf), kBlock, 0, s>>>(M.nf, fluxTmp.p, fluxh[cur].p);
    BSIM_LAUNCH_CHECK();
}

StepRep PisoSolver::step(double dt) {
    struct Record {  // RunTime::recordIteration at every way out of the step
        RunClock& r;
        ~Record() { r.record(); }
    } record{run};
    if (!(dt > 0) || !std::isfinite(dt)) throw std::invalid_argument("history requires positive finite dt");
    if (!started) start();
    StepRep rep;
    rep.dt = dt;
    if (initial.status != kStepCompleted) return rep;  // runPiso ended in the initial projection
    const DeviceMesh& M = *C.mesh;
    const MeshView mv = M.view();
    hipStream_t s = C.stream;
    const int nc = M.nc, nf = M.nf;
    const int nb = blocksFor(nc);
    const long long n3 = 3LL * nc;
    double* part = red.p + kPartOffset;
    auto t0 = Clock::now();
    // History saves (main.cpp:163-164; history.h:8-13): the previous level
    // becomes the older one and U, phi the previous one.
    cur = 1 - cur;
    kCopyD<<<blocksFor(n3), kBlock, 0, s>>>(n3, U.p, Uh[cur].p);
    kCopyD<<<blocksFor(nf), kBlock, 0, s>>>(nf, phi.p, phih[cur].p);
    dtPrev = dtCur;
    dtCur = dt;
    ++levels;
    if (turb) turb->saveOld(dt);  // main.cpp:165
    computeTemporalFlux();
    stepped = true;
    // ddt weights (procedural_equation.cpp:204-212; Euler below two levels, 231-237)
    double a0 = 1 / dtCur, a1 = 1 / dtCur, a2 = 0;
    if (c.timeMethod == kBdf2 && levels >= 2) {
        const double olddt = dtPrev == 0 ? dtCur : dtPrev;
        const double r = dtCur / olddt;
        a0 = (1 + 2 * r) / ((1 + r) * dtCur);
        a1 = (1 + r) / dtCur;
        a2 = -r * r / ((1 + r) * dtCur);
    }
    const int old = 1 - cur;
    const int pressureSolves = c.pDiff == kOrthogonal ? 1 : c.nonOrthCorr + 1;
    // Without a coupling tolerance the corrector count is fixed, so the
    // coupling sums are read once after the last pass.
    const bool adaptive = c.maxCorrectors > c.nCorrectors || c.couplingTol < std::numeric_limits<double>::max();
    timing.history += since(t0);
    bool converged = false;
    for (int iter = 0; iter < c.maxIterations; ++iter) {
        t0 = Clock::now();
        rep.passes = iter + 1;
        kCopyD<<<blocksFor(n3), kBlock, 0, s>>>(n3, U.p, Uiter.p);
        // Momentum predictor (main.cpp:174-181).
        double* diag = momSys->candDiag();
        double* off = momSys->candVal();
        kFillD<<<nb, kBlock, 0, s>>>(nc, diag, 0.0);
        kFillD<<<blocksFor(std::max(M.nnzSliced, 1)), kBlock, 0, s>>>(M.nnzSliced, off, 0.0);
        kFillD<<<blocksFor(n3), kBlock, 0, s>>>(n3, src.p, 0.0);
        kTimeTerm<<<nb, kBlock, 0, s>>>(mv, c.rho, a0, a1, a2, Uh[cur].p, Uh[old].p, diag, src.p);
        const FieldArg Uf = fieldU(phi.p);
        const bool needGradU = c.mConv == kLinearUpwind || (c.mConv == kCentral && c.mInterp == kCorrected) ||
                               c.mDiff != kOrthogonal || turb;
        if (needGradU) ops::gradient(mv, M.orthogonal, 3, Uf, c.mGrad, gradU.p, work9.p, C.err.p, s);
        if (turb) {
            // laplacian(muEff) with the current eddy viscosity (main.cpp:176)
            const FieldArg mf{turb->muEff(), zgType.p, zgVal.p, nullptr, turb->muEffTrace()};
            const double* gm = nullptr;
            if (c.mCoefInterp == kCorrected) { ops::gradient(mv, M.orthogonal, 1, mf, c.mCoefGrad, gradP.p, work9.p, C.err.p, s); gm = gradP.p; }
            ops::interpolate(mv, 1, mf, c.mCoefInterp, gm, kfMom.p, C.err.p, s);
        }
        ops::assembleTransport(mv, 3, Uf, phi.p, c.mConv, c.mInterp, c.rho, kfMom.p, c.mDiff, -(-1.0), gradU.p, diag,
                               off, src.p, faceWork.p, C.err.p, s);
        ops::gradient(mv, M.orthogonal, 1, fieldP(), c.pGrad, gradP.p, work9.p, C.err.p, s);
        kMomentumPressureSource<<<nb, kBlock, 0, s>>>(mv, gradP.p, src.p);
        if (turb) turb->addStressSource(Uf, gradU.p, c.mInterp, c.mGrad, src.p);  // main.cpp:179-182
        // rU before relaxation (main.cpp:184)
        kMomentumResidual<<<reduceBlocks(nc), kBlock, 0, s>>>(mv, diag, off, src.p, U.p, part);
        kReduceTail<3><<<1, kTail, 0, s>>>(reduceBlocks(nc), part, StoreSums<3>{red.p + kSlotRU});
        kPisoRelax<<<nb, kBlock, 0, s>>>(mv, diag, src.p, Uiter.p, c.alphaU, gradP.p, rAU.p, invAP.p, npr.p);
        if (nf > M.nfi) kOwnerTrace<<<blocksFor(nf - M.nfi), kBlock, 0, s>>>(mv, rAU.p, rAUtr.p);
        {
            const double* gR = nullptr;
            if (c.pCoefInterp == kCorrected) { ops::gradient(mv, M.orthogonal, 1, fieldRAU(), c.pCoefGrad, work9.p + 3LL * nc, work9.p, C.err.p, s); gR = work9.p + 3LL * nc; }
            ops::interpolate(mv, 1, fieldRAU(), c.pCoefInterp, gR, rAUf.p, C.err.p, s);
        }
        BSIM_LAUNCH_CHECK();
        double hr[3];
        BSIM_HIP(hipMemcpyAsync(hr, red.p + kSlotRU, sizeof(hr), hipMemcpyDeviceToHost, s));
        ++run.reductions;
        C.checkErrors();
        rep.rU = std::sqrt(hr[0]) / std::max(std::sqrt(hr[1]) + std::sqrt(hr[2]), 1e-30);
        timing.assembly += since(t0);
        if (!std::isfinite(rep.rU)) { rep.status = kStepNumericalFailure; return rep; }
        t0 = Clock::now();
        // Momentum solve, three right-hand sides (procedural_equation.cpp:357-370).
        momSys->update();
        momSys->prepare(c.mom);
        kSplit<<<nb, kBlock, 0, s>>>(nc, src.p, bcomp[0].p, bcomp[1].p, bcomp[2].p);
        kSplit<<<nb, kBlock, 0, s>>>(nc, U.p, xcomp[0].p, xcomp[1].p, xcomp[2].p);
        int momStatus = kConverged;
        double linU = 0.0;
        {
            const double* const b3[3] = {bcomp[0].p, bcomp[1].p, bcomp[2].p};
            double* const x3[3] = {xcomp[0].p, xcomp[1].p, xcomp[2].p};
            SolveRes r3[3];
            momSys->solve3(b3, x3, r3, momHist);
            for (int comp = 0; comp < 3; ++comp) {
                const SolveRes& r = r3[comp];
                rep.momIters[comp] = r.iterations;
                if (r.status == kNumericalFailure) momStatus = kNumericalFailure;
                else if (r.status != kConverged && momStatus != kNumericalFailure) momStatus = kMaxIterations;
                linU = std::max(linU, r.relative);
            }
        }
        kMerge<<<nb, kBlock, 0, s>>>(nc, xcomp[0].p, xcomp[1].p, xcomp[2].p, U.p);
        BSIM_LAUNCH_CHECK();
        rep.linU = linU;
        timing.momentumSolve += since(t0);
        if (momStatus == kNumericalFailure) { rep.status = kStepNumericalFailure; return rep; }
        // Correctors (main.cpp:202-249).
        const double* Ad = momSys->diag();
        const double* Ao = momSys->val();
        bool pressureConverged = true;
        double pLin = 0.0, dP = 0.0, first = 0.0, coupling = 0.0;
        int used = 0;
        rep.pIters = 0;
        auto couplingOf = [&](const double* h) {
            const double n0 = std::sqrt(h[0]), n1 = std::sqrt(h[1]), n2 = std::sqrt(h[2]), n3 = std::sqrt(h[3]);
            const double scale = n0 + n1 + n2;
            return scale > 0.0 ? n3 / scale : n3;
        };
        auto dPOf = [](const double* h) { return std::sqrt(h[4]) / std::max(std::sqrt(h[5]), 1e-30); };
        for (int corrector = 0; corrector < c.maxCorrectors; ++corrector) {
            t0 = Clock::now();
            kHbyA<<<nb, kBlock, 0, s>>>(mv, Ad, Ao, U.p, npr.p, invAP.p, HbyA.p);
            kCopyD<<<nb, kBlock, 0, s>>>(nc, p.p, pOld.p);
            // HbyA keeps U's boundary conditions and takes phi as its flux context.
            const FieldArg Hf{HbyA.p, uType.p, uVal.p, phi.p, nullptr};
            const double* gH = nullptr;
            if (c.pInterp == kCorrected) { ops::gradient(mv, M.orthogonal, 3, Hf, c.pGrad, gradU.p, work9.p, C.err.p, s); gH = gradU.p; }
            ops::flux(mv, Hf, gH, full.p, C.err.p, s);  // phiH
            double* phiH = full.p;
            kAddTemporal<<<blocksFor(std::max(M.nfi, 1)), kBlock, 0, s>>>(M.nfi, rAUf.p, tf.p, phiH);
            ops::divergenceFace(mv, phiH, imbalance.p, s);
            BSIM_LAUNCH_CHECK();
            timing.corrector += since(t0);
            // Total-pressure solves (main.cpp:216-225).
            for (int k = 0; k < pressureSolves; ++k) {
                t0 = Clock::now();
                double* pd = pSys->candDiag();
                double* po = pSys->candVal();
                kFillD<<<nb, kBlock, 0, s>>>(nc, pd, 0.0);
                kFillD<<<blocksFor(std::max(M.nnzSliced, 1)), kBlock, 0, s>>>(M.nnzSliced, po, 0.0);
                kFillD<<<nb, kBlock, 0, s>>>(nc, psrc.p, 0.0);
                if (c.pDiff != kOrthogonal) {
                    ops::gradient(mv, M.orthogonal, 1, fieldP(), c.pGrad, gradP.p, work9.p, C.err.p, s);
                    // The assembly and diffusionCorrection (procedural_equation.cpp:77-87) from one face pass.
                    ops::assembleDiffusionCorrected(mv, fieldP(), rAUf.p, gradP.p, c.pDiff, -(-1.0), pd, po, psrc.p,
                                                    corr.p, faceWork.p, C.err.p, s);
                } else {
                    ops::assembleDiffusion(mv, 1, fieldP(), rAUf.p, nullptr, kOrthogonal, -(-1.0), pd, po, psrc.p, C.err.p, s);
                    kFillD<<<blocksFor(nf), kBlock, 0, s>>>(nf, corr.p, 0.0);
                }
                kScalarSource<<<nb, kBlock, 0, s>>>(mv, imbalance.p, psrc.p);
                int* notConst = reinterpret_cast<int*>(red.p + kSlotFlag);
                kFillD<<<1, 64, 0, s>>>(1, red.p + kSlotFlag, 0.0);
                kReferenceTest<<<nb, kBlock, 0, s>>>(mv, pd, po, notConst);
                kReferenceApply<<<1, 64, 0, s>>>(pd, psrc.p, refCell, 0.0, notConst, C.err.p);
                BSIM_LAUNCH_CHECK();
                timing.pressureAssembly += since(t0);
                t0 = Clock::now();
                pSys->update();
                pSys->prepare(c.pre);
                const SolveRes r = pSys->solve(psrc.p, p.p, 0);
                pHist = pSys->hist;
                C.checkErrors();
                rep.pIters += r.iterations;
                timing.pressureSolve += since(t0);
                if (r.status == kNumericalFailure) { rep.status = kStepNumericalFailure; return rep; }
                pressureConverged = pressureConverged && r.status == kConverged;
                pLin = std::max(pLin, r.relative);
            }
            t0 = Clock::now();
            // phi = phiH + faceFlux(pEq, p) (main.cpp:228), U = HbyA - rAU grad(p).
            kPressureFluxPhi<<<blocksFor(nf), kBlock, 0, s>>>(mv, fieldP(), rAUf.p, corr.p, phiH, phi.p, C.err.p);
            ops::gradient(mv, M.orthogonal, 1, fieldP(), c.pGrad, gradP.p, work9.p, C.err.p, s);
            kPisoCorrectU<<<nb, kBlock, 0, s>>>(nc, HbyA.p, rAU.p, gradP.p, U.p);
            double* slot = red.p + kSlotCoupling + 8 * corrector;
            kCoupling<<<reduceBlocks(nc), kBlock, 0, s>>>(mv, Ad, Ao, U.p, npr.p, gradP.p, p.p, pOld.p, part);
            kReduceTail<6><<<1, kTail, 0, s>>>(reduceBlocks(nc), part, StoreSums<6>{slot});
            BSIM_LAUNCH_CHECK();
            used = corrector + 1;
            bool done = used >= c.nCorrectors && !adaptive;
            if (adaptive) {
                double h[6];
                BSIM_HIP(hipMemcpyAsync(h, slot, sizeof(h), hipMemcpyDeviceToHost, s));
                ++run.reductions;
                BSIM_HIP(hipStreamSynchronize(s));
                done = used >= c.nCorrectors && couplingOf(h) <= c.couplingTol;
            }
            timing.corrector += since(t0);
            if (done) break;
        }
        // Turbulence transport at this time level (main.cpp:251-266).
        bool turbulenceConverged = true, turbulenceLinear = true;
        if (turb) {
            rep.kIters = rep.wIters = 0;
            for (int correction = 0; correction < tc.correctors; ++correction) {
                t0 = Clock::now();
                const TransportRep t = turb->solveTransport(fieldU(phi.p), phi.p, c.mGrad);
                timing.turbulence += since(t0);
                rep.turbCorrections = correction + 1;
                rep.kIters += t.itK;
                rep.wIters += t.itOmega;
                rep.dTurb = std::max(t.dOmega, t.dK);
                rep.rTurb = std::max(t.rOmega, t.rK);
                if (!t.healthy || !std::isfinite(rep.dTurb) || !std::isfinite(rep.rTurb)) {
                    rep.status = kStepNumericalFailure;
                    return rep;
                }
                turbulenceLinear = t.statusK == kConverged && t.statusOmega == kConverged;
                turbulenceConverged = turbulenceLinear && rep.dTurb <= turb->tolerance() && rep.rTurb <= turb->tolerance();
                if (!turbulenceLinear || turbulenceConverged) break;
            }
            rep.turbLinear = turbulenceLinear;
            rep.turbConverged = turbulenceConverged;
        }
        t0 = Clock::now();
        // Step diagnostics (main.cpp:268-270; fvm_execution.cpp:91-205).
        kStepDiagnostics<<<reduceBlocks(nc), kBlock, 0, s>>>(mv, U.p, Uiter.p, phi.p, dt, part);
        kDiagnosticsTail<<<1, kTail, 0, s>>>(reduceBlocks(nc), part, red.p + kSlotDiag);
        BSIM_LAUNCH_CHECK();
        std::vector<double> hs(static_cast<std::size_t>(8) * used + 8);
        BSIM_HIP(hipMemcpyAsync(hs.data(), red.p + kSlotDiag, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
        BSIM_HIP(hipMemcpyAsync(hs.data() + 8, red.p + kSlotCoupling, sizeof(double) * 8 * used,
                                hipMemcpyDeviceToHost, s));
        ++run.reductions;
        C.checkErrors();
        for (int k = 0; k < used; ++k) {
            const double* h = hs.data() + 8 + 8 * k;
            dP = std::max(dP, dPOf(h));
            coupling = couplingOf(h);
            if (k == 0) first = coupling;
        }
        const double* d = hs.data();
        rep.dU = std::sqrt(d[0]) / std::max(std::sqrt(d[1]), 1e-30);
        rep.mass = d[2] / std::max(d[3], 1e-30);
        rep.coMean = d[4] / d[5];
        rep.coMax = d[6];
        rep.correctors = used;
        rep.dP = dP;
        rep.rCouplingFirst = first;
        rep.rCoupling = coupling;
        rep.linP = pLin;
        timing.diagnostics += since(t0);
        if (!(std::isfinite(rep.rU) && std::isfinite(rep.dU) && std::isfinite(rep.dP) && std::isfinite(coupling) &&
              std::isfinite(rep.mass))) {
            rep.status = kStepNumericalFailure;
            return rep;
        }
        const bool linearConverged = momStatus == kConverged && pressureConverged && turbulenceLinear;
        rep.linearOk = linearConverged;
        const bool accepted = linearConverged && rep.mass <= c.tolMass && coupling <= c.couplingTol;
        const bool settled = rep.rU <= c.tolMom && rep.dU <= c.tolU && rep.dP <= c.tolP && turbulenceConverged;
        converged = accepted && (c.maxIterations == 1 ? (tc.coupling == 1 || turbulenceConverged) : settled);
        rep.accepted = accepted;
        rep.converged = converged;
        if (!accepted) { rep.status = kStepNotConverged; return rep; }
        if (converged) break;
    }
    ++timing.steps;
    rep.status = converged ? kStepCompleted : kStepNotConverged;
    return rep;
}

}  // namespace bsim
