// Wilcox (1988) k-omega on the device.  Each kernel names the reference line
// it follows; derived fields carry calculated boundary traces computed by the
// same pointwise formula as their cell values (include/babelsim/math.h:185-258).
#include "flow_kernels.hpp"
#include "komega.hpp"

#include <algorithm>

namespace bsim {
namespace {

using namespace dev;
using namespace flow;

// 2 dev(symm(G)):dev(symm(G)), G row-major with rows the velocity components
// (k_omega.cpp:12-23).
__device__ inline double strainSquared(const double* g) {
    const double divergence = g[0] + g[4] + g[8];
    double sum = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double strain = 0.5 * (g[3 * i + j] + g[3 * j + i]) - (i == j ? divergence / 3.0 : 0.0);
            sum += strain * strain;
        }
    return 2.0 * sum;
}

// mut = rho k / max(omega, omegaMin), muEff = mu + mut, cells and boundary
// traces (k_omega.cpp:146-149).
__global__ void __launch_bounds__(kBlock) kUpdateViscosity(MeshView m, FieldArg k, FieldArg w, double rho, double mu, double omegaMin,
                                 double* mut, double* muEff, double* mutT, double* muEffT, int* err) {
    const int i = globalThread();
    if (i < m.nc) {
        const double v = (rho * k.v[i]) / fmax(w.v[i], omegaMin);
        mut[i] = v;
        muEff[i] = v + mu;
    }
    const int f = m.nfi + i;
    if (f < m.nf) {
        double kb, wb;
        traceValue<1>(k, m, f, false, 0.0, &kb, err);
        traceValue<1>(w, m, f, false, 0.0, &wb, err);
        const double v = (rho * kb) / fmax(wb, omegaMin);
        mutT[f] = v;
        muEffT[f] = v + mu;
    }
}

// Production mut 2S:S, diffusivities mu + sigma mut (with traces), the omega
// sink rate beta rho omega and source gamma P omega / max(k, kMin)
// (k_omega.cpp:70-77).
__global__ void __launch_bounds__(kBlock) kTransportCoefficients(MeshView m, const double* gradU, const double* mut, const double* mutT,
                                       const double* k, const double* w, double rho, double mu, double sigmaK,
                                       double sigmaW, double beta, double gamma, double kMin, double* prod,
                                       double* kd, double* wd, double* rate, double* srcW, double* kdT, double* wdT) {
    const int i = globalThread();
    if (i < m.nc) {
        const double P = mut[i] * strainSquared(gradU + 9 * static_cast<long long>(i));
        prod[i] = P;
        kd[i] = sigmaK * mut[i] + mu;
        wd[i] = sigmaW * mut[i] + mu;
        rate[i] = (beta * rho) * w[i];
        srcW[i] = ((gamma * P) * w[i]) / fmax(k[i], kMin);
    }
    const int f = m.nfi + i;
    if (f < m.nf) {
        kdT[f] = sigmaK * mutT[f] + mu;
        wdT[f] = sigmaW * mutT[f] + mu;
    }
}

// The k sink rate beta* rho omega with the new omega (k_omega.cpp:99).
__global__ void __launch_bounds__(kBlock) kRateK(int n, const double* w, double betaStarRho, double* rate) {
    const int i = globalThread();
    if (i < n) rate[i] = betaStarRho * w[i];
}

// reaction (procedural_equation.cpp:181-187) and source (192-196).
// equ::ddt for a scalar history (procedural_equation.cpp:213-222): w = rho V,
// diag += w a0, src += w a1 previous (+ w a2 older).
__global__ void __launch_bounds__(kBlock) kScalarTimeTerm(int n, const double* V, double rho, double a0, double a1,
                                                          double a2, const double* prev, const double* older,
                                                          double* diag, double* src) {
    const int i = globalThread();
    if (i >= n) return;
    const double w = rho * V[i];
    diag[i] += w * a0;
    src[i] += w * a1 * prev[i];
    if (a2 != 0) src[i] += w * a2 * older[i];
}
__global__ void __launch_bounds__(kBlock) kReactionSource(int n, const double* V, const double* rate, const double* s, double* diag,
                                double* b) {
    const int i = globalThread();
    if (i >= n) return;
    diag[i] += 1.0 * rate[i] * V[i];
    b[i] += 1.0 * V[i] * s[i];
}

// relativeResidual sums |b - Ax|^2, |Ax|^2, |b|^2 of a scalar equation
// (procedural_equation.cpp:429-433).
__global__ void __launch_bounds__(kBlock) kScalarResidual(MeshView m, const double* diag, const double* off, const double* b, const double* x,
                                double* part) {
    __shared__ double smem[3 * kBlock / 64];
    double a[3] = {0, 0, 0};
    for (int i = globalThread(); i < m.nc; i += gridDim.x * blockDim.x) {
        double ax = diag[i] * x[i];
        const RowRange r = m.row(i);
        for (int k_ = 0; r.valid(k_); ++k_) {
            const int e = r.at(k_);
            if (m.mcol[e] < 0) continue;
            ax += off[e] * x[m.mcol[e]];
        }
        const double res = b[i] - ax;
        a[0] += res * res;
        a[1] += ax * ax;
        a[2] += b[i] * b[i];
    }
    blockPartial<3>(a, part, smem);
}

// relax without rescaling (procedural_equation.cpp:298-305).
__global__ void __launch_bounds__(kBlock) kRelaxScalar(int n, double* diag, double* b, const double* prev, double alpha) {
    const int i = globalThread();
    if (i >= n) return;
    const double extra = diag[i] * (1 / alpha - 1);
    diag[i] += extra;
    b[i] += extra * prev[i];
}

// omega = max(omega, omegaMin, rho k / (ratioMax mu)) and the count of cells
// it raised (k_omega.cpp:131-144).
__global__ void __launch_bounds__(kBlock) kBoundOmega(int n, double* w, const double* k, double omegaMin, double rho, double ratioMu,
                            double* part) {
    __shared__ double smem[kBlock / 64];
    double cnt = 0.0;
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x) {
        const double wi = w[i];
        const double floor = rho * k[i] / ratioMu;
        cnt += wi < fmax(omegaMin, floor) ? 1.0 : 0.0;
        w[i] = fmax(fmax(wi, omegaMin), floor);
    }
    double v[1] = {cnt};
    blockPartial<1>(v, part, smem);
}

__global__ void __launch_bounds__(kBlock) kBoundK(int n, double* k, double kMin) {
    const int i = globalThread();
    if (i < n) k[i] = fmax(k[i], kMin);
}

// relativeChange sums for k and omega (fvm_execution.cpp:110-128).
__global__ void __launch_bounds__(kBlock) kChange(int n, const double* k, const double* kp, const double* w, const double* wp, double* part) {
    __shared__ double smem[4 * kBlock / 64];
    double a[4] = {0, 0, 0, 0};
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x) {
        const double dk = k[i] - kp[i], dw = w[i] - wp[i];
        a[0] += dk * dk;
        a[1] += k[i] * k[i];
        a[2] += dw * dw;
        a[3] += w[i] * w[i];
    }
    blockPartial<4>(a, part, smem);
}

// Vector-gradient boundary traces (fvm_execution.cpp:243-265), with the
// symmetry reflection of include/babelsim/vector.h:107-115.
__global__ void __launch_bounds__(kBlock) kVectorGradientTraces(MeshView m, FieldArg U, const double* g, double* gt, int* err) {
    const int f = m.nfi + globalThread();
    if (f >= m.nf) return;
    const int o = m.own[f];
    const D3 n = ld3(m.nrm, f);
    const D3 d = ld3(m.Cf, f) - ld3(m.C, o);
    double G[9];
    for (int q = 0; q < 9; ++q) G[q] = g[9 * static_cast<long long>(o) + q];
    const Cnd<3> cnd = condition<3>(U, m, f, false, 0.0, err);
    double out[9];
    if (cnd.type == kSymmetry) {
        const double nn[3] = {n.x, n.y, n.z};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double r = 0.0;
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l)
                        r += ((i == k ? 1.0 : 0.0) - 2 * nn[i] * nn[k]) * G[3 * k + l] *
                             ((l == j ? 1.0 : 0.0) - 2 * nn[l] * nn[j]);
                out[3 * i + j] = 0.5 * (G[3 * i + j] + r);
            }
    } else {
        const double dn = dot(d, n);
        const D3 t = d - dn * n;
        for (int c = 0; c < 3; ++c) {
            D3 gc = d3(G[3 * c], G[3 * c + 1], G[3 * c + 2]);
            const double sn = cnd.type == kFixedValue
                                  ? (cnd.v[c] - U.v[3 * static_cast<long long>(o) + c] - dot(gc, t)) / dn
                              : cnd.type == kFixedGradient ? cnd.v[c] : 0.0;
            gc = gc + (sn - dot(gc, n)) * n;
            out[3 * c] = gc.x;
            out[3 * c + 1] = gc.y;
            out[3 * c + 2] = gc.z;
        }
    }
    for (int q = 0; q < 9; ++q) gt[9 * static_cast<long long>(f) + q] = out[q];
}

// Rows of muEff (grad(U)^T - 2/3 tr(grad U) I), cells and boundary traces
// (model.cpp:67-71).
__device__ inline void stressRows(double mu, const double* G, double* r0, double* r1, double* r2) {
    const double tr = G[0] + G[4] + G[8];
    double* rows[3] = {r0, r1, r2};
    for (int c = 0; c < 3; ++c)
        for (int j = 0; j < 3; ++j) rows[c][j] = mu * (G[3 * j + c] - (2.0 / 3.0) * (c == j ? tr : 0.0));
}
__global__ void __launch_bounds__(kBlock) kStressRows(MeshView m, const double* muEff, const double* muEffT, const double* g,
                            const double* gt, double* r0, double* r1, double* r2, double* t0, double* t1,
                            double* t2) {
    const long long i = globalThread();
    if (i < m.nc) stressRows(muEff[i], g + 9 * i, r0 + 3 * i, r1 + 3 * i, r2 + 3 * i);
    const long long f = m.nfi + i;
    if (f < m.nf) stressRows(muEffT[f], gt + 9 * f, t0 + 3 * f, t1 + 3 * f, t2 + 3 * f);
}

// Component `comp` of div(T): the signed row fluxes of the cell over its
// volume, one face at a time (operators.cpp:1580-1586).
__global__ void __launch_bounds__(kBlock) kDivRow(MeshView m, const double* q, int comp, double* divT) {
    const int c = globalThread();
    if (c >= m.nc) return;
    const double v = m.V[c];
    double acc = 0.0;
    const RowRange r = m.row(c);
    for (int k_ = 0; r.valid(k_); ++k_) {
        const int f = m.mface[r.at(k_)];
        if (f < 0) continue;
        acc = m.own[f] == c ? acc + q[f] / v : acc - q[f] / v;
    }
    for (int e = m.bptr[c]; e < m.bptr[c + 1]; ++e) acc += q[m.bface[e]] / v;
    divT[3 * static_cast<long long>(c) + comp] = acc;
}

// equ::source(momentum, div T) (procedural_equation.cpp:192-196).
__global__ void __launch_bounds__(kBlock) kAddVectorSource(int n, const double* V, const double* s, double* src) {
    const int i = globalThread();
    if (i >= n) return;
    const double v = 1.0 * V[i];
    for (int k = 0; k < 3; ++k) src[3 * static_cast<long long>(i) + k] += v * s[3 * static_cast<long long>(i) + k];
}

}  // namespace

KOmegaDevice::KOmegaDevice(Context& ctx, const KOmegaCfg& cfg, double density, double viscosity, const int* kbct,
                           const double* kbcv, const int* wbct, const double* wbcv)
    : C(ctx), c(cfg), rho(density), mu(viscosity) {
    if (!(c.relaxation > 0 && c.relaxation <= 1)) throw std::invalid_argument("relaxation must be in (0,1]");
    if (!(c.kMin > 0) || !(c.omegaMin > 0) || !(c.viscosityRatioMax > 0))
        throw std::invalid_argument("k-omega bounds must be positive");
    const DeviceMesh& M = *C.mesh;
    hipStream_t s = C.stream;
    const int nc = M.nc, nf = M.nf, np = M.np;
    for (int k_ = 0; k_ < np; ++k_)
        if (kbct[k_] < 0 || kbct[k_] > 4 || wbct[k_] < 0 || wbct[k_] > 4)
            throw std::invalid_argument("invalid boundary type");
    std::vector<int> zg(np, kZeroGradient);
    std::vector<double> zgv(np, 0.0);
    kType.upload(kbct, np, s); kVal.upload(kbcv, np, s);
    wType.upload(wbct, np, s); wVal.upload(wbcv, np, s);
    zgType.upload(zg, s); zgVal.upload(zgv, s);
    for (DBuf<double>* b : {&k, &w, &kPrev, &wPrev, &mut, &muEffC, &prod, &rate, &srcW, &kd, &wd, &bsrc}) b->alloc(nc);
    for (DBuf<double>* b : {&gradX, &divT}) b->alloc(3 * static_cast<std::size_t>(nc));
    for (DBuf<double>* b : {&work, &gradR}) b->alloc(9 * static_cast<std::size_t>(nc));
    for (DBuf<double>* b : {&mutT, &muEffT, &kdT, &wdT, &ctxK, &ctxW, &kf, &qf}) b->alloc(nf);
    for (int q = 0; q < 3; ++q) {
        rows[q].alloc(3 * static_cast<std::size_t>(nc));
        rowsT[q].alloc(3 * static_cast<std::size_t>(nf));
    }
    gtr.alloc(9 * static_cast<std::size_t>(nf));
    red.alloc(16 + 4 * static_cast<std::size_t>(kMaxReduceBlocks));
    for (DBuf<double>* b : {&k, &w, &mut, &muEffC, &mutT, &muEffT, &kdT, &wdT, &ctxK, &ctxW}) b->zero(s);
    kSys = std::make_unique<DeviceSystem>(s, meshPattern(M));
    wSys = std::make_unique<DeviceSystem>(s, meshPattern(M));
}

void KOmegaDevice::setFields(const double* kc, const double* wc) {
    hipStream_t s = C.stream;
    if (kc) { auto v = C.cellsIn(kc, 1); k.upload(v, s); }
    if (wc) { auto v = C.cellsIn(wc, 1); w.upload(v, s); }
    BSIM_HIP(hipStreamSynchronize(s));
}

void KOmegaDevice::getFields(double* kc, double* wc, double* mc) {
    hipStream_t s = C.stream;
    if (kc) C.cellsOut(k.toHost(s), kc, 1);
    if (wc) C.cellsOut(w.toHost(s), wc, 1);
    if (mc) C.cellsOut(mut.toHost(s), mc, 1);
}

void KOmegaDevice::updateViscosity() {
    const DeviceMesh& M = *C.mesh;
    const int n = std::max(M.nc, M.nf - M.nfi);
    kUpdateViscosity<<<blocksFor(n), kBlock, 0, C.stream>>>(M.view(), fieldK(), fieldW(), rho, mu, c.omegaMin, mut.p,
                                                            muEffC.p, mutT.p, muEffT.p, C.err.p);
    BSIM_LAUNCH_CHECK();
}

double KOmegaDevice::boundOmega() {
    const DeviceMesh& M = *C.mesh;
    hipStream_t s = C.stream;
    double* part = red.p + 16;
    kBoundOmega<<<reduceBlocks(M.nc), kBlock, 0, s>>>(M.nc, w.p, k.p, c.omegaMin, rho, c.viscosityRatioMax * mu, part);
    kReduceTail<1><<<1, kTail, 0, s>>>(reduceBlocks(M.nc), part, StoreSums<1>{red.p});
    BSIM_LAUNCH_CHECK();
    double v = 0.0;
    BSIM_HIP(hipMemcpyAsync(&v, red.p, sizeof(double), hipMemcpyDeviceToHost, s));
    ++reads;
    BSIM_HIP(hipStreamSynchronize(s));
    return v;
}

void KOmegaDevice::start() {
    // k_omega.cpp:50-53 and 124-129: contexts are the (zero) face flux of
    // the moment, k >= kMin, omega bounded, then the viscosity.
    const DeviceMesh& M = *C.mesh;
    hipStream_t s = C.stream;
    kFillD<<<blocksFor(M.nf), kBlock, 0, s>>>(M.nf, ctxK.p, 0.0);
    kFillD<<<blocksFor(M.nf), kBlock, 0, s>>>(M.nf, ctxW.p, 0.0);
    kBoundK<<<blocksFor(M.nc), kBlock, 0, s>>>(M.nc, k.p, c.kMin);
    BSIM_LAUNCH_CHECK();
    (void)boundOmega();
    updateViscosity();
    C.checkErrors();
}

void KOmegaDevice::addStressSource(FieldArg U, const double* gradU, int mInterp, int mGrad, double* src) {
    const DeviceMesh& M = *C.mesh;
    const MeshView mv = M.view();
    hipStream_t s = C.stream;
    const int nc = M.nc, nf = M.nf;
    if (nf > M.nfi)
        kVectorGradientTraces<<<blocksFor(nf - M.nfi), kBlock, 0, s>>>(mv, U, gradU, gtr.p, C.err.p);
    kStressRows<<<blocksFor(std::max(nc, nf - M.nfi)), kBlock, 0, s>>>(mv, muEffC.p, muEffT.p, gradU, gtr.p,
                                                                       rows[0].p, rows[1].p, rows[2].p, rowsT[0].p,
                                                                       rowsT[1].p, rowsT[2].p);
    for (int q = 0; q < 3; ++q) {
        const FieldArg R{rows[q].p, zgType.p, zgVal.p, nullptr, rowsT[q].p};
        const double* gR = nullptr;
        if (mInterp == kCorrected) { ops::gradient(mv, M.orthogonal, 3, R, mGrad, gradR.p, work.p, C.err.p, s); gR = gradR.p; }
        ops::flux(mv, R, gR, qf.p, C.err.p, s);
        kDivRow<<<blocksFor(nc), kBlock, 0, s>>>(mv, qf.p, q, divT.p);
    }
    kAddVectorSource<<<blocksFor(nc), kBlock, 0, s>>>(nc, M.V.p, divT.p, src);
    BSIM_LAUNCH_CHECK();
}

void KOmegaDevice::setTransient(int method) {
    if (method != 1 && method != 2) throw std::invalid_argument("transient k-omega needs Euler or BDF2");
    const int nc = C.mesh->nc;
    for (int q = 0; q < 2; ++q) { kh[q].ensure(nc); wh[q].ensure(nc); }
    timeMethod = method;
    hcur = 0;
    hlevels = 0;
    hdt = hdtPrev = 0;
}

void KOmegaDevice::saveOld(double dt) {
    if (!timeMethod) throw std::logic_error("saveOld needs the transient form");
    const int nc = C.mesh->nc;
    hipStream_t s = C.stream;
    hcur = 1 - hcur;
    kCopyD<<<blocksFor(nc), kBlock, 0, s>>>(nc, k.p, kh[hcur].p);
    kCopyD<<<blocksFor(nc), kBlock, 0, s>>>(nc, w.p, wh[hcur].p);
    if (hlevels == 0) {  // the older slot of a one-level history holds the same values (history.h:7)
        kCopyD<<<blocksFor(nc), kBlock, 0, s>>>(nc, k.p, kh[1 - hcur].p);
        kCopyD<<<blocksFor(nc), kBlock, 0, s>>>(nc, w.p, wh[1 - hcur].p);
    }
    BSIM_LAUNCH_CHECK();
    hdtPrev = hdt;
    hdt = dt;
    ++hlevels;
}

void KOmegaDevice::assemble(DeviceSystem& sys, FieldArg x, double* ctx, const double* phi, const TransportOpts& o,
                            const double* diffusivity, const double* diffusivityTrace, const double* rt,
                            const double* source, const TimeTerm* tt) {
    const DeviceMesh& M = *C.mesh;
    const MeshView mv = M.view();
    hipStream_t s = C.stream;
    const int nc = M.nc, nf = M.nf;
    double* diag = sys.candDiag();
    double* off = sys.candVal();
    kFillD<<<blocksFor(nc), kBlock, 0, s>>>(nc, diag, 0.0);
    kFillD<<<blocksFor(std::max(M.nnzSliced, 1)), kBlock, 0, s>>>(M.nnzSliced, off, 0.0);
    kFillD<<<blocksFor(nc), kBlock, 0, s>>>(nc, bsrc.p, 0.0);
    // equ::ddt(equation, rho, history) first (k_omega.cpp:82, 96)
    if (tt) kScalarTimeTerm<<<blocksFor(nc), kBlock, 0, s>>>(nc, M.V.p, rho, tt->a0, tt->a1, tt->a2, tt->prev, tt->older,
                                                             diag, bsrc.p);
    kCopyD<<<blocksFor(nf), kBlock, 0, s>>>(nf, phi, ctx);  // equ::div binds phi (procedural_equation.cpp:121-125)
    const bool needGrad = o.conv == kLinearUpwind || (o.conv == kCentral && o.interp == kCorrected) ||
                          o.diff != kOrthogonal;
    if (needGrad) ops::gradient(mv, M.orthogonal, 1, x, o.grad, gradX.p, work.p, C.err.p, s);
    ops::assembleConvection(mv, 1, x, phi, o.conv, o.interp, gradX.p, rho, diag, off, bsrc.p, C.err.p, s);
    const FieldArg D{diffusivity, zgType.p, zgVal.p, nullptr, diffusivityTrace};
    const double* gD = nullptr;
    if (o.coefInterp == kCorrected) { ops::gradient(mv, M.orthogonal, 1, D, o.coefGrad, gradR.p, work.p, C.err.p, s); gD = gradR.p; }
    ops::interpolate(mv, 1, D, o.coefInterp, gD, kf.p, C.err.p, s);
    ops::assembleDiffusion(mv, 1, x, kf.p, gradX.p, o.diff, -(-1.0), diag, off, bsrc.p, C.err.p, s);
    kReactionSource<<<blocksFor(nc), kBlock, 0, s>>>(nc, M.V.p, rt, source, diag, bsrc.p);
    BSIM_LAUNCH_CHECK();
}

double KOmegaDevice::residual(DeviceSystem& sys, const double* x) {
    const DeviceMesh& M = *C.mesh;
    hipStream_t s = C.stream;
    double* part = red.p + 16;
    kScalarResidual<<<reduceBlocks(M.nc), kBlock, 0, s>>>(M.view(), sys.candDiag(), sys.candVal(), bsrc.p, x, part);
    kReduceTail<3><<<1, kTail, 0, s>>>(reduceBlocks(M.nc), part, StoreSums<3>{red.p});
    BSIM_LAUNCH_CHECK();
    double h[3];
    BSIM_HIP(hipMemcpyAsync(h, red.p, sizeof(h), hipMemcpyDeviceToHost, s));
    ++reads;
    BSIM_HIP(hipStreamSynchronize(s));
    return std::sqrt(h[0]) / std::max(std::sqrt(h[1]) + std::sqrt(h[2]), 1e-30);
}

TransportRep KOmegaDevice::solveTransport(FieldArg U, const double* phi, int mGrad) {
    const DeviceMesh& M = *C.mesh;
    const MeshView mv = M.view();
    hipStream_t s = C.stream;
    const int nc = M.nc, nf = M.nf;
    TransportRep rep;
    TimeTerm tk, tw;
    const TimeTerm* ptk = nullptr;
    const TimeTerm* ptw = nullptr;
    if (timeMethod) {  // procedural_equation.cpp:198-239
        if (hlevels == 0) throw std::logic_error("saveOld must precede a transient assembly");
        const int method = timeMethod == 2 && hlevels < 2 ? 1 : timeMethod;
        double a0 = 1 / hdt, a1 = 1 / hdt, a2 = 0;
        if (method == 2) {
            const double olddt = hdtPrev == 0 ? hdt : hdtPrev;
            const double r = hdt / olddt;
            a0 = (1 + 2 * r) / ((1 + r) * hdt);
            a1 = (1 + r) / hdt;
            a2 = -r * r / ((1 + r) * hdt);
        }
        tk = TimeTerm{a0, a1, a2, kh[hcur].p, kh[1 - hcur].p};
        tw = TimeTerm{a0, a1, a2, wh[hcur].p, wh[1 - hcur].p};
        ptk = &tk;
        ptw = &tw;
    }
    kCopyD<<<blocksFor(nc), kBlock, 0, s>>>(nc, k.p, kPrev.p);
    kCopyD<<<blocksFor(nc), kBlock, 0, s>>>(nc, w.p, wPrev.p);
    // k_omega.cpp:62-77: coefficients frozen at the current iterate.
    updateViscosity();
    ops::gradient(mv, M.orthogonal, 3, U, mGrad, gradR.p, work.p, C.err.p, s);
    kTransportCoefficients<<<blocksFor(std::max(nc, nf - M.nfi)), kBlock, 0, s>>>(
        mv, gradR.p, mut.p, mutT.p, k.p, w.p, rho, mu, c.sigmaK, c.sigmaOmega, c.beta, c.gamma, c.kMin, prod.p,
        kd.p, wd.p, rate.p, srcW.p, kdT.p, wdT.p);
    BSIM_LAUNCH_CHECK();
    // omega (k_omega.cpp:80-93)
    assemble(*wSys, fieldW(), ctxW.p, phi, c.wOpts, wd.p, wdT.p, rate.p, srcW.p, ptw);
    rep.rOmega = residual(*wSys, w.p);
    kRelaxScalar<<<blocksFor(nc), kBlock, 0, s>>>(nc, wSys->candDiag(), bsrc.p, wPrev.p, c.relaxation);
    BSIM_LAUNCH_CHECK();
    wSys->update();
    wSys->prepare(c.wSolver);
    SolveRes r = wSys->solve(bsrc.p, w.p, 0);
    wHist = wSys->hist;
    C.checkErrors();
    rep.itOmega = r.iterations;
    rep.statusOmega = r.status;
    if (r.status == kNumericalFailure) { rep.healthy = false; return rep; }
    double bounded = boundOmega();
    // k with the new omega in its sink (k_omega.cpp:95-106)
    kRateK<<<blocksFor(nc), kBlock, 0, s>>>(nc, w.p, c.betaStar * rho, rate.p);
    assemble(*kSys, fieldK(), ctxK.p, phi, c.kOpts, kd.p, kdT.p, rate.p, prod.p, ptk);
    rep.rK = residual(*kSys, k.p);
    kRelaxScalar<<<blocksFor(nc), kBlock, 0, s>>>(nc, kSys->candDiag(), bsrc.p, kPrev.p, c.relaxation);
    BSIM_LAUNCH_CHECK();
    kSys->update();
    kSys->prepare(c.kSolver);
    r = kSys->solve(bsrc.p, k.p, 0);
    kHist = kSys->hist;
    C.checkErrors();
    rep.itK = r.iterations;
    rep.statusK = r.status;
    if (r.status == kNumericalFailure) { rep.healthy = false; return rep; }
    // bound and publish the viscosity (k_omega.cpp:108-111, 124-129)
    kCopyD<<<blocksFor(nf), kBlock, 0, s>>>(nf, phi, ctxK.p);
    kCopyD<<<blocksFor(nf), kBlock, 0, s>>>(nf, phi, ctxW.p);
    kBoundK<<<blocksFor(nc), kBlock, 0, s>>>(nc, k.p, c.kMin);
    BSIM_LAUNCH_CHECK();
    bounded = std::max(bounded, boundOmega());
    updateViscosity();
    double* part = red.p + 16;
    kChange<<<reduceBlocks(nc), kBlock, 0, s>>>(nc, k.p, kPrev.p, w.p, wPrev.p, part);
    kReduceTail<4><<<1, kTail, 0, s>>>(reduceBlocks(nc), part, StoreSums<4>{red.p});
    BSIM_LAUNCH_CHECK();
    double h[4];
    BSIM_HIP(hipMemcpyAsync(h, red.p, sizeof(h), hipMemcpyDeviceToHost, s));
    ++reads;
    C.checkErrors();
    rep.dK = std::sqrt(h[0]) / std::max(std::sqrt(h[1]), 1e-30);
    rep.dOmega = std::sqrt(h[2]) / std::max(std::sqrt(h[3]), 1e-30);
    rep.boundedOmega = bounded;
    rep.healthy = std::isfinite(rep.rOmega) && std::isfinite(rep.rK) && std::isfinite(rep.dOmega) &&
                  std::isfinite(rep.dK);
    return rep;
}

}  // namespace bsim
