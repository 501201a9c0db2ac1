// The procedural equation seam on the device (see equ.hpp).  Each routine
// cites the reference function it restates; the arithmetic of each update
// keeps the reference's operand order.
#include "equ.hpp"

#include "flow_kernels.hpp"

#include <algorithm>
#include <cmath>

namespace bsim {

namespace {

using namespace flow;

enum EquFlag : int { kFlagCapacity = 1, kFlagNonFinite = 2, kFlagDiagonal = 4 };

__global__ void __launch_bounds__(kBlock) kAccumulate(long long n, double f, const double* b, double* a) {  // a += f*b
    const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    if (i < n) a[i] += f * b[i];
}
__global__ void __launch_bounds__(kBlock) kScaleArr(long long n, double f, double* x) {
    const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    if (i < n) x[i] = f * x[i];
}
// reaction (procedural_equation.cpp:194-204): c*V, or factor*c_i*V with a finite c_i.
__global__ void __launch_bounds__(kBlock) kReaction(MeshView m, double* diag, const double* c, double value, double factor, int* flag) {
    const int i = globalThread();
    if (i >= m.nc) return;
    if (c) {
        if (!isfinite(c[i])) atomicOr(flag, kFlagNonFinite);
        diag[i] += factor * c[i] * m.V[i];
    } else {
        diag[i] += value * m.V[i];
    }
}
// source (procedural_equation.cpp:205-213): V*value, or factor*V*value_i.
__global__ void __launch_bounds__(kBlock) kSource(MeshView m, int nk, double* src, const double* field, double c0, double c1, double c2,
                        double factor) {
    const int i = globalThread();
    if (i >= m.nc) return;
    const double cst[3] = {c0, c1, c2};
    for (int k = 0; k < nk; ++k) {
        const long long q = static_cast<long long>(i) * nk + k;
        if (field) src[q] += factor * m.V[i] * field[q];
        else src[q] += m.V[i] * cst[k];
    }
}
// timeTerm (procedural_equation.cpp:214-240).
__global__ void __launch_bounds__(kBlock) kTime(MeshView m, int nk, double* diag, double* src, const double* capField, double cap,
                      const double* prev, const double* older, double a0, double a1, double a2, int* flag) {
    const int i = globalThread();
    if (i >= m.nc) return;
    const double c = capField ? capField[i] : cap;
    if (!(c > 0) || !isfinite(c)) atomicOr(flag, kFlagCapacity);
    const double w = c * m.V[i];
    diag[i] += w * a0;
    for (int k = 0; k < nk; ++k) {
        const long long q = static_cast<long long>(i) * nk + k;
        src[q] += w * a1 * prev[q];
        if (a2 != 0) src[q] += w * a2 * older[q];
    }
}
// relax (procedural_equation.cpp:317-324).
__global__ void __launch_bounds__(kBlock) kRelax(int n, int nk, double* diag, double* src, const double* prev, double alpha) {
    const int i = globalThread();
    if (i >= n) return;
    const double extra = diag[i] * (1 / alpha - 1);
    diag[i] += extra;
    for (int k = 0; k < nk; ++k) {
        const long long q = static_cast<long long>(i) * nk + k;
        src[q] += extra * prev[q];
    }
}
// reference(cell, value) (procedural_equation.cpp:325-334); skipped when the
// row-sum test found a row that is not a constant null mode.
__global__ void __launch_bounds__(kBlock) kRefApply(double* diag, double* src, int cell, double value, const int* notConstant, int* flag) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (notConstant && notConstant[0]) return;
    if (!(diag[cell] > 0)) {
        atomicOr(flag, kFlagDiagonal);
        return;
    }
    src[cell] += diag[cell] * value;
    diag[cell] *= 2;
}
// A x (and b - A x): the diagonal first, then the off-diagonal entries of the
// row in column order, per component.
__global__ void __launch_bounds__(kBlock) kApplyRows(MeshView m, int nk, const double* diag, const double* off, const double* x, const double* b,
                           double* y) {
    const int i = globalThread();
    if (i >= m.nc) return;
    double ax[3];
    for (int k = 0; k < nk; ++k) ax[k] = diag[i] * x[static_cast<long long>(i) * nk + k];
    for (int k_ = 0, e = 0; m.row(i).valid(k_) && (e = m.row(i).at(k_), true); ++k_) {
        const int c = m.mcol[e];
        if (c < 0) continue;
        for (int k = 0; k < nk; ++k) ax[k] += off[e] * x[static_cast<long long>(c) * nk + k];
    }
    for (int k = 0; k < nk; ++k) {
        const long long q = static_cast<long long>(i) * nk + k;
        y[q] = b ? b[q] - ax[k] : ax[k];
    }
}
// Partial sums of |b - A x|^2, |A x|^2, |b|^2 (relativeResidual, procedural_equation.cpp:447-452).
__global__ void __launch_bounds__(kBlock) kRelResidual(MeshView m, int nk, const double* diag, const double* off, const double* src,
                             const double* x, double* part) {
    __shared__ double smem[3 * kBlock / 64];
    double a[3] = {0, 0, 0};
    for (int i = globalThread(); i < m.nc; i += gridDim.x * blockDim.x) {
        double ax[3];
        for (int k = 0; k < nk; ++k) ax[k] = diag[i] * x[static_cast<long long>(i) * nk + k];
        for (int k_ = 0, e = 0; m.row(i).valid(k_) && (e = m.row(i).at(k_), true); ++k_) {
            const int c = m.mcol[e];
            if (c < 0) continue;
            for (int k = 0; k < nk; ++k) ax[k] += off[e] * x[static_cast<long long>(c) * nk + k];
        }
        for (int k = 0; k < nk; ++k) {
            const double b = src[static_cast<long long>(i) * nk + k], r = b - ax[k];
            a[0] += r * r;
            a[1] += ax[k] * ax[k];
            a[2] += b * b;
        }
    }
    blockPartial<3>(a, part, smem);
}
// One diffusion snapshot's face flux (procedural_equation.cpp:352-370): the
// orthogonal flux of the solution with the snapshot's boundary conditions and
// frozen coefficient, plus the frozen correction, times the factor.
__global__ void __launch_bounds__(kBlock) kFaceFluxAdd(MeshView m, FieldArg sol, const double* kf, const double* corr, double factor,
                             double* out, int* err) {
    const int f = globalThread();
    if (f >= m.nf) return;
    const double k = kf[f];
    if (!(k >= 0.0) || !isfinite(k)) atomicOr(err, kErrNegativeDiffusivity);
    double g[1];
    dev::integratedNormalGradient<1>(sol, nullptr, m, f, kOrthogonal, g, err);
    double inc = k * g[0];
    inc += corr[f];
    out[f] += factor * inc;
}
// Sums of squares of up to 4 fields (normsL2, field_math.cpp:30-52).
struct NormArgs {
    const double* v[4];
    long long n[4];
};
__global__ void __launch_bounds__(kBlock) kNorms(NormArgs a, int count, long long nmax, double* part) {
    __shared__ double smem[4 * kBlock / 64];
    double s[4] = {0, 0, 0, 0};
    for (long long i = globalThread(); i < nmax; i += static_cast<long long>(gridDim.x) * blockDim.x)
        for (int k = 0; k < count; ++k)
            if (i < a.n[k]) s[k] += a.v[k][i] * a.v[k][i];
    blockPartial<4>(s, part, smem);
}
// relativeChange / relativeMagnitude partial sums (fvm_execution.cpp:91-146):
// change: |a-b|^2 and |a|^2 per cell; magnitude: |a|^2 and |b|^2.
__global__ void __launch_bounds__(kBlock) kPairSums(int nc, int nk, const double* a, const double* b, int change, double* part) {
    __shared__ double smem[2 * kBlock / 64];
    double s[2] = {0, 0};
    for (int i = globalThread(); i < nc; i += gridDim.x * blockDim.x) {
        double d2 = 0, a2 = 0, b2 = 0;
        for (int k = 0; k < nk; ++k) {
            const long long q = static_cast<long long>(i) * nk + k;
            const double d = a[q] - b[q];
            d2 += d * d;
            a2 += a[q] * a[q];
            b2 += b[q] * b[q];
        }
        if (change) { s[0] += d2; s[1] += a2; }
        else { s[0] += a2; s[1] += b2; }
    }
    blockPartial<2>(s, part, smem);
}
// Block maximum of one value per thread, stored under the logical block index.
__device__ inline void blockMax(double v, double* partial, double* smem) {
    const double m = blockMaxValue(v, smem);
    if (threadIdx.x == 0) partial[xcdBlock()] = m;
}
__global__ void __launch_bounds__(kTail) kMaxTail(int nb, const double* partial, double* out) {
    __shared__ double smem[kTail / 64];
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += kTail) v = fmax(v, partial[b]);
    v = blockMaxValue<kTail>(v, smem);
    if (threadIdx.x == 0) out[0] = v;
}
// fluxBalance (fvm_execution.cpp:148-175): per cell the outward flux sum;
// partial sums of |imbalance|, imbalance^2 and |outward| and the block maximum
// of |imbalance|.
__global__ void __launch_bounds__(kBlock) kFluxBalance(MeshView m, const double* phi, double* part, double* maxPart) {
    __shared__ double smem[3 * kBlock / 64];
    double s[3] = {0, 0, 0}, mx = 0.0;
    for (int c = globalThread(); c < m.nc; c += gridDim.x * blockDim.x) {
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
        s[0] += fabs(imb);
        s[1] += imb * imb;
        s[2] += sc;
        mx = fmax(mx, fabs(imb));
    }
    blockPartial<3>(s, part, smem);
    blockMax(mx, maxPart, smem);
}
// courantNumber (fvm_execution.cpp:177-199): co = 0.5 dt sum|phi_f| / V per
// cell; partial sums of co V and V and the block maximum of co.
__global__ void __launch_bounds__(kBlock) kCourant(MeshView m, const double* phi, double dt, double* part, double* maxPart) {
    __shared__ double smem[2 * kBlock / 64];
    double s[2] = {0, 0}, mx = 0.0;
    for (int c = globalThread(); c < m.nc; c += gridDim.x * blockDim.x) {
        double sum = 0.0;
        for (int k_ = 0, e = 0; m.row(c).valid(k_) && (e = m.row(c).at(k_), true); ++k_) {
            const int f = m.mface[e];
            if (f >= 0) sum += fabs(phi[f]);
        }
        for (int e = m.bptr[c]; e < m.bptr[c + 1]; ++e) sum += fabs(phi[m.bface[e]]);
        const double v = m.V[c];
        const double co = 0.5 * dt * sum / v;
        mx = fmax(mx, co);
        s[0] += co * v;
        s[1] += v;
    }
    blockPartial<2>(s, part, smem);
    blockMax(mx, maxPart, smem);
}
__global__ void __launch_bounds__(kBlock) kFaceAdd(int n, double factor, const double* inc, double* target) {
    const int f = globalThread();
    if (f < n) target[f] += factor * inc[f];
}

void finite(double v) {
    if (!std::isfinite(v)) throw std::invalid_argument("nonfinite equ coefficient");
}

// Reduction scratch of a context for the free functions below: partials of
// up to 4 sums, then the block maxima, then the totals.
struct Reducer {
    DBuf<double> buf;
    explicit Reducer(hipStream_t s) { buf.alloc(5 * static_cast<std::size_t>(kMaxReduceBlocks) + 8); (void)s; }
    double* part() { return buf.p; }
    double* maxPart() { return buf.p + 4 * static_cast<std::size_t>(kMaxReduceBlocks); }
    double* out() { return buf.p + 5 * static_cast<std::size_t>(kMaxReduceBlocks); }
};

}  // namespace

// ------------------------------------------------------------------ fields

DeviceField::DeviceField(Context& ctx, int loc, int nk) : C(ctx), location(loc), ncomp(nk) {
    if (loc != kCellField && loc != kFaceField) throw std::invalid_argument("fields live on cells or faces");
    if (loc == kCellField && nk != 1 && nk != 3) throw std::invalid_argument("cell fields have 1 or 3 components");
    if (loc == kFaceField && nk != 1) throw std::invalid_argument("face fields have one value per face");
    v.alloc(std::max<std::size_t>(size(), 1));
    kFillD<<<blocksFor(static_cast<long long>(size())), kBlock, 0, C.stream>>>(static_cast<long long>(size()), v.p, 0.0);
    if (loc == kCellField) {
        const int np = C.mesh->np;
        std::vector<int> t(np, kZeroGradient);
        std::vector<double> z(static_cast<std::size_t>(np) * nk, 0.0);
        setBoundary(t.data(), z.data());
    }
    BSIM_LAUNCH_CHECK();
}

void DeviceField::setBoundaryFlux(const DeviceField& flux) {
    if (location != kCellField || &flux.C != &C || flux.location != kFaceField)
        throw std::invalid_argument("boundary flux must be a face field on the same mesh");
    const int nf = C.mesh->nf;
    ctx.ensure(nf);
    kCopyD<<<blocksFor(nf), kBlock, 0, C.stream>>>(nf, flux.v.p, ctx.p);
    BSIM_LAUNCH_CHECK();
    hasCtx = true;
}

std::size_t DeviceField::size() const {
    return location == kCellField ? static_cast<std::size_t>(C.mesh->nc) * ncomp : static_cast<std::size_t>(C.mesh->nf);
}

void DeviceField::setBoundary(const int* types, const double* values) {
    if (location != kCellField) throw std::invalid_argument("face fields carry no boundary conditions");
    const int np = C.mesh->np;
    for (int p = 0; p < np; ++p)
        if (types[p] < 0 || types[p] > 4) throw std::invalid_argument("invalid boundary type");
    hbct.assign(types, types + np);
    bct.upload(types, np, C.stream);
    bcv.upload(values, static_cast<std::size_t>(np) * ncomp, C.stream);
    BSIM_HIP(hipStreamSynchronize(C.stream));
}

void DeviceField::setValues(const double* caller) {
    const std::vector<double> in = location == kCellField ? C.cellsIn(caller, ncomp) : C.facesIn(caller, 1);
    for (double x : in)
        if (!std::isfinite(x)) throw std::invalid_argument("field values must be finite");
    v.upload(in, C.stream);
    BSIM_HIP(hipStreamSynchronize(C.stream));
}

void DeviceField::getValues(double* caller) const {
    const std::vector<double> h = v.toHost(C.stream, size());
    if (location == kCellField) C.cellsOut(h, caller, ncomp);
    else C.facesOut(h, caller, 1);
}

// --------------------------------------------------------------- equations

DeviceEquation::DeviceEquation(DeviceField& unknown, const EquOptions& o, const SolverCfg& lin)
    : x(unknown), ncomp(unknown.ncomp), opt(o), linear(lin) {
    if (unknown.location != kCellField) throw std::invalid_argument("equ requires compatible cell fields");
    const DeviceMesh& M = *x.C.mesh;
    diag.alloc(M.nc);
    off.alloc(std::max(M.nnzSliced, 1));
    src.alloc(static_cast<std::size_t>(M.nc) * ncomp);
    red.alloc(4 * static_cast<std::size_t>(kMaxReduceBlocks) + 16);
    reset();
}

void DeviceEquation::requireCell(const DeviceField& f, int nk) const {
    if (&f.C != &x.C || f.location != kCellField || (nk > 0 && f.ncomp != nk))
        throw std::invalid_argument("equ requires compatible cell fields");
}

const double* DeviceEquation::gradientOf(const DeviceField& f, int method) {
    const DeviceMesh& M = *x.C.mesh;
    const std::size_t n = 3 * static_cast<std::size_t>(f.ncomp) * M.nc;
    grad.ensure(n);
    work.ensure(n);
    ops::gradient(M.view(), M.orthogonal, f.ncomp, f.arg(), method, grad.p, work.p, x.C.err.p, x.C.stream);
    return grad.p;
}

void DeviceEquation::reset() {
    const DeviceMesh& M = *x.C.mesh;
    hipStream_t s = x.C.stream;
    kFillD<<<blocksFor(M.nc), kBlock, 0, s>>>(M.nc, diag.p, 0.0);
    kFillD<<<blocksFor(std::max(M.nnzSliced, 1)), kBlock, 0, s>>>(std::max(M.nnzSliced, 1), off.p, 0.0);
    kFillD<<<blocksFor(static_cast<long long>(M.nc) * ncomp), kBlock, 0, s>>>(static_cast<long long>(M.nc) * ncomp,
                                                                            src.p, 0.0);
    BSIM_LAUNCH_CHECK();
    diffusion.clear();
    boundaryFlux = nullptr;
}

std::unique_ptr<DeviceEquation> DeviceEquation::copy() const {
    auto b = std::make_unique<DeviceEquation>(x, opt, linear);
    const DeviceMesh& M = *x.C.mesh;
    hipStream_t s = x.C.stream;
    kCopyD<<<blocksFor(M.nc), kBlock, 0, s>>>(M.nc, diag.p, b->diag.p);
    kCopyD<<<blocksFor(std::max(M.nnzSliced, 1)), kBlock, 0, s>>>(std::max(M.nnzSliced, 1), off.p, b->off.p);
    kCopyD<<<blocksFor(static_cast<long long>(M.nc) * ncomp), kBlock, 0, s>>>(static_cast<long long>(M.nc) * ncomp,
                                                                            src.p, b->src.p);
    BSIM_LAUNCH_CHECK();
    b->diffusion = diffusion;
    b->boundaryFlux = boundaryFlux;
    return b;
}

void DeviceEquation::div(const DeviceField& phi, double factor) {
    finite(factor);
    if (&phi.C != &x.C || phi.location != kFaceField) throw std::invalid_argument("equ::div requires a face flux");
    if (boundaryFlux && boundaryFlux != &phi)
        throw std::invalid_argument("one equation requires one boundary flux context");
    boundaryFlux = &phi;
    x.setBoundaryFlux(phi);
    const DeviceMesh& M = *x.C.mesh;
    const bool needGrad = opt.conv == kLinearUpwind || (opt.conv == kCentral && opt.interp == kCorrected);
    const double* g = needGrad ? gradientOf(x, opt.grad) : nullptr;
    ops::assembleConvection(M.view(), ncomp, x.arg(), phi.v.p, opt.conv, opt.interp, g, factor, diag.p, off.p, src.p,
                            x.C.err.p, x.C.stream);
    BSIM_LAUNCH_CHECK();
    x.C.checkErrors();
}

void DeviceEquation::laplacian(const DeviceField* k, double kc, double multiplier) {
    finite(multiplier);
    const DeviceMesh& M = *x.C.mesh;
    const MeshView mv = M.view();
    hipStream_t s = x.C.stream;
    // The face diffusivity: a constant, a face field, or a cell field
    // interpolated with the coefficient options (procedural_equation.cpp:155-166).
    kfBuf.ensure(M.nf);
    if (!k) {
        finite(kc);
        kFillD<<<blocksFor(M.nf), kBlock, 0, s>>>(M.nf, kfBuf.p, kc);
    } else if (k->location == kFaceField) {
        if (&k->C != &x.C) throw std::invalid_argument("equ diffusivity must be a compatible cell or face field");
        kCopyD<<<blocksFor(M.nf), kBlock, 0, s>>>(M.nf, k->v.p, kfBuf.p);
    } else {
        requireCell(*k, 1);
        const double* gk = opt.coefInterp == kCorrected ? gradientOf(*k, opt.coefGrad) : nullptr;
        ops::interpolate(mv, 1, k->arg(), opt.coefInterp, gk, kfBuf.p, x.C.err.p, s);
    }
    const bool corrected = opt.diff != kOrthogonal;
    const double* g = (corrected || ncomp == 3) ? gradientOf(x, opt.grad) : nullptr;
    ops::assembleDiffusion(mv, ncomp, x.arg(), kfBuf.p, g, opt.diff, -multiplier, diag.p, off.p, src.p, x.C.err.p, s);
    if (ncomp == 1) {
        auto snap = std::make_shared<Snapshot>();
        snap->kf.alloc(M.nf);
        snap->corr.alloc(M.nf);
        kCopyD<<<blocksFor(M.nf), kBlock, 0, s>>>(M.nf, kfBuf.p, snap->kf.p);
        if (corrected) {
            // diffusionCorrection (procedural_equation.cpp:77-87): full minus orthogonal flux.
            work.ensure(M.nf);
            ops::diffusionFlux(mv, x.arg(), kfBuf.p, g, opt.diff, snap->corr.p, x.C.err.p, s);
            ops::diffusionFlux(mv, x.arg(), kfBuf.p, g, kOrthogonal, work.p, x.C.err.p, s);
            kCorrection<<<blocksFor(M.nf), kBlock, 0, s>>>(M.nf, snap->corr.p, work.p, snap->corr.p);
        } else {
            kFillD<<<blocksFor(M.nf), kBlock, 0, s>>>(M.nf, snap->corr.p, 0.0);
        }
        snap->bct.upload(x.hbct, s);
        snap->bcv.alloc(M.np);
        kCopyD<<<1, kBlock, 0, s>>>(M.np, x.bcv.p, snap->bcv.p);
        if (x.hasCtx) {
            snap->ctx.alloc(M.nf);
            kCopyD<<<blocksFor(M.nf), kBlock, 0, s>>>(M.nf, x.ctx.p, snap->ctx.p);
            snap->hasCtx = true;
        }
        diffusion.push_back({std::move(snap), multiplier});
    }
    BSIM_LAUNCH_CHECK();
    x.C.checkErrors();
}

void DeviceEquation::reaction(const DeviceField* c, double value, double factor) {
    finite(factor);
    if (c) requireCell(*c, 1);
    else finite(value);
    const DeviceMesh& M = *x.C.mesh;
    int* flag = reinterpret_cast<int*>(red.p);
    BSIM_HIP(hipMemsetAsync(flag, 0, sizeof(int), x.C.stream));
    kReaction<<<blocksFor(M.nc), kBlock, 0, x.C.stream>>>(M.view(), diag.p, c ? c->v.p : nullptr, value, factor, flag);
    BSIM_LAUNCH_CHECK();
    int h = 0;
    BSIM_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, x.C.stream));
    BSIM_HIP(hipStreamSynchronize(x.C.stream));
    if (h) throw std::invalid_argument("nonfinite equ coefficient");
}

void DeviceEquation::source(const DeviceField* value, const double* constant, double factor) {
    finite(factor);
    double c[3] = {0, 0, 0};
    if (value) requireCell(*value, ncomp);
    else
        for (int k = 0; k < ncomp; ++k) c[k] = constant[k];
    const DeviceMesh& M = *x.C.mesh;
    kSource<<<blocksFor(M.nc), kBlock, 0, x.C.stream>>>(M.view(), ncomp, src.p, value ? value->v.p : nullptr, c[0],
                                                        c[1], c[2], factor);
    BSIM_LAUNCH_CHECK();
}

void DeviceEquation::ddt(const DeviceField* capacity, double cap, const DeviceField& previous, double dt, int method,
                         const DeviceField* older, double olddt) {
    requireCell(previous, ncomp);
    if (method == 0) return;  // Steady
    if (!(dt > 0) || !std::isfinite(dt)) throw std::invalid_argument("equ::ddt requires positive finite dt");
    if (capacity) requireCell(*capacity, 1);
    double a0 = 1 / dt, a1 = 1 / dt, a2 = 0;
    if (method == 2) {
        if (!older) throw std::invalid_argument("BDF2 requires explicit older field; use Euler to start");
        requireCell(*older, ncomp);
        if (olddt == 0) olddt = dt;
        if (!(olddt > 0) || !std::isfinite(olddt)) throw std::invalid_argument("invalid previous dt");
        const double r = dt / olddt;
        a0 = (1 + 2 * r) / ((1 + r) * dt);
        a1 = (1 + r) / dt;
        a2 = -r * r / ((1 + r) * dt);
    } else if (method != 1) {
        throw std::invalid_argument("unknown time method");
    }
    const DeviceMesh& M = *x.C.mesh;
    int* flag = reinterpret_cast<int*>(red.p);
    BSIM_HIP(hipMemsetAsync(flag, 0, sizeof(int), x.C.stream));
    kTime<<<blocksFor(M.nc), kBlock, 0, x.C.stream>>>(M.view(), ncomp, diag.p, src.p, capacity ? capacity->v.p : nullptr,
                                                      cap, previous.v.p, older ? older->v.p : nullptr, a0, a1, a2, flag);
    BSIM_LAUNCH_CHECK();
    int h = 0;
    BSIM_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, x.C.stream));
    BSIM_HIP(hipStreamSynchronize(x.C.stream));
    if (h) throw std::invalid_argument("time capacity must be positive and finite");
}

void DeviceEquation::scale(double f) {
    finite(f);
    const DeviceMesh& M = *x.C.mesh;
    hipStream_t s = x.C.stream;
    kScaleArr<<<blocksFor(M.nc), kBlock, 0, s>>>(M.nc, f, diag.p);
    kScaleArr<<<blocksFor(std::max(M.nnzSliced, 1)), kBlock, 0, s>>>(M.nnzSliced, f, off.p);
    kScaleArr<<<blocksFor(static_cast<long long>(M.nc) * ncomp), kBlock, 0, s>>>(static_cast<long long>(M.nc) * ncomp,
                                                                               f, src.p);
    BSIM_LAUNCH_CHECK();
    for (Diffusion& d : diffusion) d.factor *= f;
}

void DeviceEquation::add(const DeviceEquation& b, double f) {
    if (&b.x != &x) throw std::invalid_argument("matrix addition requires same unknown binding");
    if (b.boundaryFlux) {
        if (boundaryFlux && boundaryFlux != b.boundaryFlux)
            throw std::invalid_argument("one equation requires one boundary flux context");
        boundaryFlux = b.boundaryFlux;
    }
    finite(f);
    std::vector<Diffusion> contributions = b.diffusion;  // self-addition safe
    const DeviceMesh& M = *x.C.mesh;
    hipStream_t s = x.C.stream;
    kAccumulate<<<blocksFor(M.nc), kBlock, 0, s>>>(M.nc, f, b.diag.p, diag.p);
    kAccumulate<<<blocksFor(static_cast<long long>(M.nc) * ncomp), kBlock, 0, s>>>(static_cast<long long>(M.nc) * ncomp,
                                                                                 f, b.src.p, src.p);
    kAccumulate<<<blocksFor(std::max(M.nnzSliced, 1)), kBlock, 0, s>>>(M.nnzSliced, f, b.off.p, off.p);
    BSIM_LAUNCH_CHECK();
    for (Diffusion& d : contributions) {
        d.factor *= f;
        diffusion.push_back(std::move(d));
    }
}

void DeviceEquation::relax(const DeviceField& previous, double alpha) {
    if (!(alpha > 0 && alpha <= 1)) throw std::invalid_argument("relaxation must be in (0,1]");
    requireCell(previous, ncomp);
    const DeviceMesh& M = *x.C.mesh;
    kRelax<<<blocksFor(M.nc), kBlock, 0, x.C.stream>>>(M.nc, ncomp, diag.p, src.p, previous.v.p, alpha);
    BSIM_LAUNCH_CHECK();
}

void DeviceEquation::reference(int callerCell, double value) {
    finite(value);
    if (ncomp != 1) throw std::invalid_argument("reference requires a scalar equation");
    if (callerCell < 0 || callerCell >= x.C.mesh->nc) throw std::invalid_argument("reference global cell does not exist");
    int* flag = reinterpret_cast<int*>(red.p);
    BSIM_HIP(hipMemsetAsync(flag, 0, sizeof(int), x.C.stream));
    kRefApply<<<1, 64, 0, x.C.stream>>>(diag.p, src.p, x.C.cellOldToNew[callerCell], value, nullptr, flag);
    BSIM_LAUNCH_CHECK();
    int h = 0;
    BSIM_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, x.C.stream));
    BSIM_HIP(hipStreamSynchronize(x.C.stream));
    if (h) throw std::invalid_argument("reference requires a positive diagonal");
}

bool DeviceEquation::referenceIfUnanchored(double value) {
    finite(value);
    if (ncomp != 1) throw std::invalid_argument("reference requires a scalar equation");
    const DeviceMesh& M = *x.C.mesh;
    int* flags = reinterpret_cast<int*>(red.p);  // [0] not constant, [1] diagonal
    BSIM_HIP(hipMemsetAsync(flags, 0, 2 * sizeof(int), x.C.stream));
    kReferenceTest<<<blocksFor(M.nc), kBlock, 0, x.C.stream>>>(M.view(), diag.p, off.p, flags);
    kRefApply<<<1, 64, 0, x.C.stream>>>(diag.p, src.p, x.C.cellOldToNew[0], value, flags, flags + 1);
    BSIM_LAUNCH_CHECK();
    int h[2] = {0, 0};
    BSIM_HIP(hipMemcpyAsync(h, flags, sizeof(h), hipMemcpyDeviceToHost, x.C.stream));
    BSIM_HIP(hipStreamSynchronize(x.C.stream));
    if (h[1]) throw std::invalid_argument("reference requires a positive diagonal");
    return h[0] == 0;
}

void DeviceEquation::diagonal(DeviceField& out) const {
    requireCell(out, 1);
    kCopyD<<<blocksFor(x.C.mesh->nc), kBlock, 0, x.C.stream>>>(x.C.mesh->nc, diag.p, out.v.p);
    BSIM_LAUNCH_CHECK();
}

void DeviceEquation::rhs(DeviceField& out) const {
    requireCell(out, ncomp);
    const long long n = static_cast<long long>(x.C.mesh->nc) * ncomp;
    kCopyD<<<blocksFor(n), kBlock, 0, x.C.stream>>>(n, src.p, out.v.p);
    BSIM_LAUNCH_CHECK();
}

void DeviceEquation::apply(const DeviceField& xin, DeviceField& out) const {
    requireCell(xin, ncomp);
    requireCell(out, ncomp);
    if (&xin == &out) throw std::invalid_argument("apply needs an output field distinct from its input");
    const DeviceMesh& M = *x.C.mesh;
    kApplyRows<<<blocksFor(M.nc), kBlock, 0, x.C.stream>>>(M.view(), ncomp, diag.p, off.p, xin.v.p, nullptr, out.v.p);
    BSIM_LAUNCH_CHECK();
}

void DeviceEquation::residual(const DeviceField& xin, DeviceField& out) const {
    requireCell(xin, ncomp);
    requireCell(out, ncomp);
    if (&xin == &out) throw std::invalid_argument("residual needs an output field distinct from its input");
    const DeviceMesh& M = *x.C.mesh;
    kApplyRows<<<blocksFor(M.nc), kBlock, 0, x.C.stream>>>(M.view(), ncomp, diag.p, off.p, xin.v.p, src.p, out.v.p);
    BSIM_LAUNCH_CHECK();
}

double DeviceEquation::relativeResidual(const DeviceField& xin) const {
    requireCell(xin, ncomp);
    const DeviceMesh& M = *x.C.mesh;
    const int rb = reduceBlocks(M.nc);
    kRelResidual<<<rb, kBlock, 0, x.C.stream>>>(M.view(), ncomp, diag.p, off.p, src.p, xin.v.p, red.p);
    double* tot = red.p + 4 * static_cast<std::size_t>(kMaxReduceBlocks);
    kReduceTail<3><<<1, kTail, 0, x.C.stream>>>(rb, red.p, StoreSums<3>{tot});
    BSIM_LAUNCH_CHECK();
    double h[3];
    BSIM_HIP(hipMemcpyAsync(h, tot, sizeof(h), hipMemcpyDeviceToHost, x.C.stream));
    BSIM_HIP(hipStreamSynchronize(x.C.stream));
    return std::sqrt(h[0]) / std::max(std::sqrt(h[1]) + std::sqrt(h[2]), 1e-30);
}

void DeviceEquation::faceFlux(const DeviceField& solution, DeviceField& out) const {
    if (ncomp != 1) throw std::invalid_argument("faceFlux requires a scalar equation");
    requireCell(solution, 1);
    if (&out.C != &x.C || out.location != kFaceField) throw std::invalid_argument("faceFlux writes a face field");
    const DeviceMesh& M = *x.C.mesh;
    kFillD<<<blocksFor(M.nf), kBlock, 0, x.C.stream>>>(M.nf, out.v.p, 0.0);
    for (const Diffusion& d : diffusion) {
        // The solution's values with the boundary conditions frozen at assembly.
        const FieldArg sol{solution.v.p, d.snap->bct.p, d.snap->bcv.p, d.snap->hasCtx ? d.snap->ctx.p : nullptr,
                           nullptr};
        kFaceFluxAdd<<<blocksFor(M.nf), kBlock, 0, x.C.stream>>>(M.view(), sol, d.snap->kf.p, d.snap->corr.p, d.factor,
                                                                 out.v.p, x.C.err.p);
    }
    BSIM_LAUNCH_CHECK();
    x.C.checkErrors();
}

void DeviceEquation::solve(const SolverCfg* cfg, SolveRes* results) {
    const DeviceMesh& M = *x.C.mesh;
    hipStream_t s = x.C.stream;
    if (!sys) sys = std::make_unique<DeviceSystem>(s, meshPattern(M));
    BSIM_HIP(hipMemcpyAsync(sys->candDiag(), diag.p, sizeof(double) * M.nc, hipMemcpyDeviceToDevice, s));
    if (M.nnzSliced)
        BSIM_HIP(hipMemcpyAsync(sys->candVal(), off.p, sizeof(double) * M.nnzSliced, hipMemcpyDeviceToDevice, s));
    sys->update();
    sys->prepare(cfg ? *cfg : linear);
    if (ncomp == 1) {
        results[0] = sys->solve(src.p, x.v.p, 0);
        return;
    }
    for (int k = 0; k < 6; ++k) comp[k].ensure(M.nc);
    const int nb = blocksFor(M.nc);
    kSplit<<<nb, kBlock, 0, s>>>(M.nc, src.p, comp[0].p, comp[1].p, comp[2].p);
    kSplit<<<nb, kBlock, 0, s>>>(M.nc, x.v.p, comp[3].p, comp[4].p, comp[5].p);
    const double* const b3[3] = {comp[0].p, comp[1].p, comp[2].p};
    double* const x3[3] = {comp[3].p, comp[4].p, comp[5].p};
    std::vector<double> hists[3];
    sys->solve3(b3, x3, results, hists);
    kMerge<<<nb, kBlock, 0, s>>>(M.nc, comp[3].p, comp[4].p, comp[5].p, x.v.p);
    BSIM_LAUNCH_CHECK();
    BSIM_HIP(hipStreamSynchronize(s));
}

// -------------------------------------------------------------- reductions

std::vector<double> normsL2(const std::vector<const DeviceField*>& fields) {
    std::vector<double> out;
    if (fields.empty()) return out;
    Context& C = fields[0]->C;
    Reducer R(C.stream);
    for (std::size_t a = 0; a < fields.size(); a += 4) {
        NormArgs na{};
        const int count = static_cast<int>(std::min<std::size_t>(4, fields.size() - a));
        long long nmax = 0;
        for (int k = 0; k < count; ++k) {
            const DeviceField* f = fields[a + k];
            if (&f->C != &C) throw std::invalid_argument("math norm batch requires run mesh cell/face fields");
            na.v[k] = f->v.p;
            na.n[k] = static_cast<long long>(f->size());
            nmax = std::max(nmax, na.n[k]);
        }
        const int rb = reduceBlocks(nmax);
        kNorms<<<rb, kBlock, 0, C.stream>>>(na, count, nmax, R.part());
        kReduceTail<4><<<1, kTail, 0, C.stream>>>(rb, R.part(), StoreSums<4>{R.out()});
        BSIM_LAUNCH_CHECK();
        double h[4];
        BSIM_HIP(hipMemcpyAsync(h, R.out(), sizeof(h), hipMemcpyDeviceToHost, C.stream));
        BSIM_HIP(hipStreamSynchronize(C.stream));
        for (int k = 0; k < count; ++k) out.push_back(std::sqrt(h[k]));
    }
    return out;
}

namespace {
double pairRatio(const DeviceField& a, const DeviceField& b, int change) {
    if (&a.C != &b.C || a.location != kCellField || b.location != kCellField || a.ncomp != b.ncomp)
        throw std::invalid_argument("relative measures require two compatible cell fields");
    Context& C = a.C;
    Reducer R(C.stream);
    const int rb = reduceBlocks(C.mesh->nc);
    kPairSums<<<rb, kBlock, 0, C.stream>>>(C.mesh->nc, a.ncomp, a.v.p, b.v.p, change, R.part());
    kReduceTail<2><<<1, kTail, 0, C.stream>>>(rb, R.part(), StoreSums<2>{R.out()});
    BSIM_LAUNCH_CHECK();
    double h[2];
    BSIM_HIP(hipMemcpyAsync(h, R.out(), sizeof(h), hipMemcpyDeviceToHost, C.stream));
    BSIM_HIP(hipStreamSynchronize(C.stream));
    return std::sqrt(h[0]) / std::max(std::sqrt(h[1]), 1e-30);
}
}  // namespace

double relativeChange(const DeviceField& current, const DeviceField& previous) {
    return pairRatio(current, previous, 1);
}

double relativeMagnitude(const DeviceField& value, const DeviceField& reference) {
    return pairRatio(value, reference, 0);
}

void fluxBalance(const DeviceField& phi, double out[4]) {
    if (phi.location != kFaceField) throw std::invalid_argument("face flux");
    Context& C = phi.C;
    Reducer R(C.stream);
    const int rb = reduceBlocks(C.mesh->nc);
    kFluxBalance<<<rb, kBlock, 0, C.stream>>>(C.mesh->view(), phi.v.p, R.part(), R.maxPart());
    kReduceTail<3><<<1, kTail, 0, C.stream>>>(rb, R.part(), StoreSums<3>{R.out()});
    kMaxTail<<<1, kTail, 0, C.stream>>>(rb, R.maxPart(), R.out() + 3);
    BSIM_LAUNCH_CHECK();
    double h[4];
    BSIM_HIP(hipMemcpyAsync(h, R.out(), sizeof(h), hipMemcpyDeviceToHost, C.stream));
    BSIM_HIP(hipStreamSynchronize(C.stream));
    out[0] = h[0];
    out[1] = std::sqrt(h[1]);
    out[2] = h[3];
    out[3] = h[0] / std::max(h[2], 1e-30);
}

void courantNumber(const DeviceField& phi, double dt, double out[2]) {
    if (phi.location != kFaceField) throw std::invalid_argument("Courant flux");
    if (!(dt > 0.0) || !std::isfinite(dt)) throw std::invalid_argument("Courant time step must be finite and positive");
    Context& C = phi.C;
    Reducer R(C.stream);
    const int rb = reduceBlocks(C.mesh->nc);
    kCourant<<<rb, kBlock, 0, C.stream>>>(C.mesh->view(), phi.v.p, dt, R.part(), R.maxPart());
    kReduceTail<2><<<1, kTail, 0, C.stream>>>(rb, R.part(), StoreSums<2>{R.out()});
    kMaxTail<<<1, kTail, 0, C.stream>>>(rb, R.maxPart(), R.out() + 2);
    BSIM_LAUNCH_CHECK();
    double h[3];
    BSIM_HIP(hipMemcpyAsync(h, R.out(), sizeof(h), hipMemcpyDeviceToHost, C.stream));
    BSIM_HIP(hipStreamSynchronize(C.stream));
    out[0] = h[2];
    out[1] = h[0] / h[1];
}

void faceIncrement(const DeviceField& increment, DeviceField& target, double factor, int region) {
    if (&increment.C != &target.C || increment.location != kFaceField || target.location != kFaceField)
        throw std::invalid_argument("face increment requires matching run mesh face fields");
    if (region != 0 && region != 1) throw std::invalid_argument("invalid face region");
    const DeviceMesh& M = *target.C.mesh;
    const int n = region == 1 ? M.nfi : M.nf;  // interior faces come first
    kFaceAdd<<<blocksFor(std::max(n, 1)), kBlock, 0, target.C.stream>>>(n, factor, increment.v.p, target.v.p);
    BSIM_LAUNCH_CHECK();
}

}  // namespace bsim
