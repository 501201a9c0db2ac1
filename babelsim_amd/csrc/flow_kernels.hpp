// Kernels shared by the SIMPLE and PISO drivers: field copies, the momentum
// pressure source and residual, calculated-field traces, the pressure
// equation's source and reference, and the frozen pressure face flux.  Each
// names the reference line it follows.
#pragma once

#include "fvm_device.hpp"

namespace bsim {
// Internal linkage: every translation unit that includes this header gets
// its own copy of these kernels.
namespace flow {
namespace {

using namespace dev;

__global__ void kFillD(long long n, double* x, double v) {
    const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    if (i < n) x[i] = v;
}
__global__ void kCopyD(long long n, const double* a, double* b) {
    const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
    if (i < n) b[i] = a[i];
}
// Vec3 cell arrays <-> one array per component.
__global__ void kSplit(int n, const double* v3, double* c0, double* c1, double* c2) {
    const int i = globalThread();
    if (i >= n) return;
    c0[i] = v3[3 * static_cast<long long>(i)];
    c1[i] = v3[3 * static_cast<long long>(i) + 1];
    c2[i] = v3[3 * static_cast<long long>(i) + 2];
}
__global__ void kMerge(int n, const double* c0, const double* c1, const double* c2, double* v3) {
    const int i = globalThread();
    if (i >= n) return;
    v3[3 * static_cast<long long>(i)] = c0[i];
    v3[3 * static_cast<long long>(i) + 1] = c1[i];
    v3[3 * static_cast<long long>(i) + 2] = c2[i];
}

// equ::source(momentum, -grad(p)): source += (1*V) * (-1*gradP)  (procedural_equation.cpp:192-196)
__global__ void kMomentumPressureSource(MeshView m, const double* gradP, double* src) {
    const int i = globalThread();
    if (i >= m.nc) return;
    const double v = 1.0 * m.V[i];
    for (int k = 0; k < 3; ++k) src[3 * static_cast<long long>(i) + k] += v * (-1.0 * gradP[3 * static_cast<long long>(i) + k]);
}

// diagnostics::relativeResidual partial sums of |b-Ax|^2, |Ax|^2, |b|^2
// (procedural_equation.cpp:429-433) on the unrelaxed momentum rows.
__global__ void kMomentumResidual(MeshView m, const double* diag, const double* off, const double* src, const double* U,
                                  double* part) {
    __shared__ double smem[3 * kBlock / 64];
    double a[3] = {0, 0, 0};
    for (int i = globalThread(); i < m.nc; i += gridDim.x * blockDim.x) {
        double ax[3];
        for (int k = 0; k < 3; ++k) ax[k] = diag[i] * U[3 * static_cast<long long>(i) + k];
        for (int k_ = 0, e = 0; m.row(i).valid(k_) && (e = m.row(i).at(k_), true); ++k_) {
            if (m.mcol[e] < 0) continue;
            for (int k = 0; k < 3; ++k) ax[k] += off[e] * U[3 * static_cast<long long>(m.mcol[e]) + k];
        }
        double r2 = 0, a2 = 0, b2 = 0;
        for (int k = 0; k < 3; ++k) {
            const double b = src[3 * static_cast<long long>(i) + k];
            const double r = b - ax[k];
            r2 += r * r; a2 += ax[k] * ax[k]; b2 += b * b;
        }
        a[0] += r2; a[1] += a2; a[2] += b2;
    }
    blockPartial<3>(a, part, smem);
}

// Calculated-field traces of V/aP: V_o / aP_o on boundary faces.
__global__ void kOwnerTrace(MeshView m, const double* cell, double* trace) {
    const int f = m.nfi + globalThread();
    if (f < m.nf) trace[f] = cell[m.own[f]];
}

// equ::source(pEq, -div(phiH)) (procedural_equation.cpp:192-196)
__global__ void kScalarSource(MeshView m, const double* divPhiH, double* src) {
    const int i = globalThread();
    if (i < m.nc) src[i] += (1.0 * m.V[i]) * (-1.0 * divPhiH[i]);
}

// referenceIfUnanchored constant-null-mode test (procedural_equation.cpp:316-333).
__global__ void kReferenceTest(MeshView m, const double* diag, const double* off, int* notConstant) {
    const int i = globalThread();
    if (i >= m.nc) return;
    double rowSum = diag[i], mag = fabs(diag[i]);
    for (int k_ = 0, e = 0; m.row(i).valid(k_) && (e = m.row(i).at(k_), true); ++k_) {
        if (m.mcol[e] < 0) continue;
        rowSum += off[e];
        mag += fabs(off[e]);
    }
    if (!(isfinite(rowSum) && fabs(rowSum) <= 64 * 2.220446049250313e-16 * mag)) atomicOr(notConstant, 1);
}
// reference(global cell 0, value) (procedural_equation.cpp:306-315).
__global__ void kReferenceApply(double* diag, double* src, int cell, double value, const int* notConstant, int* err) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || notConstant[0]) return;
    if (!(diag[cell] > 0)) { atomicOr(err, kErrNonFinite); return; }
    src[cell] += diag[cell] * value;
    diag[cell] *= 2;
}

__global__ void kCorrection(int nf, const double* full, const double* orth, double* corr) {
    const int f = globalThread();
    if (f < nf) corr[f] = full[f] + (-1.0) * orth[f];
}

// equ::faceFlux (procedural_equation.cpp:335-352) with phi = phiH + pressureFlux (main.cpp:98).
__global__ void kPressureFluxPhi(MeshView m, FieldArg pp, const double* kf, const double* corr, const double* phiH,
                                 double* phi, int* err) {
    const int f = globalThread();
    if (f >= m.nf) return;
    const double k = kf[f];
    if (!(k >= 0.0) || !isfinite(k)) atomicOr(err, kErrNegativeDiffusivity);
    double g[1];
    integratedNormalGradient<1>(pp, nullptr, m, f, kOrthogonal, g, err);
    double inc = k * g[0];
    inc += 1.0 * corr[f];
    const double pf = 0.0 + -1.0 * inc;
    phi[f] = phiH[f] + pf;
}

}  // namespace
}  // namespace flow
}  // namespace bsim
