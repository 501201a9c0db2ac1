// Device-side face rules shared by the operator kernels and the fused SIMPLE
// kernels: BabelSim's boundary resolution (include/babelsim/field.h:285-313),
// face interpolation and reconstruction (src/discretization/operators.cpp:
// 35-188) and the integrated normal gradient (operators.cpp:857-954).
#pragma once

#include "fvm_ops.hpp"

namespace bsim {
namespace dev {

template <int NC> struct Cnd {
    int type;
    double v[NC];
};

// Face-condition resolution, field.h:285-301 (inletOutlet flips on flux sign).
template <int NC>
__device__ inline Cnd<NC> condition(const FieldArg& x, const MeshView& m, int f, bool hasFlux, double flux, int* err) {
    Cnd<NC> c;
    if (x.trace) {
        c.type = kFixedValue;
        for (int k = 0; k < NC; ++k) c.v[k] = x.trace[static_cast<long long>(NC) * f + k];
        return c;
    }
    const int p = m.patch[f];
    c.type = x.bcType[p];
    for (int k = 0; k < NC; ++k) c.v[k] = x.bcVal[NC * p + k];
    if (c.type == kInletOutlet) {
        if (!hasFlux) {
            if (!x.ctx) { atomicOr(err, kErrNoFluxContext); flux = 0.0; }
            else flux = x.ctx[f];
        }
        if (flux < 0.0) c.type = kFixedValue;
        else { c.type = kZeroGradient; for (int k = 0; k < NC; ++k) c.v[k] = 0.0; }
    }
    return c;
}

template <int NC> __device__ inline void cellValue(const FieldArg& x, int c, double* out) {
    for (int k = 0; k < NC; ++k) out[k] = x.v[static_cast<long long>(NC) * c + k];
}

// Boundary trace, field.h:302-313.
template <int NC>
__device__ inline void traceValue(const FieldArg& x, const MeshView& m, int f, bool hasFlux, double flux, double* out,
                                  int* err) {
    const Cnd<NC> c = condition<NC>(x, m, f, hasFlux, flux, err);
    const int o = m.own[f];
    if (c.type == kFixedValue) { for (int k = 0; k < NC; ++k) out[k] = c.v[k]; return; }
    cellValue<NC>(x, o, out);
    const D3 n = ld3(m.nrm, f);
    if (c.type == kFixedGradient) {
        const double d = dot(ld3(m.Cf, f) - ld3(m.C, o), n);
        for (int k = 0; k < NC; ++k) out[k] = out[k] + d * c.v[k];
    } else if (c.type == kSymmetry && NC == 3) {
        const D3 u = d3(out[0], out[NC > 1 ? 1 : 0], out[NC > 2 ? 2 : 0]);
        const D3 r = u - dot(u, n) * n;
        out[0] = r.x;
        if (NC == 3) { out[NC > 1 ? 1 : 0] = r.y; out[NC > 2 ? 2 : 0] = r.z; }
    }
}

template <int NC> __device__ inline D3 grow(const double* g, int c, int comp) {
    return ld3(g, static_cast<long long>(NC) * c + comp);
}

template <int NC> __device__ inline void internalValue(const FieldArg& x, const MeshView& m, int f, double* out) {
    const double w = m.w[f];
    const long long o = m.own[f], n = m.nei[f];
    for (int k = 0; k < NC; ++k) out[k] = w * x.v[NC * o + k] + (1.0 - w) * x.v[NC * n + k];
}

template <int NC>
__device__ inline void correctedInternal(const FieldArg& x, const double* g, const MeshView& m, int f, double* out) {
    internalValue<NC>(x, m, f, out);
    const double w = m.w[f];
    const int o = m.own[f], n = m.nei[f];
    const D3 sk = ld3(m.skew, f);
    for (int k = 0; k < NC; ++k) {
        const D3 fg = w * grow<NC>(g, o, k) + (1.0 - w) * grow<NC>(g, n, k);
        out[k] += dot(sk, fg);
    }
}

// Constrained-gradient extrapolation, operators.cpp:111-161.
template <int NC>
__device__ inline void reconstructedBoundary(const FieldArg& x, const double* g, const MeshView& m, int f,
                                             bool hasFlux, double flux, double* out, int* err) {
    const Cnd<NC> c = condition<NC>(x, m, f, hasFlux, flux, err);
    if (c.type == kFixedValue) { for (int k = 0; k < NC; ++k) out[k] = c.v[k]; return; }
    const int o = m.own[f];
    const D3 n = ld3(m.nrm, f);
    const D3 off = ld3(m.Cf, f) - ld3(m.C, o);
    cellValue<NC>(x, o, out);
    for (int k = 0; k < NC; ++k) {
        const double pn = c.type == kFixedGradient ? c.v[k] : 0.0;
        const D3 go = grow<NC>(g, o, k);
        const D3 cg = go + (pn - dot(go, n)) * n;
        out[k] += dot(cg, off);
    }
    if (NC == 3 && c.type == kSymmetry) {
        const D3 u = d3(out[0], out[NC > 1 ? 1 : 0], out[NC > 2 ? 2 : 0]);
        const D3 r = u - dot(u, n) * n;
        out[0] = r.x;
        out[NC > 1 ? 1 : 0] = r.y;
        out[NC > 2 ? 2 : 0] = r.z;
    }
}

template <int NC>
__device__ inline void luCorrection(const double* g, const MeshView& m, int cell, int f, double* out) {
    const D3 off = ld3(m.Cf, f) - ld3(m.C, cell);
    for (int k = 0; k < NC; ++k) out[k] = dot(grow<NC>(g, cell, k), off);
}

__device__ inline double limitCorrection(double orthogonal, double correction, int method) {
    if (method == kDiffCorrected) return correction;
    if (method != kLimitedCorrected) return 0.0;
    const double b = fabs(orthogonal);
    return fmin(fmax(correction, -b), b);
}
__device__ inline bool correctedDiffusion(int d) { return d == kDiffCorrected || d == kLimitedCorrected; }

// operators.cpp:857-954
template <int NC>
__device__ inline void integratedNormalGradient(const FieldArg& x, const double* g, const MeshView& m, int f, int dm,
                                                double* out, int* err) {
    const int o = m.own[f], n = m.nei[f];
    const double oc = m.orth[f];
    const D3 kf = ld3(m.kf, f);
    if (n >= 0) {
        const double w = m.w[f];
        for (int k = 0; k < NC; ++k) {
            const double orthogonal = oc * (x.v[static_cast<long long>(NC) * n + k] - x.v[static_cast<long long>(NC) * o + k]);
            double r = orthogonal;
            if (correctedDiffusion(dm)) {
                const D3 fg = w * grow<NC>(g, o, k) + (1.0 - w) * grow<NC>(g, n, k);
                r = orthogonal + limitCorrection(orthogonal, dot(kf, fg), dm);
            }
            out[k] = r;
        }
        return;
    }
    const Cnd<NC> c = condition<NC>(x, m, f, false, 0.0, err);
    if (c.type == kFixedGradient) {
        for (int k = 0; k < NC; ++k) out[k] = NC == 1 ? c.v[k] * m.Af[f] : m.Af[f] * c.v[k];
        return;
    }
    if (c.type != kFixedValue && !(NC == 3 && c.type == kSymmetry)) {
        for (int k = 0; k < NC; ++k) out[k] = 0.0;
        return;
    }
    double bv[NC];
    if (NC == 3) traceValue<NC>(x, m, f, false, 0.0, bv, err);
    else for (int k = 0; k < NC; ++k) bv[k] = c.v[k];
    const D3 nr = ld3(m.nrm, f);
    for (int k = 0; k < NC; ++k) {
        const double orthogonal = oc * (bv[k] - x.v[static_cast<long long>(NC) * o + k]);
        double r = orthogonal;
        if (correctedDiffusion(dm)) {
            D3 bg = grow<NC>(g, o, k);
            if (c.type == kSymmetry) bg = bg - dot(bg, nr) * nr;
            r = orthogonal + limitCorrection(orthogonal, dot(kf, bg), dm);
        }
        out[k] = r;
    }
}

// Symmetric pivoted LDL^T solve of a 3x3 system with NR right-hand sides
// (b and x row-major 3 x NR).  Pivots swap rows/columns through constant-index
// branches and permute the right-hand side as they happen, so every array
// stays in registers; the arithmetic is the textbook loop order.
template <int A_, int B_, int K_, int NR>
__device__ inline void symSwap(double (&M)[3][3], double (&L)[3][3], double (&bp)[3][NR]) {
    for (int j = 0; j < 3; ++j) { const double q = M[A_][j]; M[A_][j] = M[B_][j]; M[B_][j] = q; }
    for (int i = 0; i < 3; ++i) { const double q = M[i][A_]; M[i][A_] = M[i][B_]; M[i][B_] = q; }
    for (int j = 0; j < K_; ++j) { const double q = L[A_][j]; L[A_][j] = L[B_][j]; L[B_][j] = q; }
    for (int r = 0; r < NR; ++r) { const double q = bp[A_][r]; bp[A_][r] = bp[B_][r]; bp[B_][r] = q; }
}
template <int NR> __device__ inline bool solveSym3(const double (&A)[3][3], const double* b, double* x) {
    double M[3][3], L[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, D[3], bp[3][NR];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) M[i][j] = A[i][j];
    for (int i = 0; i < 3; ++i) for (int r = 0; r < NR; ++r) bp[i][r] = b[i * NR + r];
    auto ok = [](double d) { return fabs(d) > 0.0 && isfinite(d); };
    // step 0: largest diagonal of rows 0..2
    int p0 = 0;
    double best = fabs(M[0][0]);
    if (fabs(M[1][1]) > best) { p0 = 1; best = fabs(M[1][1]); }
    if (fabs(M[2][2]) > best) p0 = 2;
    if (p0 == 1) symSwap<0, 1, 0, NR>(M, L, bp);
    else if (p0 == 2) symSwap<0, 2, 0, NR>(M, L, bp);
    D[0] = M[0][0];
    if (!ok(D[0])) return false;
    L[1][0] = M[1][0] / D[0];
    M[1][1] -= L[1][0] * M[0][1];
    M[1][2] -= L[1][0] * M[0][2];
    L[2][0] = M[2][0] / D[0];
    M[2][1] -= L[2][0] * M[0][1];
    M[2][2] -= L[2][0] * M[0][2];
    // step 1: rows 1..2
    const int p1 = fabs(M[2][2]) > fabs(M[1][1]) ? 2 : 1;
    if (p1 == 2) symSwap<1, 2, 1, NR>(M, L, bp);
    D[1] = M[1][1];
    if (!ok(D[1])) return false;
    L[2][1] = M[2][1] / D[1];
    M[2][2] -= L[2][1] * M[1][2];
    D[2] = M[2][2];
    if (!ok(D[2])) return false;
    for (int r = 0; r < NR; ++r) {
        double y0 = bp[0][r], y1 = bp[1][r], y2 = bp[2][r];
        y1 -= L[1][0] * y0;
        y2 -= L[2][0] * y0;
        y2 -= L[2][1] * y1;
        y0 /= D[0];
        y1 /= D[1];
        y2 /= D[2];
        y1 -= L[2][1] * y2;
        y0 -= L[1][0] * y1;
        y0 -= L[2][0] * y2;
        // undo the pivots in reverse order
        if (p1 == 2) { const double q = y1; y1 = y2; y2 = q; }
        if (p0 == 1) { const double q = y0; y0 = y1; y1 = q; }
        else if (p0 == 2) { const double q = y0; y0 = y2; y2 = q; }
        x[0 * NR + r] = y0;
        x[1 * NR + r] = y1;
        x[2 * NR + r] = y2;
    }
    for (int i = 0; i < 3 * NR; ++i) if (!isfinite(x[i])) return false;
    return true;
}

}  // namespace dev
}  // namespace bsim
