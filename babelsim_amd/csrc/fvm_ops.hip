// Row-gather HIP kernels for the face operators of
// src/discretization/operators.cpp.  One thread owns one cell (or one face)
// and reads the data of its faces; values that BabelSim scatters from a face
// to its owner and neighbour are recomputed on both sides instead.
#include "fvm_device.hpp"

namespace bsim {
namespace {

using namespace dev;

// ---------------------------------------------------------------- gradients
template <int NC, bool PASS2>
__global__ void __launch_bounds__(kBlock) kGradGG(MeshView m, FieldArg x, const double* g0, double* out, int* err) {
    const int c = globalThread();
    if (c >= m.nc) return;
    D3 acc[NC];
    for (int k = 0; k < NC; ++k) acc[k] = d3(0, 0, 0);
    const double iv = m.invV[c];
    double val[NC];
    for (int k_ = 0, e = 0; m.row(c).valid(k_) && (e = m.row(c).at(k_), true); ++k_) {
        if (m.mface[e] < 0) continue;
        const int f = m.mface[e];
        if (PASS2) correctedInternal<NC>(x, g0, m, f, val);
        else internalValue<NC>(x, m, f, val);
        const D3 sf = ld3(m.Sf, f);
        const bool owner = m.own[f] == c;
        for (int k = 0; k < NC; ++k) {
            const D3 contrib = val[k] * sf;
            acc[k] = owner ? acc[k] + contrib * iv : acc[k] - contrib * iv;
        }
    }
    for (int e = m.bptr[c]; e < m.bptr[c + 1]; ++e) {
        const int f = m.bface[e];
        if (PASS2) reconstructedBoundary<NC>(x, g0, m, f, false, 0.0, val, err);
        else traceValue<NC>(x, m, f, false, 0.0, val, err);
        const D3 sf = ld3(m.Sf, f);
        for (int k = 0; k < NC; ++k) acc[k] = acc[k] + (val[k] * sf) * iv;
    }
    for (int k = 0; k < NC; ++k) st3(out, static_cast<long long>(NC) * c + k, acc[k]);
}


template <int NC> __global__ void __launch_bounds__(kBlock) kGradLS(MeshView m, FieldArg x, double* out, int* err) {
    const int c = globalThread();
    if (c >= m.nc) return;
    double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double rhs[3 * NC];
    for (int i = 0; i < 3 * NC; ++i) rhs[i] = 0.0;
    const D3 cc = ld3(m.C, c);
    double xc[NC], val[NC];
    cellValue<NC>(x, c, xc);
    auto addRow = [&](D3 d, const double* diff) {
        const double d2 = dot(d, d);
        if (!(d2 > 0.0)) { atomicOr(err, kErrSingularLs); return; }
        const double wgt = 1.0 / d2;
        const D3 wd = wgt * d;
        const double wv[3] = {wd.x, wd.y, wd.z}, dv[3] = {d.x, d.y, d.z};
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) A[i][j] += wv[i] * dv[j];
        for (int i = 0; i < 3; ++i) for (int k = 0; k < NC; ++k) rhs[i * NC + k] += wv[i] * diff[k];
    };
    // Interior neighbours kC at a time: their columns, then their centres and
    // values, are loaded together; the rows are added in entry order.
    constexpr int kC = NC == 1 ? 8 : 4;
    const RowRange R = m.row(c);
    for (int k0 = 0; k0 < R.K && R.valid(k0); k0 += kC) {
        int other[kC];
#pragma unroll
        for (int j = 0; j < kC; ++j) {
            const int k = k0 + j;
            const int e = k < R.K && R.valid(k) ? R.at(k) : -1;
            other[j] = e >= 0 && m.mface[e] >= 0 ? m.mcol[e] : -1;
        }
        double nv[kC][NC];
        D3 nc_[kC];
#pragma unroll
        for (int j = 0; j < kC; ++j)
            if (other[j] >= 0) {
                cellValue<NC>(x, other[j], nv[j]);
                nc_[j] = ld3(m.C, other[j]);
            }
#pragma unroll
        for (int j = 0; j < kC; ++j)
            if (other[j] >= 0) {
                double diff[NC];
                for (int k = 0; k < NC; ++k) diff[k] = nv[j][k] - xc[k];
                addRow(nc_[j] - cc, diff);
            }
    }
    for (int e = m.bptr[c]; e < m.bptr[c + 1]; ++e) {
        const int f = m.bface[e];
        const int type = condition<NC>(x, m, f, false, 0.0, err).type;
        const D3 off = ld3(m.Cf, f) - cc;
        const D3 n = ld3(m.nrm, f);
        const D3 d = type == kFixedValue ? 2.0 * off : (2.0 * dot(off, n)) * n;
        traceValue<NC>(x, m, f, false, 0.0, val, err);
        double diff[NC];
        for (int k = 0; k < NC; ++k) diff[k] = 2.0 * (val[k] - xc[k]);
        addRow(d, diff);
    }
    double sol[3 * NC];
    if (!solveSym3<NC>(A, rhs, sol)) { atomicOr(err, kErrSingularLs); return; }
    for (int k = 0; k < NC; ++k) st3(out, static_cast<long long>(NC) * c + k, d3(sol[0 * NC + k], sol[1 * NC + k], sol[2 * NC + k]));
}

// -------------------------------------------------------- face quantities
template <int NC> __global__ void __launch_bounds__(kBlock) kInterp(MeshView m, FieldArg x, const double* g, double* out, int* err) {
    const int f = globalThread();
    if (f >= m.nf) return;
    double v[NC];
    if (f < m.nfi) {
        if (g) correctedInternal<NC>(x, g, m, f, v);
        else internalValue<NC>(x, m, f, v);
    } else {
        if (g) reconstructedBoundary<NC>(x, g, m, f, false, 0.0, v, err);
        else traceValue<NC>(x, m, f, false, 0.0, v, err);
    }
    for (int k = 0; k < NC; ++k) out[static_cast<long long>(NC) * f + k] = v[k];
}

__global__ void __launch_bounds__(kBlock) kSeedCtx(MeshView m, const double* U, double* ctx) {
    const int f = globalThread();
    if (f >= m.nf) return;
    ctx[f] = f < m.nfi ? 0.0 : dot(ld3(U, m.own[f]), ld3(m.Sf, f));
}

__global__ void __launch_bounds__(kBlock) kFlux(MeshView m, FieldArg U, const double* g, double* out, int* err) {
    const int f = globalThread();
    if (f >= m.nf) return;
    double v[3];
    if (f < m.nfi) {
        if (g) correctedInternal<3>(U, g, m, f, v);
        else internalValue<3>(U, m, f, v);
    } else {
        if (g) reconstructedBoundary<3>(U, g, m, f, false, 0.0, v, err);
        else traceValue<3>(U, m, f, false, 0.0, v, err);
    }
    out[f] = dot(d3(v[0], v[1], v[2]), ld3(m.Sf, f));
}

__global__ void __launch_bounds__(kBlock) kSnGrad(MeshView m, FieldArg x, const double* g, int dm, double* out, int* err) {
    const int f = globalThread();
    if (f >= m.nf) return;
    double v[1];
    integratedNormalGradient<1>(x, g, m, f, dm, v, err);
    out[f] = v[0] / m.Af[f];
}

__global__ void __launch_bounds__(kBlock) kDiffusionFlux(MeshView m, FieldArg x, const double* kf, const double* g, int dm, double* out,
                               int* err) {
    const int f = globalThread();
    if (f >= m.nf) return;
    const double k = kf[f];
    if (!(k >= 0.0) || !isfinite(k)) atomicOr(err, kErrNegativeDiffusivity);
    double v[1];
    integratedNormalGradient<1>(x, g, m, f, dm, v, err);
    out[f] = k * v[0];
}

__global__ void __launch_bounds__(kBlock) kDivFace(MeshView m, const double* phi, double* out) {
    const int c = globalThread();
    if (c >= m.nc) return;
    const double iv = m.invV[c];
    double acc = 0.0;
    for (int k_ = 0, e = 0; m.row(c).valid(k_) && (e = m.row(c).at(k_), true); ++k_) {
        if (m.mface[e] < 0) continue;
        const int f = m.mface[e];
        acc = m.own[f] == c ? acc + phi[f] * iv : acc - phi[f] * iv;
    }
    for (int e = m.bptr[c]; e < m.bptr[c + 1]; ++e) acc += phi[m.bface[e]] * iv;
    out[c] = acc;
}

__global__ void __launch_bounds__(kBlock) kExplicitLaplacian(MeshView m, FieldArg x, const double* kf, double kc, const double* g, int dm,
                                   double* out, int* err) {
    const int c = globalThread();
    if (c >= m.nc) return;
    const double iv = m.invV[c];
    double acc = 0.0;
    for (int k_ = 0, e = 0; m.row(c).valid(k_) && (e = m.row(c).at(k_), true); ++k_) {
        if (m.mface[e] < 0) continue;
        const int f = m.mface[e];
        double q[1];
        if (correctedDiffusion(dm)) integratedNormalGradient<1>(x, g, m, f, dm, q, err);
        else q[0] = m.orth[f] * (x.v[m.nei[f]] - x.v[m.own[f]]);
        const double integ = (kf ? kf[f] : kc) * q[0];
        acc = m.own[f] == c ? acc + integ * iv : acc - integ * iv;
    }
    for (int e = m.bptr[c]; e < m.bptr[c + 1]; ++e) {
        const int f = m.bface[e];
        const Cnd<1> cd = condition<1>(x, m, f, false, 0.0, err);
        double q = 0.0;
        if (cd.type == kFixedValue) {
            double tv[1];
            traceValue<1>(x, m, f, true, -1.0, tv, err);
            q = m.orth[f] * (tv[0] - x.v[c]);
            if (correctedDiffusion(dm)) { double r[1]; integratedNormalGradient<1>(x, g, m, f, dm, r, err); q = r[0]; }
        } else if (cd.type == kFixedGradient) {
            q = cd.v[0] * m.Af[f];
        }
        acc += ((kf ? kf[f] : kc) * q) * iv;
    }
    out[c] = acc;
}

template <int NC>
__global__ void __launch_bounds__(kBlock) kExplicitConvection(MeshView m, FieldArg x, const double* phi, int conv, int interp, const double* g,
                                    double* out, int* err) {
    const int c = globalThread();
    if (c >= m.nc) return;
    const bool haveG = conv == kLinearUpwind || (conv == kCentral && interp == kCorrected);
    const double iv = m.invV[c];
    double acc[NC], fv[NC], corr[NC];
    for (int k = 0; k < NC; ++k) acc[k] = 0.0;
    for (int k_ = 0, e = 0; m.row(c).valid(k_) && (e = m.row(c).at(k_), true); ++k_) {
        if (m.mface[e] < 0) continue;
        const int f = m.mface[e];
        const double q = phi[f];
        if (conv == kUpwind || conv == kLinearUpwind) {
            const int up = q >= 0.0 ? m.own[f] : m.nei[f];
            cellValue<NC>(x, up, fv);
            if (conv == kLinearUpwind) { luCorrection<NC>(g, m, up, f, corr); for (int k = 0; k < NC; ++k) fv[k] += corr[k]; }
        } else if (haveG) {
            correctedInternal<NC>(x, g, m, f, fv);
        } else {
            internalValue<NC>(x, m, f, fv);
        }
        const bool owner = m.own[f] == c;
        for (int k = 0; k < NC; ++k) {
            const double integ = q * fv[k];
            acc[k] = owner ? acc[k] + integ * iv : acc[k] - integ * iv;
        }
    }
    for (int e = m.bptr[c]; e < m.bptr[c + 1]; ++e) {
        const int f = m.bface[e];
        const double q = phi[f];
        if (conv == kUpwind || conv == kLinearUpwind) {
            if (q >= 0.0) {
                cellValue<NC>(x, c, fv);
                if (conv == kLinearUpwind) { luCorrection<NC>(g, m, c, f, corr); for (int k = 0; k < NC; ++k) fv[k] += corr[k]; }
            } else {
                traceValue<NC>(x, m, f, true, q, fv, err);
            }
        } else if (haveG) {
            reconstructedBoundary<NC>(x, g, m, f, true, q, fv, err);
        } else {
            traceValue<NC>(x, m, f, true, q, fv, err);
        }
        for (int k = 0; k < NC; ++k) acc[k] += (q * fv[k]) * iv;
    }
    for (int k = 0; k < NC; ++k) out[static_cast<long long>(NC) * c + k] = acc[k];
}

// ---------------------------------------------------------------- assembly
template <int NC>
__global__ void __launch_bounds__(kBlock) kAssembleConvection(MeshView m, FieldArg x, const double* phi, int conv, int interp, const double* g,
                                    double scale, double* diag, double* off, double* src, int* err) {
    const int c = globalThread();
    if (c >= m.nc) return;
    const bool haveG = conv == kLinearUpwind || (conv == kCentral && interp == kCorrected);
    double dAcc = 0.0, sAcc[NC], corr[NC], a[NC], b[NC];
    for (int k = 0; k < NC; ++k) sAcc[k] = 0.0;
    for (int k_ = 0, e = 0; m.row(c).valid(k_) && (e = m.row(c).at(k_), true); ++k_) {
        if (m.mface[e] < 0) continue;
        const int f = m.mface[e];
        const int o = m.own[f], n = m.nei[f];
        const bool owner = o == c;
        const double F = scale * phi[f];
        if (conv == kUpwind || conv == kLinearUpwind) {
            if (owner) { dAcc += fmax(F, 0.0); off[e] += fmin(F, 0.0); }
            else { dAcc += fmax(-F, 0.0); off[e] += fmin(-F, 0.0); }
            if (conv == kLinearUpwind) {
                luCorrection<NC>(g, m, F >= 0.0 ? o : n, f, corr);
                for (int k = 0; k < NC; ++k) sAcc[k] = owner ? sAcc[k] - F * corr[k] : sAcc[k] + F * corr[k];
            }
        } else {
            const double w = m.w[f];
            if (owner) { dAcc += F * w; off[e] += F * (1.0 - w); }
            else { off[e] -= F * w; dAcc -= F * (1.0 - w); }
            if (haveG) {
                correctedInternal<NC>(x, g, m, f, a);
                internalValue<NC>(x, m, f, b);
                for (int k = 0; k < NC; ++k) {
                    const double cr = a[k] - b[k];
                    sAcc[k] = owner ? sAcc[k] - F * cr : sAcc[k] + F * cr;
                }
            }
        }
    }
    for (int e = m.bptr[c]; e < m.bptr[c + 1]; ++e) {
        const int f = m.bface[e];
        const double F = scale * phi[f];
        const Cnd<NC> cd = condition<NC>(x, m, f, true, F, err);
        if (conv == kUpwind || conv == kLinearUpwind) {
            if (F >= 0.0) {
                dAcc += F;
                if (conv == kLinearUpwind) {
                    luCorrection<NC>(g, m, c, f, corr);
                    for (int k = 0; k < NC; ++k) sAcc[k] -= F * corr[k];
                }
            } else {
                traceValue<NC>(x, m, f, true, F, a, err);
                for (int k = 0; k < NC; ++k) sAcc[k] -= F * a[k];
            }
            continue;
        }
        if (cd.type == kFixedValue) {
            for (int k = 0; k < NC; ++k) sAcc[k] -= F * cd.v[k];
        } else if (cd.type == kFixedGradient) {
            const double dist = dot(ld3(m.Cf, f) - ld3(m.C, c), ld3(m.nrm, f));
            dAcc += F;
            for (int k = 0; k < NC; ++k) sAcc[k] -= (F * dist) * cd.v[k];
        } else {
            dAcc += F;
        }
        if (haveG) {
            reconstructedBoundary<NC>(x, g, m, f, true, F, a, err);
            traceValue<NC>(x, m, f, true, F, b, err);
            for (int k = 0; k < NC; ++k) sAcc[k] -= F * (a[k] - b[k]);
        }
    }
    diag[c] += dAcc;
    for (int k = 0; k < NC; ++k) src[static_cast<long long>(NC) * c + k] += sAcc[k];
}

template <int NC>
__global__ void __launch_bounds__(kBlock) kAssembleDiffusion(MeshView m, FieldArg x, const double* kf, const double* g, int dm, double factor,
                                   double* diag, double* off, double* src, int* err) {
    const int c = globalThread();
    if (c >= m.nc) return;
    const bool corrected = correctedDiffusion(dm);
    double dAcc = 0.0, sAcc[NC], ing[NC], bv[NC];
    for (int k = 0; k < NC; ++k) sAcc[k] = 0.0;
    for (int k_ = 0, e = 0; m.row(c).valid(k_) && (e = m.row(c).at(k_), true); ++k_) {
        if (m.mface[e] < 0) continue;
        const int f = m.mface[e];
        const double gamma = kf[f];
        if (gamma < 0.0 || !isfinite(gamma)) { atomicOr(err, kErrNegativeDiffusivity); continue; }
        if (gamma == 0.0) continue;
        const double coef = gamma * m.orth[f];
        dAcc += coef;
        off[e] += factor * (0.0 - coef);
        if (corrected) {
            const long long o = m.own[f], n = m.nei[f];
            integratedNormalGradient<NC>(x, g, m, f, dm, ing, err);
            const bool owner = o == c;
            for (int k = 0; k < NC; ++k) {
                const double orthogonal = m.orth[f] * (x.v[NC * n + k] - x.v[NC * o + k]);
                const double corr = gamma * (ing[k] - orthogonal);
                sAcc[k] = owner ? sAcc[k] + corr : sAcc[k] - corr;
            }
        }
    }
    for (int e = m.bptr[c]; e < m.bptr[c + 1]; ++e) {
        const int f = m.bface[e];
        const double gamma = kf[f];
        if (gamma < 0.0 || !isfinite(gamma)) { atomicOr(err, kErrNegativeDiffusivity); continue; }
        if (gamma == 0.0) continue;
        const double coef = gamma * m.orth[f];
        const Cnd<NC> cd = condition<NC>(x, m, f, false, 0.0, err);
        if (cd.type == kFixedValue || (NC == 3 && cd.type == kSymmetry)) {
            if (NC == 3) traceValue<NC>(x, m, f, true, -1.0, bv, err);
            else for (int k = 0; k < NC; ++k) bv[k] = cd.v[k];
            dAcc += coef;
            for (int k = 0; k < NC; ++k) sAcc[k] += coef * bv[k];
            if (corrected) {
                integratedNormalGradient<NC>(x, g, m, f, dm, ing, err);
                for (int k = 0; k < NC; ++k) {
                    const double orthogonal = m.orth[f] * (bv[k] - x.v[static_cast<long long>(NC) * c + k]);
                    sAcc[k] += gamma * (ing[k] - orthogonal);
                }
            }
        } else if (cd.type == kFixedGradient) {
            for (int k = 0; k < NC; ++k) sAcc[k] += NC == 1 ? (gamma * cd.v[k]) * m.Af[f] : (gamma * m.Af[f]) * cd.v[k];
        }
    }
    diag[c] += factor * dAcc;
    for (int k = 0; k < NC; ++k) src[static_cast<long long>(NC) * c + k] += factor * sAcc[k];
}

// ------------------------------------------------ face-ordered assembly
// The parts of an interior face's contribution that need the gradients (the
// linearUpwind or corrected-interpolation convection flux and the
// non-orthogonal diffusion correction) are formed once per face here, in face
// order, instead of once from each of the face's two cells; the cell pass
// (kAssembleTransport) then gathers them.  Each stored value is the product
// the cell kernels form inline (contraction is off), so the matrices are
// those of kAssembleConvection then kAssembleDiffusion bit for bit.
//   qc[NC*f + k] = F * correction   (F = scale * phi[f]);
//   qd[NC*f + k] = gamma * (ing - orth * (x_n - x_o)).
template <int NC>
__global__ void __launch_bounds__(kBlock) kFaceCorrections(MeshView m, FieldArg x, const double* phi, int conv, int interp,
                                                           const double* g, double scale, const double* kf, int dm,
                                                           double* qc, double* qd) {
    const int f = globalThread();
    if (f >= m.nfi) return;
    const long long o = m.own[f], n = m.nei[f];
    if (qc) {
        const double F = scale * phi[f];
        double a[NC], b[NC];
        if (conv == kLinearUpwind) {
            luCorrection<NC>(g, m, F >= 0.0 ? static_cast<int>(o) : static_cast<int>(n), f, a);
            for (int k = 0; k < NC; ++k) qc[static_cast<long long>(NC) * f + k] = F * a[k];
        } else {
            correctedInternal<NC>(x, g, m, f, a);
            internalValue<NC>(x, m, f, b);
            for (int k = 0; k < NC; ++k) {
                const double cr = a[k] - b[k];
                qc[static_cast<long long>(NC) * f + k] = F * cr;
            }
        }
    }
    if (qd) {
        const double gamma = kf[f];
        double ing[NC];
        int dummy = 0;
        if (gamma > 0.0 && isfinite(gamma)) {
            integratedNormalGradient<NC>(x, g, m, f, dm, ing, &dummy);
            for (int k = 0; k < NC; ++k) {
                const double orthogonal = m.orth[f] * (x.v[NC * n + k] - x.v[NC * o + k]);
                qd[static_cast<long long>(NC) * f + k] = gamma * (ing[k] - orthogonal);
            }
        }
    }
}

// kAssembleConvection (when conv >= 0) followed by kAssembleDiffusion (when
// kf is given) in one cell pass: each cell keeps the two kernels' separate
// accumulators and adds them to diag, off and src in their order.
template <int NC>
__global__ void __launch_bounds__(kBlock) kAssembleTransport(MeshView m, FieldArg x, const double* phi, int conv, int interp,
                                                             const double* g, double scale, const double* kf, int dm,
                                                             double factor, const double* qc, const double* qd,
                                                             double* diag, double* off, double* src, int* err) {
    const int c = globalThread();
    if (c >= m.nc) return;
    const bool doConv = conv >= 0, doDiff = kf != nullptr;
    const bool haveG = conv == kLinearUpwind || (conv == kCentral && interp == kCorrected);
    const bool corrected = doDiff && correctedDiffusion(dm);
    double dC = 0.0, dD = 0.0, sC[NC], sD[NC], a[NC], b[NC], corr[NC];
    for (int k = 0; k < NC; ++k) { sC[k] = 0.0; sD[k] = 0.0; }
    for (int k_ = 0, e = 0; m.row(c).valid(k_) && (e = m.row(c).at(k_), true); ++k_) {
        if (m.mface[e] < 0) continue;
        const int f = m.mface[e];
        const bool owner = m.own[f] == c;
        double oe = off[e];
        if (doConv) {
            const double F = scale * phi[f];
            if (conv == kUpwind || conv == kLinearUpwind) {
                if (owner) { dC += fmax(F, 0.0); oe += fmin(F, 0.0); }
                else { dC += fmax(-F, 0.0); oe += fmin(-F, 0.0); }
            } else {
                const double w = m.w[f];
                if (owner) { dC += F * w; oe += F * (1.0 - w); }
                else { oe -= F * w; dC -= F * (1.0 - w); }
            }
            if (haveG)
                for (int k = 0; k < NC; ++k) {
                    const double q = qc[static_cast<long long>(NC) * f + k];
                    sC[k] = owner ? sC[k] - q : sC[k] + q;
                }
        }
        if (doDiff) {
            const double gamma = kf[f];
            if (gamma < 0.0 || !isfinite(gamma)) atomicOr(err, kErrNegativeDiffusivity);
            else if (gamma != 0.0) {
                const double coef = gamma * m.orth[f];
                dD += coef;
                oe += factor * (0.0 - coef);
                if (corrected)
                    for (int k = 0; k < NC; ++k) {
                        const double q = qd[static_cast<long long>(NC) * f + k];
                        sD[k] = owner ? sD[k] + q : sD[k] - q;
                    }
            }
        }
        off[e] = oe;
    }
    if (doConv)
        for (int e = m.bptr[c]; e < m.bptr[c + 1]; ++e) {
            const int f = m.bface[e];
            const double F = scale * phi[f];
            const Cnd<NC> cd = condition<NC>(x, m, f, true, F, err);
            if (conv == kUpwind || conv == kLinearUpwind) {
                if (F >= 0.0) {
                    dC += F;
                    if (conv == kLinearUpwind) {
                        luCorrection<NC>(g, m, c, f, corr);
                        for (int k = 0; k < NC; ++k) sC[k] -= F * corr[k];
                    }
                } else {
                    traceValue<NC>(x, m, f, true, F, a, err);
                    for (int k = 0; k < NC; ++k) sC[k] -= F * a[k];
                }
                continue;
            }
            if (cd.type == kFixedValue) {
                for (int k = 0; k < NC; ++k) sC[k] -= F * cd.v[k];
            } else if (cd.type == kFixedGradient) {
                const double dist = dot(ld3(m.Cf, f) - ld3(m.C, c), ld3(m.nrm, f));
                dC += F;
                for (int k = 0; k < NC; ++k) sC[k] -= (F * dist) * cd.v[k];
            } else {
                dC += F;
            }
            if (haveG) {
                reconstructedBoundary<NC>(x, g, m, f, true, F, a, err);
                traceValue<NC>(x, m, f, true, F, b, err);
                for (int k = 0; k < NC; ++k) sC[k] -= F * (a[k] - b[k]);
            }
        }
    if (doDiff)
        for (int e = m.bptr[c]; e < m.bptr[c + 1]; ++e) {
            const int f = m.bface[e];
            const double gamma = kf[f];
            if (gamma < 0.0 || !isfinite(gamma)) { atomicOr(err, kErrNegativeDiffusivity); continue; }
            if (gamma == 0.0) continue;
            const double coef = gamma * m.orth[f];
            const Cnd<NC> cd = condition<NC>(x, m, f, false, 0.0, err);
            if (cd.type == kFixedValue || (NC == 3 && cd.type == kSymmetry)) {
                if (NC == 3) traceValue<NC>(x, m, f, true, -1.0, a, err);
                else for (int k = 0; k < NC; ++k) a[k] = cd.v[k];
                dD += coef;
                for (int k = 0; k < NC; ++k) sD[k] += coef * a[k];
                if (corrected) {
                    integratedNormalGradient<NC>(x, g, m, f, dm, b, err);
                    for (int k = 0; k < NC; ++k) {
                        const double orthogonal = m.orth[f] * (a[k] - x.v[static_cast<long long>(NC) * c + k]);
                        sD[k] += gamma * (b[k] - orthogonal);
                    }
                }
            } else if (cd.type == kFixedGradient) {
                for (int k = 0; k < NC; ++k) sD[k] += NC == 1 ? (gamma * cd.v[k]) * m.Af[f] : (gamma * m.Af[f]) * cd.v[k];
            }
        }
    double d = diag[c];
    if (doConv) d += dC;
    if (doDiff) d += factor * dD;
    diag[c] = d;
    for (int k = 0; k < NC; ++k) {
        double v = src[static_cast<long long>(NC) * c + k];
        if (doConv) v += sC[k];
        if (doDiff) v += factor * sD[k];
        src[static_cast<long long>(NC) * c + k] = v;
    }
}

// diffusionFlux with the method and orthogonal, then kCorrection, for every
// face (procedural_equation.cpp:77-87): corr = k ing - k ing_orth, the
// products and the difference formed as those kernels form them.  Interior
// faces also store the assembly's correction qd (kFaceCorrections).
__global__ void __launch_bounds__(kBlock) kDiffusionFaces(MeshView m, FieldArg x, const double* kf, const double* g, int dm,
                                                          double* corr, double* qd, int* err) {
    const int f = globalThread();
    if (f >= m.nf) return;
    const double k = kf[f];
    const bool valid = k >= 0.0 && isfinite(k);
    if (!valid) atomicOr(err, kErrNegativeDiffusivity);
    double v[1], v0[1];
    integratedNormalGradient<1>(x, g, m, f, dm, v, err);
    integratedNormalGradient<1>(x, g, m, f, kOrthogonal, v0, err);
    const double full = k * v[0], sn = k * v0[0];
    corr[f] = full + (-1.0) * sn;
    if (f < m.nfi && valid && k != 0.0) qd[f] = k * (v[0] - v0[0]);
}

__global__ void __launch_bounds__(kBlock) kGradTraces(MeshView m, FieldArg x, const double* g, double* traces) {
    const int b = globalThread();
    const int f = m.nfi + b;
    if (f >= m.nf) return;
    int dummy = 0;
    const int o = m.own[f];
    const D3 n = ld3(m.nrm, f);
    const D3 d = ld3(m.Cf, f) - ld3(m.C, o);
    D3 gv = ld3(g, o);
    const Cnd<1> bc = condition<1>(x, m, f, false, 0.0, &dummy);
    const double sn = bc.type == kFixedValue ? (bc.v[0] - x.v[o] - dot(gv, d - dot(d, n) * n)) / dot(d, n)
                    : bc.type == kFixedGradient ? bc.v[0] : 0.0;
    gv = gv + (sn - dot(gv, n)) * n;
    st3(traces, f, gv);
}

}  // namespace

namespace ops {

void gradient(const MeshView& m, bool orthogonal, int ncomp, FieldArg x, int method, double* out, double* work,
              int* err, hipStream_t s) {
    const int nb = blocksFor(m.nc);
    if (method == kLeastSquares) {
        if (ncomp == 1) kGradLS<1><<<nb, kBlock, 0, s>>>(m, x, out, err);
        else kGradLS<3><<<nb, kBlock, 0, s>>>(m, x, out, err);
    } else if (method == kGreenGauss) {
        double* first = orthogonal ? out : work;
        if (ncomp == 1) {
            kGradGG<1, false><<<nb, kBlock, 0, s>>>(m, x, nullptr, first, err);
            if (!orthogonal) kGradGG<1, true><<<nb, kBlock, 0, s>>>(m, x, first, out, err);
        } else {
            kGradGG<3, false><<<nb, kBlock, 0, s>>>(m, x, nullptr, first, err);
            if (!orthogonal) kGradGG<3, true><<<nb, kBlock, 0, s>>>(m, x, first, out, err);
        }
    } else {
        throw std::invalid_argument("unsupported gradient method");
    }
    BSIM_LAUNCH_CHECK();
}

void interpolate(const MeshView& m, int ncomp, FieldArg x, int method, const double* grad, double* out, int* err,
                 hipStream_t s) {
    if (method != kLinear && method != kCorrected) throw std::invalid_argument("unsupported interpolation method");
    const double* g = method == kCorrected ? grad : nullptr;
    if (ncomp == 1) kInterp<1><<<blocksFor(m.nf), kBlock, 0, s>>>(m, x, g, out, err);
    else kInterp<3><<<blocksFor(m.nf), kBlock, 0, s>>>(m, x, g, out, err);
    BSIM_LAUNCH_CHECK();
}

void seedFluxContext(const MeshView& m, const double* U, double* ctx, hipStream_t s) {
    kSeedCtx<<<blocksFor(m.nf), kBlock, 0, s>>>(m, U, ctx);
    BSIM_LAUNCH_CHECK();
}

void flux(const MeshView& m, FieldArg U, const double* gradU, double* out, int* err, hipStream_t s) {
    kFlux<<<blocksFor(m.nf), kBlock, 0, s>>>(m, U, gradU, out, err);
    BSIM_LAUNCH_CHECK();
}

void normalGradient(const MeshView& m, FieldArg x, const double* grad, int dmethod, double* out, int* err,
                    hipStream_t s) {
    kSnGrad<<<blocksFor(m.nf), kBlock, 0, s>>>(m, x, grad, dmethod, out, err);
    BSIM_LAUNCH_CHECK();
}

void diffusionFlux(const MeshView& m, FieldArg x, const double* kf, const double* grad, int dmethod, double* out,
                   int* err, hipStream_t s) {
    kDiffusionFlux<<<blocksFor(m.nf), kBlock, 0, s>>>(m, x, kf, grad, dmethod, out, err);
    BSIM_LAUNCH_CHECK();
}

void divergenceFace(const MeshView& m, const double* phi, double* out, hipStream_t s) {
    kDivFace<<<blocksFor(m.nc), kBlock, 0, s>>>(m, phi, out);
    BSIM_LAUNCH_CHECK();
}

void explicitLaplacian(const MeshView& m, FieldArg x, const double* kf, double k, const double* grad, int dmethod,
                       double* out, int* err, hipStream_t s) {
    kExplicitLaplacian<<<blocksFor(m.nc), kBlock, 0, s>>>(m, x, kf, k, grad, dmethod, out, err);
    BSIM_LAUNCH_CHECK();
}

void explicitConvection(const MeshView& m, int ncomp, FieldArg x, const double* phi, int conv, int interp,
                        const double* grad, double* out, int* err, hipStream_t s) {
    if (ncomp == 1) kExplicitConvection<1><<<blocksFor(m.nc), kBlock, 0, s>>>(m, x, phi, conv, interp, grad, out, err);
    else kExplicitConvection<3><<<blocksFor(m.nc), kBlock, 0, s>>>(m, x, phi, conv, interp, grad, out, err);
    BSIM_LAUNCH_CHECK();
}

void assembleConvection(const MeshView& m, int ncomp, FieldArg x, const double* phi, int conv, int interp,
                        const double* grad, double scale, double* diag, double* off, double* src, int* err,
                        hipStream_t s) {
    if (conv != kUpwind && conv != kLinearUpwind && conv != kCentral) throw std::invalid_argument("unsupported convection method");
    if (interp != kLinear && interp != kCorrected) throw std::invalid_argument("unsupported convection interpolation method");
    if (!std::isfinite(scale)) throw std::invalid_argument("convection flux scale must be finite");
    if (ncomp == 1) kAssembleConvection<1><<<blocksFor(m.nc), kBlock, 0, s>>>(m, x, phi, conv, interp, grad, scale, diag, off, src, err);
    else kAssembleConvection<3><<<blocksFor(m.nc), kBlock, 0, s>>>(m, x, phi, conv, interp, grad, scale, diag, off, src, err);
    BSIM_LAUNCH_CHECK();
}

void assembleDiffusion(const MeshView& m, int ncomp, FieldArg x, const double* kf, const double* grad, int dmethod,
                       double factor, double* diag, double* off, double* src, int* err, hipStream_t s) {
    if (dmethod != kOrthogonal && dmethod != kDiffCorrected && dmethod != kLimitedCorrected)
        throw std::invalid_argument("unsupported diffusion method");
    if (ncomp == 1) kAssembleDiffusion<1><<<blocksFor(m.nc), kBlock, 0, s>>>(m, x, kf, grad, dmethod, factor, diag, off, src, err);
    else kAssembleDiffusion<3><<<blocksFor(m.nc), kBlock, 0, s>>>(m, x, kf, grad, dmethod, factor, diag, off, src, err);
    BSIM_LAUNCH_CHECK();
}

void assembleTransport(const MeshView& m, int ncomp, FieldArg x, const double* phi, int conv, int interp, double scale,
                       const double* kf, int dmethod, double factor, const double* grad, double* diag, double* off,
                       double* src, double* faceWork, int* err, hipStream_t s) {
    if (phi) {
        if (conv != kUpwind && conv != kLinearUpwind && conv != kCentral) throw std::invalid_argument("unsupported convection method");
        if (interp != kLinear && interp != kCorrected) throw std::invalid_argument("unsupported convection interpolation method");
        if (!std::isfinite(scale)) throw std::invalid_argument("convection flux scale must be finite");
    }
    if (kf && dmethod != kOrthogonal && dmethod != kDiffCorrected && dmethod != kLimitedCorrected)
        throw std::invalid_argument("unsupported diffusion method");
    const int cv = phi ? conv : -1;
    const bool needC = phi && (conv == kLinearUpwind || (conv == kCentral && interp == kCorrected));
    const bool needD = kf && (dmethod == kDiffCorrected || dmethod == kLimitedCorrected);
    const long long span = static_cast<long long>(ncomp) * m.nfi;
    double* qc = needC ? faceWork : nullptr;
    double* qd = needD ? faceWork + (needC ? span : 0) : nullptr;
    if ((needC || needD) && m.nfi > 0) {
        if (ncomp == 1) kFaceCorrections<1><<<blocksFor(m.nfi), kBlock, 0, s>>>(m, x, phi, conv, interp, grad, scale, kf, dmethod, qc, qd);
        else kFaceCorrections<3><<<blocksFor(m.nfi), kBlock, 0, s>>>(m, x, phi, conv, interp, grad, scale, kf, dmethod, qc, qd);
    }
    if (ncomp == 1)
        kAssembleTransport<1><<<blocksFor(m.nc), kBlock, 0, s>>>(m, x, phi, cv, interp, grad, scale, kf, dmethod, factor, qc,
                                                                 qd, diag, off, src, err);
    else
        kAssembleTransport<3><<<blocksFor(m.nc), kBlock, 0, s>>>(m, x, phi, cv, interp, grad, scale, kf, dmethod, factor, qc,
                                                                 qd, diag, off, src, err);
    BSIM_LAUNCH_CHECK();
}

void assembleDiffusionCorrected(const MeshView& m, FieldArg x, const double* kf, const double* grad, int dmethod,
                                double factor, double* diag, double* off, double* src, double* corr, double* faceWork,
                                int* err, hipStream_t s) {
    if (dmethod != kDiffCorrected && dmethod != kLimitedCorrected)
        throw std::invalid_argument("the corrected assembly needs a corrected diffusion method");
    kDiffusionFaces<<<blocksFor(m.nf), kBlock, 0, s>>>(m, x, kf, grad, dmethod, corr, faceWork, err);
    kAssembleTransport<1><<<blocksFor(m.nc), kBlock, 0, s>>>(m, x, nullptr, -1, kLinear, grad, 0.0, kf, dmethod, factor,
                                                             nullptr, faceWork, diag, off, src, err);
    BSIM_LAUNCH_CHECK();
}

void gradientTraces(const MeshView& m, FieldArg x, const double* grad, double* traces, hipStream_t s) {
    if (m.nf > m.nfi) kGradTraces<<<blocksFor(m.nf - m.nfi), kBlock, 0, s>>>(m, x, grad, traces);
    BSIM_LAUNCH_CHECK();
}

}  // namespace ops
}  // namespace bsim
