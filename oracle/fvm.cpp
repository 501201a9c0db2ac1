// Oracle face operators: restates src/discretization/operators.cpp and the
// boundary rules of include/babelsim/field.h.  Test infrastructure only.
#include "internal.hpp"

#include <algorithm>

namespace orc {
namespace {

inline double symmetric(double v, const V3&) { return v; }                 // vector.h:103
inline V3 symmetric(const V3& v, const V3& n) { return v - dot(v, n) * n; } // vector.h:104-106

inline double boundaryNormalDistance(const Mesh& m, Index f) {  // boundary_evaluation.h:7-13
    return dot(m.Cf[f] - m.C[m.own[f]], m.nrm[f]);
}

template <class T> T internalFaceValue(const Field<T>& x, Index f) {  // operators.cpp:35-44
    const Mesh& m = *x.m;
    const double w = m.w[f];
    return w * x.v[m.own[f]] + (1.0 - w) * x.v[m.nei[f]];
}

template <class T> T interpolatedFaceValue(const Field<T>& x, Index f,
                                           double flux = std::numeric_limits<double>::quiet_NaN()) {
    // operators.cpp:242-248
    return x.m->boundary(f) ? boundaryTrace(x, f, flux) : internalFaceValue(x, f);
}

double correctedInternal(const SField& x, const V3* g, Index f) {  // operators.cpp:46-60
    const Mesh& m = *x.m;
    const Index o = m.own[f], n = m.nei[f];
    const double w = m.w[f];
    const V3 fg = w * g[o] + (1.0 - w) * g[n];
    return internalFaceValue(x, f) + dot(m.skew[f], fg);
}
V3 correctedInternal(const VField& x, const T3* g, Index f) {  // operators.cpp:62-79
    const Mesh& m = *x.m;
    const Index o = m.own[f], n = m.nei[f];
    const double w = m.w[f];
    V3 value = internalFaceValue(x, f);
    for (int c = 0; c < 3; ++c) {
        const V3 fg = w * g[o].r[c] + (1.0 - w) * g[n].r[c];
        value[c] += dot(m.skew[f], fg);
    }
    return value;
}

double reconstructedBoundary(const SField& x, const V3* g, Index f, double flux) {  // operators.cpp:111-130
    const Mesh& m = *x.m;
    const Index o = m.own[f];
    const BC<double> cond = faceCondition(x, f, flux);
    if (cond.type == FixedValue || (cond.type == InletOutlet && flux < 0.0)) return cond.value;
    const V3 n = m.nrm[f];
    const double pn = cond.type == FixedGradient ? cond.value : 0.0;
    const V3 cg = g[o] + (pn - dot(g[o], n)) * n;
    return x.v[o] + dot(cg, m.Cf[f] - m.C[o]);
}
V3 reconstructedBoundary(const VField& x, const T3* g, Index f, double flux) {  // operators.cpp:132-161
    const Mesh& m = *x.m;
    const Index o = m.own[f];
    const BC<V3> cond = faceCondition(x, f, flux);
    if (cond.type == FixedValue || (cond.type == InletOutlet && flux < 0.0)) return cond.value;
    const V3 n = m.nrm[f];
    const V3 off = m.Cf[f] - m.C[o];
    V3 value = x.v[o];
    for (int c = 0; c < 3; ++c) {
        const double pn = cond.type == FixedGradient ? cond.value[c] : 0.0;
        const V3 cg = g[o].r[c] + (pn - dot(g[o].r[c], n)) * n;
        value[c] += dot(cg, off);
    }
    if (cond.type == Symmetry) value -= dot(value, n) * n;
    return value;
}

double luCorrection(const SField& x, const V3* g, Index cell, Index f) {  // operators.cpp:81-91
    return dot(g[cell], x.m->Cf[f] - x.m->C[cell]);
}
V3 luCorrection(const VField& x, const T3* g, Index cell, Index f) {  // operators.cpp:93-109
    const V3 off = x.m->Cf[f] - x.m->C[cell];
    V3 c;
    for (int k = 0; k < 3; ++k) c[k] = dot(g[cell].r[k], off);
    return c;
}

double limitCorrection(double orthogonal, double correction, int method) {  // operators.cpp:499-513
    if (method == CorrectedDiff) return correction;
    if (method != LimitedCorrected) return 0.0;
    const double b = std::abs(orthogonal);
    return std::clamp(correction, -b, b);
}
bool correctedDiffusion(int d) { return d == CorrectedDiff || d == LimitedCorrected; }

// 3x3 symmetric solve by pivoted LDL^T (the reference uses Eigen::LDLT; the
// solution is unique for an SPD normal matrix, rounding may differ).
bool solveSym3(double A[3][3], const double* b, int nrhs, double* x) {
    int perm[3] = {0, 1, 2};
    double M[3][3];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) M[i][j] = A[i][j];
    double L[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, D[3];
    for (int k = 0; k < 3; ++k) {
        int p = k;
        for (int i = k + 1; i < 3; ++i) if (std::abs(M[i][i]) > std::abs(M[p][p])) p = i;
        if (p != k) {
            std::swap(perm[k], perm[p]);
            for (int j = 0; j < 3; ++j) std::swap(M[k][j], M[p][j]);
            for (int i = 0; i < 3; ++i) std::swap(M[i][k], M[i][p]);
            for (int j = 0; j < k; ++j) std::swap(L[k][j], L[p][j]);
        }
        D[k] = M[k][k];
        if (!(std::abs(D[k]) > 0.0) || !std::isfinite(D[k])) return false;
        for (int i = k + 1; i < 3; ++i) {
            L[i][k] = M[i][k] / D[k];
            for (int j = k + 1; j < 3; ++j) M[i][j] -= L[i][k] * M[k][j];
        }
    }
    for (int r = 0; r < nrhs; ++r) {
        double y[3];
        for (int i = 0; i < 3; ++i) y[i] = b[perm[i] * nrhs + r];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < i; ++j) y[i] -= L[i][j] * y[j];
        for (int i = 0; i < 3; ++i) y[i] /= D[i];
        for (int i = 2; i >= 0; --i) for (int j = i + 1; j < 3; ++j) y[i] -= L[j][i] * y[j];
        for (int i = 0; i < 3; ++i) x[perm[i] * nrhs + r] = y[i];
    }
    for (int i = 0; i < 3 * nrhs; ++i) if (!std::isfinite(x[i])) return false;
    return true;
}

}  // namespace

template <class T> BC<T> faceCondition(const Field<T>& f, Index face, double flux) {  // field.h:285-301
    const Mesh& m = *f.m;
    if (!m.boundary(face)) throw std::invalid_argument("expected boundary face");
    if (f.calculated) return BC<T>{FixedValue, f.trace[face]};
    BC<T> c = f.bc[m.patch[face]];
    if (c.type == InletOutlet) {
        if (!std::isfinite(flux)) {
            if (f.fluxCtx.empty()) throw std::invalid_argument("inletOutlet requires a boundary flux context");
            flux = f.fluxCtx[face];
        }
        if (!std::isfinite(flux)) throw std::invalid_argument("nonfinite boundary flux");
        return flux < 0.0 ? BC<T>{FixedValue, c.value} : BC<T>{ZeroGradient, T{}};
    }
    return c;
}

template <class T> T boundaryTrace(const Field<T>& f, Index face, double flux) {  // field.h:302-313
    const Mesh& m = *f.m;
    const BC<T> c = faceCondition(f, face, flux);
    const T& o = f.v[m.own[face]];
    if (c.type == FixedValue) return c.value;
    if (c.type == FixedGradient) return o + dot(m.Cf[face] - m.C[m.own[face]], m.nrm[face]) * c.value;
    if (c.type == Symmetry) return symmetric(o, m.nrm[face]);
    return o;
}

template <class T> T correctedFaceValue(const Field<T>& f, const typename GradOf<T>::type* g, Index face,
                                        double flux) {  // operators.cpp:163-183
    return f.m->boundary(face) ? reconstructedBoundary(f, g, face, flux) : correctedInternal(f, g, face);
}

// Face loops that scatter into both cells of a face run as gathers: every
// cell, on its own thread, visits its faces in ascending face order and
// applies the face's update to its own entries.  Each entry then receives the
// same updates in the same order as in a serial loop over faces, so results
// are bit-identical to that loop and independent of the thread count.  The
// owner's visit also writes the face's own entries (LDU upper/lower).
template <class F> void cellFaces(const Mesh& m, F&& face) {
    parallelFor(m.nc, [&](Index c) {
        for (Index k = m.cfo[c]; k < m.cfo[c + 1]; ++k) {
            const Index f = m.cfi[k];
            face(f, c, m.own[f] == c);
        }
    });
}

// Green-Gauss, two passes on non-orthogonal meshes (operators.cpp:275-316, 365-405).
template <class T> void greenGauss(const Field<T>& x, std::vector<typename GradOf<T>::type>& out) {
    using G = typename GradOf<T>::type;
    const Mesh& m = *x.m;
    out.assign(m.nc, G{});
    std::vector<T> fv(m.nf);
    parallelFor(m.nf, [&](Index f) { fv[f] = interpolatedFaceValue(x, f); });
    cellFaces(m, [&](Index f, Index cell, bool own) {
        const T value = fv[f];
        for (int c = 0; c < ncomp<T>(); ++c) {
            const V3 contrib = comp(value, c) * m.Sf[f];
            if (own) gradRow(out[cell], c) += contrib * m.invV[cell];
            else gradRow(out[cell], c) -= contrib * m.invV[cell];
        }
    });
    if (m.orthogonal) return;
    const std::vector<G> initial = out;
    out.assign(m.nc, G{});
    parallelFor(m.nf, [&](Index f) { fv[f] = correctedFaceValue(x, initial.data(), f); });
    cellFaces(m, [&](Index f, Index cell, bool own) {
        const T value = fv[f];
        for (int c = 0; c < ncomp<T>(); ++c) {
            const V3 contrib = comp(value, c) * m.Sf[f];
            if (own) gradRow(out[cell], c) += contrib * m.invV[cell];
            else gradRow(out[cell], c) -= contrib * m.invV[cell];
        }
    });
}

// Least squares with boundary mirror rows (operators.cpp:318-363, 407-458).
template <class T> void leastSquares(const Field<T>& x, std::vector<typename GradOf<T>::type>& out) {
    using G = typename GradOf<T>::type;
    const Mesh& m = *x.m;
    constexpr int nc = ncomp<T>();
    out.assign(m.nc, G{});
    parallelFor(m.nc, [&](Index cell) {
        double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        double rhs[3 * 3] = {0};
        for (Index k = m.cfo[cell]; k < m.cfo[cell + 1]; ++k) {
            const Index f = m.cfi[k];
            const Index o = m.own[f], n = m.nei[f];
            V3 d{};
            T diff{};
            if (n != kNone) {
                const Index other = o == cell ? n : o;
                d = m.C[other] - m.C[cell];
                diff = x.v[other] - x.v[cell];
            } else {
                const int type = faceCondition(x, f).type;
                const V3 off = m.Cf[f] - m.C[cell];
                d = type == FixedValue ? 2.0 * off : (2.0 * boundaryNormalDistance(m, f)) * m.nrm[f];
                diff = 2.0 * (boundaryTrace(x, f) - x.v[cell]);
            }
            const double d2 = norm2(d);
            if (!(d2 > 0.0)) throw std::runtime_error("least-squares neighbour distance is invalid");
            const double wgt = 1.0 / d2;
            const V3 wd = wgt * d;
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) A[i][j] += wd[i] * d[j];
            for (int i = 0; i < 3; ++i) for (int c = 0; c < nc; ++c) rhs[i * nc + c] += wd[i] * comp(diff, c);
        }
        double sol[9];
        if (!solveSym3(A, rhs, nc, sol)) throw std::runtime_error("least-squares gradient matrix is singular");
        for (int c = 0; c < nc; ++c) gradRow(out[cell], c) = V3{sol[0 * nc + c], sol[1 * nc + c], sol[2 * nc + c]};
    });
}

template <class T> void gradient(const Field<T>& f, std::vector<typename GradOf<T>::type>& out, int method) {
    if (method == GreenGauss) greenGauss(f, out);
    else if (method == LeastSquares) leastSquares(f, out);
    else throw std::invalid_argument("unsupported gradient method");
}

template <class T> void interpolate(const Field<T>& f, std::vector<T>& out, int method, int gmethod) {
    // operators.cpp:250-265, 705-745
    const Mesh& m = *f.m;
    out.assign(m.nf, T{});
    if (method == Linear) {
        parallelFor(m.nf, [&](Index i) { out[i] = interpolatedFaceValue(f, i); });
        return;
    }
    if (method != Corrected) throw std::invalid_argument("unsupported interpolation method");
    std::vector<typename GradOf<T>::type> g;
    gradient(f, g, gmethod);
    parallelFor(m.nf, [&](Index i) { out[i] = correctedFaceValue(f, g.data(), i); });
}

template <class T> T integratedNormalGradient(const Field<T>& x, const std::vector<typename GradOf<T>::type>& g,
                                              Index f, int dm) {
    // operators.cpp:857-954
    const Mesh& m = *x.m;
    const Index o = m.own[f], n = m.nei[f];
    T result{};
    if (n != kNone) {
        const double w = m.w[f];
        for (int c = 0; c < ncomp<T>(); ++c) {
            const double orthogonal = m.orth[f] * (comp(x.v[n], c) - comp(x.v[o], c));
            double r = orthogonal;
            if (correctedDiffusion(dm)) {
                const V3 fg = w * gradRow(g[o], c) + (1.0 - w) * gradRow(g[n], c);
                r = orthogonal + limitCorrection(orthogonal, dot(m.kf[f], fg), dm);
            }
            setComp(result, c, r);
        }
        return result;
    }
    const BC<T> cond = faceCondition(x, f);
    if (cond.type == FixedGradient) {
        if constexpr (ncomp<T>() == 1) return cond.value * m.Af[f];
        else return m.Af[f] * cond.value;
    }
    const bool isVector = ncomp<T>() == 3;
    if (cond.type != FixedValue && !(isVector && cond.type == Symmetry)) return T{};
    const T bv = isVector ? boundaryTrace(x, f) : cond.value;
    for (int c = 0; c < ncomp<T>(); ++c) {
        const double orthogonal = m.orth[f] * (comp(bv, c) - comp(x.v[o], c));
        double r = orthogonal;
        if (correctedDiffusion(dm)) {
            V3 bg = gradRow(g[o], c);
            if (cond.type == Symmetry) bg -= dot(bg, m.nrm[f]) * m.nrm[f];
            r = orthogonal + limitCorrection(orthogonal, dot(m.kf[f], bg), dm);
        }
        setComp(result, c, r);
    }
    return result;
}

void flux(VField& U, std::vector<double>& out, int interp, int gmethod) {  // operators.cpp:957-1002
    const Mesh& m = *U.m;
    if (interp != Linear && interp != Corrected) throw std::invalid_argument("unsupported flux interpolation method");
    if (U.fluxCtx.empty()) {
        std::vector<double> init(m.nf, 0.0);
        for (Index f = 0; f < m.nf; ++f)
            if (m.boundary(f)) init[f] = dot(U.v[m.own[f]], m.Sf[f]);
        U.fluxCtx = init;
    }
    std::vector<T3> g;
    if (interp == Corrected) gradient(U, g, gmethod);
    out.assign(m.nf, 0.0);
    parallelFor(m.nf, [&](Index f) {
        const V3 fv = interp == Corrected ? correctedFaceValue(U, g.data(), f) : interpolatedFaceValue(U, f);
        out[f] = dot(fv, m.Sf[f]);
    });
    U.fluxCtx = out;  // fvm_execution.cpp:277-279
}

void divergenceFace(const Mesh& m, const std::vector<double>& phi, std::vector<double>& out) {
    // operators.cpp:1055-1063 via addFaceDivergence (460-475)
    out.assign(m.nc, 0.0);
    cellFaces(m, [&](Index f, Index c, bool own) {
        if (own) out[c] += phi[f] * m.invV[c];
        else out[c] -= phi[f] * m.invV[c];
    });
}

void divergenceVector(const VField& U, std::vector<double>& out, int interp, int gmethod) {
    // operators.cpp:1026-1053
    const Mesh& m = *U.m;
    std::vector<T3> g;
    if (interp == Corrected) gradient(U, g, gmethod);
    out.assign(m.nc, 0.0);
    std::vector<double> qf(m.nf);
    parallelFor(m.nf, [&](Index f) {
        const V3 fv = interp == Corrected ? correctedFaceValue(U, g.data(), f) : interpolatedFaceValue(U, f);
        qf[f] = dot(fv, m.Sf[f]);
    });
    cellFaces(m, [&](Index f, Index c, bool own) {
        const double q = qf[f];
        if (own) out[c] += q * m.invV[c];
        else out[c] -= q * m.invV[c];
    });
}

void normalGradient(const SField& p, const std::vector<V3>* g, int gmethod, int dm, std::vector<double>& out) {
    // fvm_execution.cpp:522-543
    const Mesh& m = *p.m;
    std::vector<V3> work;
    if (!g) { gradient(p, work, gmethod); g = &work; }
    out.assign(m.nf, 0.0);
    parallelFor(m.nf, [&](Index f) { out[f] = integratedNormalGradient(p, *g, f, dm) / m.Af[f]; });
}

void diffusionFlux(const std::vector<double>& kf, const SField& x, const std::vector<V3>& g, int dm,
                   std::vector<double>& out) {  // operators.cpp:1004-1024
    const Mesh& m = *x.m;
    out.assign(m.nf, 0.0);
    parallelFor(m.nf, [&](Index f) {
        const double k = kf[f];
        if (!(k >= 0.0) || !std::isfinite(k)) throw std::invalid_argument("face diffusivity must be non-negative and finite");
        out[f] = k * integratedNormalGradient(x, g, f, dm);
    });
}

template <class T> void addConvection(Ldu<T>& eq, const std::vector<double>& phi, Field<T>& x, int conv,
                                      int interp, int gmethod, double scale) {
    // operators.cpp:535-661
    const Mesh& m = *x.m;
    x.fluxCtx = phi;
    if (interp != Linear && interp != Corrected) throw std::invalid_argument("unsupported convection interpolation method");
    if (!std::isfinite(scale)) throw std::invalid_argument("convection flux scale must be finite");
    std::vector<typename GradOf<T>::type> g;
    const bool haveG = conv == LinearUpwind || (conv == Central && interp == Corrected);
    if (haveG) gradient(x, g, gmethod);
    // F times the deferred correction of each interior face, computed once.
    std::vector<T> fcorr;
    if (haveG) {
        fcorr.assign(m.nf, T{});
        parallelFor(m.nf, [&](Index f) {
            const Index n = m.nei[f];
            if (n == kNone) return;
            const double F = scale * phi[f];
            if (conv == LinearUpwind) fcorr[f] = F * luCorrection(x, g.data(), F >= 0.0 ? m.own[f] : n, f);
            else if (conv == Central) fcorr[f] = F * (correctedInternal(x, g.data(), f) - internalFaceValue(x, f));
        });
    }
    cellFaces(m, [&](Index f, Index o, bool own) {
        const Index n = m.nei[f];
        const double F = scale * phi[f];
        if (n != kNone) {
            // The owner's and the neighbour's halves of the face's update.
            if (conv == Upwind || conv == LinearUpwind) {
                if (own) {
                    eq.diag[o] += std::max(F, 0.0);
                    eq.upper[f] += std::min(F, 0.0);
                    eq.lower[f] += std::min(-F, 0.0);
                } else {
                    eq.diag[o] += std::max(-F, 0.0);
                }
                if (conv == LinearUpwind) {
                    if (own) eq.source[o] -= fcorr[f];
                    else eq.source[o] += fcorr[f];
                }
            } else if (conv == Central) {
                const double w = m.w[f];
                if (own) {
                    eq.diag[o] += F * w;
                    eq.upper[f] += F * (1.0 - w);
                    eq.lower[f] -= F * w;
                } else {
                    eq.diag[o] -= F * (1.0 - w);
                }
                if (haveG) {
                    if (own) eq.source[o] -= fcorr[f];
                    else eq.source[o] += fcorr[f];
                }
            } else {
                throw std::invalid_argument("unsupported convection method");
            }
            return;
        }
        const BC<T> cond = faceCondition(x, f, F);
        if (conv == Upwind || conv == LinearUpwind) {
            if (F >= 0.0) {
                eq.diag[o] += F;
                if (conv == LinearUpwind) eq.source[o] -= F * luCorrection(x, g.data(), o, f);
            } else {
                eq.source[o] -= F * boundaryTrace(x, f, F);
            }
            return;
        }
        if (conv != Central) throw std::invalid_argument("unsupported convection method");
        switch (cond.type) {
            case FixedValue: eq.source[o] -= F * cond.value; break;
            case FixedGradient: {
                const double dist = boundaryNormalDistance(m, f);
                eq.diag[o] += F;
                eq.source[o] -= (F * dist) * cond.value;
                break;
            }
            case ZeroGradient:
            case Symmetry: eq.diag[o] += F; break;
            case InletOutlet:
                if (F >= 0.0) eq.diag[o] += F; else eq.source[o] -= F * cond.value;
                break;
        }
        if (haveG) {
            const T corr = reconstructedBoundary(x, g.data(), f, F) - boundaryTrace(x, f, F);
            eq.source[o] -= F * corr;
        }
    });
}

template <class T> void addDiffusion(Ldu<T>& eq, const std::vector<double>& kf, const Field<T>& x, int gmethod,
                                     int dm, const std::vector<typename GradOf<T>::type>* given) {
    // operators.cpp:1287-1438 (scalar and vector variants)
    const Mesh& m = *x.m;
    constexpr bool isVector = ncomp<T>() == 3;
    std::vector<typename GradOf<T>::type> work;
    const std::vector<typename GradOf<T>::type>* g = given;
    if (correctedDiffusion(dm)) {
        if (!g) { gradient(x, work, gmethod); g = &work; }
    } else if (dm != Orthogonal) {
        throw std::invalid_argument("unsupported diffusion method");
    }
    // The deferred correction of each interior face, computed once.
    std::vector<T> fcorr;
    if (correctedDiffusion(dm)) {
        fcorr.assign(m.nf, T{});
        parallelFor(m.nf, [&](Index f) {
            const Index o = m.own[f], n = m.nei[f];
            const double gamma = kf[f];
            if (n == kNone || !(gamma > 0.0) || !std::isfinite(gamma)) return;
            const T orthogonal = m.orth[f] * (x.v[n] - x.v[o]);
            fcorr[f] = gamma * (integratedNormalGradient(x, *g, f, dm) - orthogonal);
        });
    }
    cellFaces(m, [&](Index f, Index c, bool own) {
        const Index o = m.own[f], n = m.nei[f];
        const double gamma = kf[f];
        if (gamma < 0.0 || !std::isfinite(gamma)) throw std::invalid_argument("face diffusivity must be non-negative and finite");
        if (gamma == 0.0) return;
        const double coef = gamma * m.orth[f];
        if (n != kNone) {
            eq.diag[c] += coef;
            if (own) {
                eq.upper[f] -= coef;
                eq.lower[f] -= coef;
            }
            if (correctedDiffusion(dm)) {
                if (own) eq.source[c] += fcorr[f];
                else eq.source[c] -= fcorr[f];
            }
            return;
        }
        const BC<T> cond = faceCondition(x, f);
        if (cond.type == FixedValue || (isVector && cond.type == Symmetry)) {
            const T bv = isVector ? boundaryTrace(x, f, -1.0) : cond.value;
            eq.diag[o] += coef;
            eq.source[o] += coef * bv;
            if (correctedDiffusion(dm)) {
                const T orthogonal = m.orth[f] * (bv - x.v[o]);
                eq.source[o] += gamma * (integratedNormalGradient(x, *g, f, dm) - orthogonal);
            }
        } else if (cond.type == FixedGradient) {
            if constexpr (isVector) eq.source[o] += (gamma * m.Af[f]) * cond.value;
            else eq.source[o] += (gamma * cond.value) * m.Af[f];
        }
    });
}

template <class T> void explicitConvection(const std::vector<double>& phi, const Field<T>& x, std::vector<T>& out,
                                           int conv, int interp, int gmethod) {
    // operators.cpp:1065-1131
    const Mesh& m = *x.m;
    std::vector<typename GradOf<T>::type> g;
    const bool haveG = conv == LinearUpwind || (conv == Central && interp == Corrected);
    if (haveG) gradient(x, g, gmethod);
    out.assign(m.nc, T{});
    cellFaces(m, [&](Index f, Index c, bool own) {
        const Index o = m.own[f], n = m.nei[f];
        const double q = phi[f];
        T fv{};
        if (conv == Upwind || conv == LinearUpwind) {
            if (n != kNone) {
                const Index up = q >= 0.0 ? o : n;
                fv = x.v[up];
                if (conv == LinearUpwind) fv += luCorrection(x, g.data(), up, f);
            } else if (q >= 0.0) {
                fv = x.v[o];
                if (conv == LinearUpwind) fv += luCorrection(x, g.data(), o, f);
            } else {
                fv = boundaryTrace(x, f, q);
            }
        } else if (haveG) {
            fv = correctedFaceValue(x, g.data(), f, q);
        } else {
            fv = interpolatedFaceValue(x, f, q);
        }
        const T integ = q * fv;
        if (own) out[c] += integ * m.invV[c];
        else out[c] -= integ * m.invV[c];
    });
}

void explicitLaplacian(const SField& x, const std::vector<double>* kf, double k, int gmethod, int dm,
                       std::vector<double>& out) {
    // operators.cpp:1157-1213
    const Mesh& m = *x.m;
    std::vector<V3> g;
    if (correctedDiffusion(dm)) gradient(x, g, gmethod);
    else if (dm != Orthogonal) throw std::invalid_argument("unsupported diffusion method");
    out.assign(m.nc, 0.0);
    cellFaces(m, [&](Index f, Index c, bool own) {
        const Index o = m.own[f], n = m.nei[f];
        double q = 0.0;
        if (n != kNone) {
            q = m.orth[f] * (x.v[n] - x.v[o]);
            if (correctedDiffusion(dm)) q = integratedNormalGradient(x, g, f, dm);
        } else {
            const BC<double> c = faceCondition(x, f);
            if (c.type == FixedValue) {
                q = m.orth[f] * (boundaryTrace(x, f, -1.0) - x.v[o]);
                if (correctedDiffusion(dm)) q = integratedNormalGradient(x, g, f, dm);
            } else if (c.type == FixedGradient) {
                q = c.value * m.Af[f];
            }
        }
        const double d = kf ? (*kf)[f] : k;
        const double integ = d * q;
        if (own) out[c] += integ * m.invV[c];
        else out[c] -= integ * m.invV[c];
    });
}

void gradientTraces(const SField& p, const std::vector<V3>& g, std::vector<V3>& traces) {
    // fvm_execution.cpp:221-235
    const Mesh& m = *p.m;
    traces.assign(m.nf, V3{});
    for (Index f = 0; f < m.nf; ++f) {
        if (!m.boundary(f)) continue;
        const Index o = m.own[f];
        const V3 n = m.nrm[f];
        const V3 d = m.Cf[f] - m.C[o];
        V3 gv = g[o];
        const BC<double> bc = faceCondition(p, f);
        const double sn = bc.type == FixedValue ? (bc.value - p.v[o] - dot(gv, d - dot(d, n) * n)) / dot(d, n)
                        : bc.type == FixedGradient ? bc.value : 0.0;
        gv += (sn - dot(gv, n)) * n;
        traces[f] = gv;
    }
}

template BC<double> faceCondition(const SField&, Index, double);
template BC<V3> faceCondition(const VField&, Index, double);
template double boundaryTrace(const SField&, Index, double);
template V3 boundaryTrace(const VField&, Index, double);
template double correctedFaceValue(const SField&, const V3*, Index, double);
template V3 correctedFaceValue(const VField&, const T3*, Index, double);
template void gradient(const SField&, std::vector<V3>&, int);
template void gradient(const VField&, std::vector<T3>&, int);
template void interpolate(const SField&, std::vector<double>&, int, int);
template void interpolate(const VField&, std::vector<V3>&, int, int);
template double integratedNormalGradient(const SField&, const std::vector<V3>&, Index, int);
template V3 integratedNormalGradient(const VField&, const std::vector<T3>&, Index, int);
template void addConvection(Ldu<double>&, const std::vector<double>&, SField&, int, int, int, double);
template void addConvection(Ldu<V3>&, const std::vector<double>&, VField&, int, int, int, double);
template void addDiffusion(Ldu<double>&, const std::vector<double>&, const SField&, int, int, const std::vector<V3>*);
template void addDiffusion(Ldu<V3>&, const std::vector<double>&, const VField&, int, int, const std::vector<T3>*);
template void explicitConvection(const std::vector<double>&, const SField&, std::vector<double>&, int, int, int);
template void explicitConvection(const std::vector<double>&, const VField&, std::vector<V3>&, int, int, int);

}  // namespace orc
