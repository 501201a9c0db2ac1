// Oracle Wilcox (1988) k-omega transport: restates
// src/physics/RANS/k_omega.cpp:12-157 (closure, segregated omega-then-k
// solves, bounding) and model.cpp:67-71 (the deviatoric stress remainder of
// the momentum equation), with the calculated boundary traces of the field
// algebra (include/babelsim/math.h:185-258) and the vector-gradient traces of
// fvm_execution.cpp:243-265.  Steady form for SIMPLE; in PISO the k and omega
// histories add the ddt terms of k_omega.cpp:59-62, 82, 96.
// Test infrastructure only.
#include "rans.hpp"

#include "simple.hpp"

#include <algorithm>

namespace orc {

double strainSquared(const T3& g) {  // k_omega.cpp:12-23
    const double divergence = g.r[0][0] + g.r[1][1] + g.r[2][2];
    double sum = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double strain = 0.5 * (g.r[i][j] + g.r[j][i]) - (i == j ? divergence / 3.0 : 0.0);
            sum += strain * strain;
        }
    return 2.0 * sum;
}

namespace {

T3 symmetricTensor(const T3& v, const V3& n) {  // include/babelsim/vector.h:107-115
    T3 reflected{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l)
                    reflected.r[i][j] += ((i == k ? 1.0 : 0.0) - 2 * n[i] * n[k]) * v.r[k][l] *
                                         ((l == j ? 1.0 : 0.0) - 2 * n[l] * n[j]);
    T3 out;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out.r[i][j] = 0.5 * (v.r[i][j] + reflected.r[i][j]);
    return out;
}

double relChange(const std::vector<double>& a, const std::vector<double>& b) {  // fvm_execution.cpp:110-128
    double d2 = 0, c2 = 0;
    for (std::size_t i = 0; i < a.size(); ++i) {
        const double d = a[i] - b[i];
        d2 += d * d;
        c2 += a[i] * a[i];
    }
    return std::sqrt(d2) / std::max(std::sqrt(c2), 1e-30);
}

}  // namespace

void gradientTracesVec(const VField& U, const std::vector<T3>& g, std::vector<T3>& traces) {
    const Mesh& m = *U.m;
    traces.assign(m.nf, T3{});
    for (Index f = 0; f < m.nf; ++f) {
        if (!m.boundary(f)) continue;
        const Index o = m.own[f];
        const V3 n = m.nrm[f];
        const V3 d = m.Cf[f] - m.C[o];
        T3 gv = g[o];
        const BC<V3> bc = faceCondition(U, f);
        if (bc.type == Symmetry) {
            gv = symmetricTensor(gv, n);
        } else {
            for (int c = 0; c < 3; ++c) {
                const double sn = bc.type == FixedValue
                                      ? (bc.value[c] - U.v[o][c] - dot(gv.r[c], d - dot(d, n) * n)) / dot(d, n)
                                  : bc.type == FixedGradient ? bc.value[c] : 0.0;
                gv.r[c] += (sn - dot(gv.r[c], n)) * n;
            }
        }
        traces[f] = gv;
    }
}

void tensorDivergence(const Mesh& m, const std::vector<T3>& T, const std::vector<T3>& Ttrace, int interp, int gmethod,
                      std::vector<V3>& out) {
    out.assign(m.nc, V3{});
    VField row(m);
    row.calculated = true;
    row.trace.assign(m.nf, V3{});
    std::vector<V3> fv;
    for (int c = 0; c < 3; ++c) {
        for (Index i = 0; i < m.nc; ++i) row.v[i] = T[i].r[c];
        for (Index f = 0; f < m.nf; ++f)
            if (m.boundary(f)) row.trace[f] = Ttrace[f].r[c];
        interpolate(row, fv, interp, gmethod);
        for (Index f = 0; f < m.nf; ++f) {
            const Index o = m.own[f], n = m.nei[f];
            const double contribution = dot(fv[f], m.Sf[f]);
            out[o][c] += contribution / m.V[o];
            if (n != kNone) out[n][c] -= contribution / m.V[n];
        }
    }
}

KOmega::KOmega(const Mesh& mesh, const KOmegaConfig& cfg, double density, double viscosity)
    : m(mesh), c(cfg), rho(density), mu(viscosity) {
    if (!(c.relaxation > 0 && c.relaxation <= 1)) throw std::invalid_argument("relaxation must be in (0,1]");
    if (!(c.kMin > 0) || !(c.omegaMin > 0) || !(c.viscosityRatioMax > 0))
        throw std::invalid_argument("k-omega bounds must be positive");
    k = SField(m);
    omega = SField(m);
    mut.assign(m.nc, 0.0);
    mutTrace.assign(m.nf, 0.0);
}

void KOmega::setBoundaries(const std::vector<BC<double>>& kbc, const std::vector<BC<double>>& wbc) {
    k.bc = kbc;
    omega.bc = wbc;
}

void KOmega::boundOmega(double& bounded) {  // k_omega.cpp:131-144
    double count = 0;
    for (Index i = 0; i < m.nc; ++i) {
        const double w = omega.v[i], energy = k.v[i];
        const double floor = rho * energy / (c.viscosityRatioMax * mu);
        count += w < std::max(c.omegaMin, floor) ? 1.0 : 0.0;
        omega.v[i] = std::max({w, c.omegaMin, floor});
    }
    bounded = std::max(bounded, count);
}

void KOmega::updateViscosity(const std::vector<double>&) {  // k_omega.cpp:146-149
    for (Index i = 0; i < m.nc; ++i) mut[i] = (rho * k.v[i]) / std::max(omega.v[i], c.omegaMin);
    for (Index f = 0; f < m.nf; ++f) {
        if (!m.boundary(f)) continue;
        const double kb = boundaryTrace(k, f), wb = boundaryTrace(omega, f);
        mutTrace[f] = (rho * kb) / std::max(wb, c.omegaMin);
    }
}

void KOmega::start(const std::vector<double>& phi) {  // constructor, k_omega.cpp:50-53
    k.fluxCtx = phi;
    omega.fluxCtx = phi;
    for (auto& v : k.v) v = std::max(v, c.kMin);
    double bounded = 0;
    boundOmega(bounded);
    updateViscosity(phi);
}

SField KOmega::effectiveViscosity() const {  // muEff = mu + mut with traces
    SField e(m);
    e.calculated = true;
    e.trace.assign(m.nf, 0.0);
    for (Index i = 0; i < m.nc; ++i) e.v[i] = mut[i] + mu;
    for (Index f = 0; f < m.nf; ++f)
        if (m.boundary(f)) e.trace[f] = mutTrace[f] + mu;
    return e;
}

void KOmega::stressDivergence(const VField& U, int mInterp, int mGrad, std::vector<V3>& out) const {
    // muEff * (transpose(grad U) - (2/3) isotropic(trace(grad U))), values and
    // traces by the same pointwise operations (model.cpp:67-71).
    std::vector<T3> g, gt;
    gradient(U, g, mGrad);
    gradientTracesVec(U, g, gt);
    const SField e = effectiveViscosity();
    auto stress = [](double mu_, const T3& G) {
        const double tr = G.r[0][0] + G.r[1][1] + G.r[2][2];
        T3 s;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double iso = i == j ? tr : 0.0;
                s.r[i][j] = mu_ * (G.r[j][i] - (2.0 / 3.0) * iso);
            }
        return s;
    };
    std::vector<T3> T(m.nc), Tt(m.nf);
    for (Index i = 0; i < m.nc; ++i) T[i] = stress(e.v[i], g[i]);
    for (Index f = 0; f < m.nf; ++f)
        if (m.boundary(f)) Tt[f] = stress(e.trace[f], gt[f]);
    tensorDivergence(m, T, Tt, mInterp, mGrad, out);
}

namespace {

void solveScalar(const Mesh& m, LinearSystem& sys, const Ldu<double>& q, const SolverConfig& sc,
                 std::vector<double>& x, SolveResult& res, std::vector<double>& hist) {
    Csr A;
    lduToCsr(m, q.diag, q.upper, q.lower, A);
    if (!sys.ready) sys.A = A;
    sys.update(A.diag, A.val);
    sys.prepare(sc);
    res = sys.solve(q.source, x, 0);
    hist = sys.hist;
}

// One transport equation: convection, diffusion with a calculated
// diffusivity, implicit reaction and explicit source (k_omega.cpp:82-87, 96-101).
// equ::ddt(equation, rho, history) (procedural_equation.cpp:198-239): weights
// a0, a1, a2 of the history's method (Euler below two saved levels).
struct TimeTerm {
    double a0 = 0, a1 = 0, a2 = 0;
    const std::vector<double>* previous = nullptr;
    const std::vector<double>* older = nullptr;
};

Ldu<double> transportEquation(const Mesh& m, SField& x, const std::vector<double>& phi, double rho,
                              const TransportOptions& o, const SField& diffusivity, const std::vector<double>& rate,
                              const std::vector<double>& src, const TimeTerm* tt) {
    Ldu<double> eq;
    eq.init(m);
    if (tt)
        for (Index i = 0; i < m.nc; ++i) {
            const double w = rho * m.V[i];
            eq.diag[i] += w * tt->a0;
            eq.source[i] += w * tt->a1 * (*tt->previous)[i];
            if (tt->a2 != 0) eq.source[i] += w * tt->a2 * (*tt->older)[i];
        }
    addConvection(eq, phi, x, o.conv, o.interp, o.grad, rho);
    {
        std::vector<double> kf;
        interpolate(diffusivity, kf, o.coefInterp, o.coefGrad);
        Ldu<double> contrib;
        contrib.init(m);
        addDiffusion<double>(contrib, kf, x, o.grad, o.diff, nullptr);
        accumulate(eq, contrib, -(-1.0));
    }
    for (Index i = 0; i < m.nc; ++i) eq.diag[i] += 1.0 * rate[i] * m.V[i];   // procedural_equation.cpp:181-187
    for (Index i = 0; i < m.nc; ++i) eq.source[i] += 1.0 * m.V[i] * src[i];  // procedural_equation.cpp:192-196
    return eq;
}

void relax(Ldu<double>& eq, const std::vector<double>& prev, double alpha) {  // procedural_equation.cpp:298-305
    for (std::size_t i = 0; i < eq.diag.size(); ++i) {
        const double extra = eq.diag[i] * (1 / alpha - 1);
        eq.diag[i] += extra;
        eq.source[i] += extra * prev[i];
    }
}

}  // namespace

void KOmega::saveOld(double dt) {
    kOlder = kPrev.empty() ? k.v : kPrev;
    wOlder = wPrev.empty() ? omega.v : wPrev;
    kPrev = k.v;
    wPrev = omega.v;
    hdtPrev = hdt;
    hdt = dt;
    ++hlevels;
}

TransportReport KOmega::solveTransport(VField& U, const std::vector<double>& phi, int mGrad) {
    TransportReport rep;
    TimeTerm tk, tw;
    const TimeTerm* ptk = nullptr;
    const TimeTerm* ptw = nullptr;
    if (timeMethod != 0) {
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
        tk = TimeTerm{a0, a1, a2, &kPrev, &kOlder};
        tw = TimeTerm{a0, a1, a2, &wPrev, &wOlder};
        ptk = &tk;
        ptw = &tw;
    }
    const std::vector<double> prevK = k.v, prevW = omega.v;
    U.fluxCtx = phi;
    updateViscosity(phi);
    std::vector<T3> g;
    gradient(U, g, mGrad);
    std::vector<double> production(m.nc);
    for (Index i = 0; i < m.nc; ++i) production[i] = mut[i] * strainSquared(g[i]);
    auto diffusivity = [&](double sigma) {
        SField d(m);
        d.calculated = true;
        d.trace.assign(m.nf, 0.0);
        for (Index i = 0; i < m.nc; ++i) d.v[i] = sigma * mut[i] + mu;
        for (Index f = 0; f < m.nf; ++f)
            if (m.boundary(f)) d.trace[f] = sigma * mutTrace[f] + mu;
        return d;
    };
    const SField kDiff = diffusivity(c.sigmaK), wDiff = diffusivity(c.sigmaOmega);
    std::vector<double> rate(m.nc), src(m.nc);
    for (Index i = 0; i < m.nc; ++i) {
        rate[i] = (c.beta * rho) * omega.v[i];
        src[i] = ((c.gamma * production[i]) * omega.v[i]) / std::max(k.v[i], c.kMin);
    }
    // omega first (k_omega.cpp:80-93)
    Ldu<double> weq = transportEquation(m, omega, phi, rho, c.wOpts, wDiff, rate, src, ptw);
    rep.rOmega = relativeResidual(m, weq, omega.v);
    relax(weq, prevW, c.relaxation);
    SolveResult r;
    solveScalar(m, wSys, weq, c.wSolver, omega.v, r, wHist);
    rep.itOmega = r.iterations;
    rep.statusOmega = r.status;
    if (r.status == NumericalFailure) { rep.healthy = false; return rep; }
    double bounded = 0;
    boundOmega(bounded);
    // then k with the new omega in its sink (k_omega.cpp:95-106)
    for (Index i = 0; i < m.nc; ++i) rate[i] = (c.betaStar * rho) * omega.v[i];
    Ldu<double> keq = transportEquation(m, k, phi, rho, c.kOpts, kDiff, rate, production, ptk);
    rep.rK = relativeResidual(m, keq, k.v);
    relax(keq, prevK, c.relaxation);
    solveScalar(m, kSys, keq, c.kSolver, k.v, r, kHist);
    rep.itK = r.iterations;
    rep.statusK = r.status;
    if (r.status == NumericalFailure) { rep.healthy = false; return rep; }
    // bound, publish the viscosity (k_omega.cpp:108-111, 124-129)
    k.fluxCtx = phi;
    omega.fluxCtx = phi;
    for (auto& v : k.v) v = std::max(v, c.kMin);
    boundOmega(bounded);
    updateViscosity(phi);
    rep.boundedOmegaCells = bounded;
    rep.dOmega = relChange(omega.v, prevW);
    rep.dK = relChange(k.v, prevK);
    rep.healthy = std::isfinite(rep.rOmega) && std::isfinite(rep.rK) && std::isfinite(rep.dOmega) &&
                  std::isfinite(rep.dK);
    return rep;
}

}  // namespace orc
