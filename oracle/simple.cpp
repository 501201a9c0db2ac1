// Oracle SIMPLE outer iteration: restates src/physics/simple/main.cpp:47-127
// together with the equ:: semantics it uses (src/discretization/
// procedural_equation.cpp) and the diagnostics of fvm_execution.cpp:91-177.
// Test infrastructure only.
#include "simple.hpp"

#include <algorithm>

namespace orc {

template <class T> void accumulate(Ldu<T>& a, const Ldu<T>& b, double f) {  // procedural_equation.cpp:62-70
    for (std::size_t i = 0; i < a.diag.size(); ++i) { a.diag[i] += f * b.diag[i]; a.source[i] += f * b.source[i]; }
    for (std::size_t i = 0; i < a.upper.size(); ++i) { a.upper[i] += f * b.upper[i]; a.lower[i] += f * b.lower[i]; }
}
template void accumulate(Ldu<double>&, const Ldu<double>&, double);
template void accumulate(Ldu<V3>&, const Ldu<V3>&, double);

template <class T> void applyLdu(const Mesh& m, const Ldu<T>& q, const std::vector<T>& x, std::vector<T>& y) {
    // procedural_equation.cpp:281-292
    y.assign(m.nc, T{});
    for (Index i = 0; i < m.nc; ++i) y[i] = q.diag[i] * x[i];
    for (Index f = 0; f < m.nf; ++f) {
        const Index o = m.own[f], n = m.nei[f];
        if (n == kNone) continue;
        y[o] += q.upper[f] * x[n];
        y[n] += q.lower[f] * x[o];
    }
}
template void applyLdu(const Mesh&, const Ldu<double>&, const std::vector<double>&, std::vector<double>&);
template void applyLdu(const Mesh&, const Ldu<V3>&, const std::vector<V3>&, std::vector<V3>&);

SField calculatedRatio(const Mesh& m, const std::vector<double>& num, const std::vector<double>& den) {
    // geometry.cpp:9-22 + math.h fieldBinary: V/aP with traces V_o/aP_o
    SField r(m);
    r.calculated = true;
    r.trace.assign(m.nf, 0.0);
    for (Index i = 0; i < m.nc; ++i) r.v[i] = num[i] / den[i];
    for (Index f = 0; f < m.nf; ++f)
        if (m.boundary(f)) r.trace[f] = num[m.own[f]] / den[m.own[f]];
    return r;
}

namespace {

double sq(double v) { return v * v; }
double sq(const V3& v) { return norm2(v); }

}  // namespace

template <class T> double relativeResidual(const Mesh& m, const Ldu<T>& q, const std::vector<T>& x) {
    // procedural_equation.cpp:429-433
    std::vector<T> ax;
    applyLdu(m, q, x, ax);
    double nr = 0, na = 0, nb = 0;
    for (Index i = 0; i < m.nc; ++i) {
        const T r = q.source[i] - ax[i];
        nr += sq(r); na += sq(ax[i]); nb += sq(q.source[i]);
    }
    nr = std::sqrt(nr); na = std::sqrt(na); nb = std::sqrt(nb);
    return nr / std::max(na + nb, 1e-30);
}
template double relativeResidual(const Mesh&, const Ldu<double>&, const std::vector<double>&);
template double relativeResidual(const Mesh&, const Ldu<V3>&, const std::vector<V3>&);

bool referenceIfUnanchored(const Mesh& m, Ldu<double>& q, double value) {
    // procedural_equation.cpp:306-333 (global cell 0 is local cell 0 on one rank)
    std::vector<double> rowSum = q.diag, mag(q.diag.size());
    for (std::size_t i = 0; i < mag.size(); ++i) mag[i] = std::abs(q.diag[i]);
    for (Index f = 0; f < m.nf; ++f) {
        const Index o = m.own[f], n = m.nei[f];
        if (n == kNone) continue;
        rowSum[o] += q.upper[f]; mag[o] += std::abs(q.upper[f]);
        rowSum[n] += q.lower[f]; mag[n] += std::abs(q.lower[f]);
    }
    bool constantMode = true;
    for (Index i = 0; i < m.nc; ++i)
        constantMode = constantMode && std::isfinite(rowSum[i]) &&
                       std::abs(rowSum[i]) <= 64 * std::numeric_limits<double>::epsilon() * mag[i];
    if (!constantMode) return false;
    if (!(q.diag[0] > 0)) throw std::invalid_argument("reference requires a positive diagonal");
    q.source[0] += q.diag[0] * value;
    q.diag[0] *= 2;
    return true;
}

Simple::Simple(const Mesh& mesh, const SimpleConfig& cfg) : m(mesh), c(cfg) {
    U = VField(m); p = SField(m); muEff = SField(m); pPrime = SField(m);
    muEff.v.assign(m.nc, c.mu);  // model.cpp:58-63: constant cell field, default (zeroGradient) patches
    if (c.turbulence == 1) turb = std::make_unique<KOmega>(m, c.kw, c.rho, c.mu);
    else if (c.turbulence != 0) throw std::invalid_argument("unsupported turbulence model");
}

void Simple::setBoundaries(const std::vector<BC<V3>>& ubc, const std::vector<BC<double>>& pbc) {
    U.bc = ubc; p.bc = pbc;
    // field.h:382-409: homogeneous counterparts for the correction.
    for (Index k = 0; k < m.np; ++k) {
        const int t = pbc[k].type;
        pPrime.bc[k] = t == FixedValue ? BC<double>{FixedValue, 0.0}
                     : t == Symmetry ? BC<double>{Symmetry, 0.0} : BC<double>{ZeroGradient, 0.0};
    }
}

void Simple::start() {
    // The model is built before phi exists (main.cpp:33, 41): its first
    // bounding sees a zero face flux.
    if (turb) turb->start(std::vector<double>(m.nf, 0.0));
    flux(U, phi, c.mInterp, c.mGrad);  // main.cpp:41: phi = math::flux(U, momentum options)
}

void Simple::solveSystem(LinearSystem& sys, const Ldu<double>& q, const SolverConfig& sc, std::vector<double>& x,
                         SolveResult& res, std::vector<double>* hist) {
    Csr A;
    lduToCsr(m, q.diag, q.upper, q.lower, A);
    if (!sys.ready) sys.A = A;
    sys.update(A.diag, A.val);
    sys.prepare(sc);
    res = sys.solve(q.source, x, 0);
    if (hist) *hist = sys.hist;
}

IterReport Simple::iterate() {
    IterReport rep{};
    const std::vector<V3> Uprev = U.v;
    // Momentum assembly (main.cpp:51-57).
    Ldu<V3> mom; mom.init(m);
    U.fluxCtx = phi;
    addConvection(mom, phi, U, c.mConv, c.mInterp, c.mGrad, c.rho);
    {
        std::vector<double> kf;
        if (turb) interpolate(turb->effectiveViscosity(), kf, c.mCoefInterp, c.mCoefGrad);
        else interpolate(muEff, kf, c.mCoefInterp, c.mCoefGrad);
        Ldu<V3> contrib; contrib.init(m);
        addDiffusion<V3>(contrib, kf, U, c.mGrad, c.mDiff, nullptr);
        accumulate(mom, contrib, -(-1.0));
    }
    std::vector<V3> gradP;
    gradient(p, gradP, c.pGrad);
#pragma omp parallel for schedule(static)
    for (Index i = 0; i < m.nc; ++i) mom.source[i] += (1.0 * m.V[i]) * (-1.0 * gradP[i]);
    if (turb) {  // main.cpp:57-59: the explicit remainder of the eddy-viscosity stress
        std::vector<V3> divT;
        turb->stressDivergence(U, c.mInterp, c.mGrad, divT);
        for (Index i = 0; i < m.nc; ++i) mom.source[i] += (1.0 * m.V[i]) * divT[i];
    }
    // main.cpp:60-66
    rep.rU = relativeResidual(m, mom, U.v);
    if (!std::isfinite(rep.rU)) {  // main.cpp:61
        rep.failed = 1;
        return rep;
    }
    const double alphaU = c.alphaU;
#pragma omp parallel for schedule(static)
    for (Index i = 0; i < m.nc; ++i) {  // relax (procedural_equation.cpp:298-305)
        const double extra = mom.diag[i] * (1 / alphaU - 1);
        mom.diag[i] += extra;
        mom.source[i] += extra * Uprev[i];
    }
    for (auto& v : mom.diag) v *= alphaU;  // scale (procedural_equation.cpp:240-247)
    for (auto& v : mom.upper) v *= alphaU;
    for (auto& v : mom.lower) v *= alphaU;
    for (auto& v : mom.source) v = alphaU * v;
    const std::vector<double> aP = mom.diag;
    const SField rAU = calculatedRatio(m, m.V, aP);
    // Momentum solve: one matrix, three right-hand sides (petsc_backend.cpp:156-198).
    {
        Csr A;
        lduToCsr(m, mom.diag, mom.upper, mom.lower, A);
        if (!momSys.ready) momSys.A = A;
        momSys.update(A.diag, A.val);
        momSys.prepare(c.mom);
        rep.momStatus = Converged;
        for (int comp = 0; comp < 3; ++comp) {
            std::vector<double> b(m.nc), x(m.nc);
            for (Index i = 0; i < m.nc; ++i) { b[i] = mom.source[i][comp]; x[i] = U.v[i][comp]; }
            const SolveResult r = momSys.solve(b, x, comp);
            momHist[comp] = momSys.hist;
            for (Index i = 0; i < m.nc; ++i) U.v[i][comp] = x[i];
            rep.momIters[comp] = r.iterations;
            if (r.status == NumericalFailure) rep.momStatus = NumericalFailure;
            else if (r.status != Converged && rep.momStatus != NumericalFailure) rep.momStatus = MaxIterations;
        }
    }
    if (rep.momStatus == NumericalFailure) {  // main.cpp:67
        rep.failed = 1;
        return rep;
    }
    // Rhie-Chow (main.cpp:71-79).
    gradient(p, gradP, c.pGrad);
    std::vector<double> phiH;
    flux(U, phiH, c.mInterp, c.mGrad);
    {
        std::vector<V3> gtr;
        gradientTraces(p, gradP, gtr);
        VField X(m);
        X.calculated = true;
        X.trace.assign(m.nf, V3{});
#pragma omp parallel for schedule(static)
        for (Index i = 0; i < m.nc; ++i) X.v[i] = rAU.v[i] * gradP[i];
        for (Index f = 0; f < m.nf; ++f) if (m.boundary(f)) X.trace[f] = rAU.trace[f] * gtr[f];
        std::vector<V3> Xf;
        interpolate(X, Xf, c.pInterp, c.pGrad);
        std::vector<double> rAUf, sn;
        interpolate(rAU, rAUf, c.pCoefInterp, c.pCoefGrad);
        normalGradient(p, &gradP, c.pGrad, c.pDiff, sn);
        for (Index f = 0; f < m.nf; ++f) {
            if (m.boundary(f)) continue;
            phiH[f] += dot(Xf[f], m.Sf[f]);
        }
        for (Index f = 0; f < m.nf; ++f) {
            if (m.boundary(f)) continue;
            phiH[f] += -1.0 * ((rAUf[f] * sn[f]) * m.Af[f]);
        }
    }
    std::vector<double> divPhiH;
    divergenceFace(m, phiH, divPhiH);
    // Pressure correction (main.cpp:81-93).
    std::fill(pPrime.v.begin(), pPrime.v.end(), 0.0);
    const int solves = c.pDiff == Orthogonal ? 1 : c.nonOrthCorr + 1;
    bool pConverged = true;
    Ldu<double> peq;
    std::vector<double> kf, correction;
    SField snapshot(m);
    rep.pIters = 0;
    rep.pStatus = Converged;
    for (int k = 0; k < solves; ++k) {
        peq.init(m);
        interpolate(rAU, kf, c.pCoefInterp, c.pCoefGrad);
        Ldu<double> contrib; contrib.init(m);
        std::vector<V3> g;
        if (c.pDiff != Orthogonal) {
            gradient(pPrime, g, c.pGrad);
            addDiffusion<double>(contrib, kf, pPrime, c.pGrad, c.pDiff, &g);
            // diffusionCorrection (procedural_equation.cpp:77-87)
            std::vector<double> full, orthF;
            diffusionFlux(kf, pPrime, g, c.pDiff, full);
            diffusionFlux(kf, pPrime, g, Orthogonal, orthF);
            correction.assign(m.nf, 0.0);
            for (Index f = 0; f < m.nf; ++f) correction[f] = full[f] + (-1.0) * orthF[f];
        } else {
            addDiffusion<double>(contrib, kf, pPrime, c.pGrad, c.pDiff, nullptr);
            correction.assign(m.nf, 0.0);
        }
        snapshot = pPrime;
        accumulate(peq, contrib, -(-1.0));
        for (Index i = 0; i < m.nc; ++i) peq.source[i] += (1.0 * m.V[i]) * (-1.0 * divPhiH[i]);
        rep.referenced = referenceIfUnanchored(m, peq, 0.0);
        SolveResult r;
        solveSystem(pSys, peq, c.pre, pPrime.v, r, &pHist);
        rep.pIters += r.iterations;
        pConverged = pConverged && r.status == Converged;
        rep.linP = r.relative;
        if (r.status == NumericalFailure) {  // main.cpp:88
            rep.pStatus = NumericalFailure;
            rep.failed = 1;
            return rep;
        }
    }
    if (!pConverged && rep.pStatus != NumericalFailure) rep.pStatus = MaxIterations;
    // Corrections (main.cpp:95-98).
#pragma omp parallel for schedule(static)
    for (Index i = 0; i < m.nc; ++i) p.v[i] += c.alphaP * pPrime.v[i];
    std::vector<double> pressureFlux(m.nf, 0.0);
    {
        SField value = snapshot;
        value.v = pPrime.v;
        std::vector<V3> zero(m.nc, V3{});
        std::vector<double> inc;
        diffusionFlux(kf, value, zero, Orthogonal, inc);
        for (Index f = 0; f < m.nf; ++f) inc[f] += 1.0 * correction[f];
        for (Index f = 0; f < m.nf; ++f) pressureFlux[f] += -1.0 * inc[f];
    }
    std::vector<V3> gpp;
    gradient(pPrime, gpp, c.pGrad);
#pragma omp parallel for schedule(static)
    for (Index i = 0; i < m.nc; ++i) U.v[i] += -1.0 * (rAU.v[i] * gpp[i]);
    phi.resize(m.nf);
#pragma omp parallel for schedule(static)
    for (Index f = 0; f < m.nf; ++f) phi[f] = phiH[f] + pressureFlux[f];
    // Turbulence transport (main.cpp:100-110).
    bool turbulenceConverged = true;
    if (turb) {
        const TransportReport t = turb->solveTransport(U, phi, c.mGrad);
        rep.kIters = t.itK;
        rep.wIters = t.itOmega;
        rep.boundedOmega = t.boundedOmegaCells;
        rep.dTurb = std::max(t.dOmega, t.dK);
        rep.rTurb = std::max(t.rOmega, t.rK);
        if (!t.healthy || !std::isfinite(rep.dTurb) || !std::isfinite(rep.rTurb)) {  // main.cpp:106
            rep.failed = 1;
            return rep;
        }
        rep.turbLinear = t.statusK == Converged && t.statusOmega == Converged;
        turbulenceConverged = rep.turbLinear && rep.dTurb <= c.kw.tolerance && rep.rTurb <= c.kw.tolerance;
        rep.turbConverged = turbulenceConverged;
    }
    // Convergence (main.cpp:112-126; fvm_execution.cpp:91-177).
    {
        double d2 = 0, c2 = 0;
        for (Index i = 0; i < m.nc; ++i) { d2 += norm2(U.v[i] - Uprev[i]); c2 += norm2(U.v[i]); }
        rep.dU = std::sqrt(d2) / std::max(std::sqrt(c2), 1e-30);
        double v2 = 0, r2 = 0;
        for (Index i = 0; i < m.nc; ++i) { v2 += pPrime.v[i] * pPrime.v[i]; r2 += p.v[i] * p.v[i]; }
        rep.dP = std::sqrt(v2) / std::max(std::sqrt(r2), 1e-30);
        double l1 = 0, scale = 0;
        for (Index cidx = 0; cidx < m.nc; ++cidx) {
            double imb = 0;
            for (Index k = m.cfo[cidx]; k < m.cfo[cidx + 1]; ++k) {
                const Index f = m.cfi[k];
                const double out = m.own[f] == cidx ? phi[f] : -phi[f];
                imb += out;
                scale += std::abs(out);
            }
            l1 += std::abs(imb);
        }
        rep.mass = l1 / std::max(scale, 1e-30);
    }
    const bool linearConverged = rep.momStatus == Converged && pConverged;
    rep.converged = linearConverged && turbulenceConverged && rep.rU <= c.tolMom && rep.dU <= c.tolU &&
                    rep.dP <= c.tolP && rep.mass <= c.tolMass;
    // main.cpp:115-116
    if (!std::isfinite(rep.dU) || !std::isfinite(rep.dP) || !std::isfinite(rep.mass)) rep.failed = 1;
    return rep;
}

}  // namespace orc
