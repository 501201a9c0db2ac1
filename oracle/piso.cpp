// Oracle transient PISO time step: restates src/physics/piso/main.cpp:19-53
// (temporalFluxCorrection) and 161-300 (one physical time step) for laminar
// flow, with the equ:: semantics it uses (src/discretization/
// procedural_equation.cpp: ddt 198-239, apply 281-292, relax 298-305,
// faceFlux 335-352) and the diagnostics of fvm_execution.cpp:91-205.
// Test infrastructure only.
#include "piso.hpp"

#include <algorithm>

namespace orc {

TimeWeights timeWeights(int method, int levels, double dt, double previousDt) {
    if (!(dt > 0) || !std::isfinite(dt)) throw std::invalid_argument("equ::ddt requires positive finite dt");
    TimeWeights w{1 / dt, 1 / dt, 0.0};
    if (method == TimeBdf2 && levels >= 2) {
        double olddt = previousDt;
        if (olddt == 0) olddt = dt;
        if (!(olddt > 0) || !std::isfinite(olddt)) throw std::invalid_argument("invalid previous dt");
        const double r = dt / olddt;
        w.a0 = (1 + 2 * r) / ((1 + r) * dt);
        w.a1 = (1 + r) / dt;
        w.a2 = -r * r / ((1 + r) * dt);
    }
    return w;
}

void temporalFluxWeights(int method, int levels, double dt, double previousDt, double& previousWeight,
                         double& olderWeight) {
    previousWeight = 1.0 / dt;
    olderWeight = 0.0;
    if (method == TimeBdf2 && levels >= 2) {
        const double ratio = dt / previousDt;
        previousWeight = (1.0 + ratio) / dt;
        olderWeight = -ratio * ratio / ((1.0 + ratio) * dt);
    }
}

Piso::Piso(const Mesh& mesh, const PisoConfig& cfg) : m(mesh), c(cfg) {
    if (!(c.alphaU > 0 && c.alphaU <= 1)) throw std::invalid_argument("relaxation must be in (0,1]");
    if (c.timeMethod != TimeEuler && c.timeMethod != TimeBdf2)
        throw std::invalid_argument("PISO requires a transient time scheme");
    if (c.maxIterations < 1 || c.nCorrectors < 1 || c.nCorrectors > 100 || c.nonOrthCorr < 0 || c.nonOrthCorr > 20)
        throw std::invalid_argument("PISO iteration counts out of range");
    if (c.maxCorrectors == 0) c.maxCorrectors = c.nCorrectors;
    if (c.maxCorrectors < c.nCorrectors || c.maxCorrectors > 100)
        throw std::invalid_argument("maxCorrectors must lie in [nCorrectors, 100]");
    if (!(c.couplingTol > 0)) throw std::invalid_argument("couplingTolerance must be positive");
    if (c.potentialInit != 0 && c.potentialInit != 1) throw std::invalid_argument("initializePotentialFlow must be 0 or 1");
    if (c.initialNonOrthCorr < 0) c.initialNonOrthCorr = c.nonOrthCorr;
    if (c.initialNonOrthCorr > 20) throw std::invalid_argument("PISO iteration counts out of range");
    U = VField(m); p = SField(m); muEff = SField(m);
    muEff.v.assign(m.nc, c.mu);  // model.cpp:58-63
}

void Piso::setBoundaries(const std::vector<BC<V3>>& ubc, const std::vector<BC<double>>& pbc) {
    U.bc = ubc; p.bc = pbc;
}

void Piso::enableKOmega(const KOmegaConfig& kw, const std::vector<BC<double>>& kbc,
                        const std::vector<BC<double>>& wbc, const PisoTurbulence& t) {
    // main.cpp:99-120
    if (t.coupling != 0 && t.coupling != 1) throw std::invalid_argument("turbulenceCoupling must be iterated or segregated");
    const bool segregated = t.coupling == 1;
    if (segregated && c.maxIterations != 1)
        throw std::invalid_argument("segregated turbulence requires one PISO momentum pass");
    if (segregated && kw.relaxation != 1.0)
        throw std::invalid_argument("single-pass transient turbulence requires turbulenceRelaxation 1");
    if (t.maxIterations < 1) throw std::invalid_argument("turbulenceMaxIterations must be positive");
    tc = t;
    if (tc.correctors <= 0) tc.correctors = c.maxIterations == 1 && !segregated ? t.maxIterations : 1;
    if (tc.correctors > tc.maxIterations) throw std::invalid_argument("turbulenceCorrectors out of range");
    if (segregated && tc.correctors != 1)
        throw std::invalid_argument("segregated turbulence requires one transport correction");
    turb = std::make_unique<KOmega>(m, kw, c.rho, c.mu);
    turb->setBoundaries(kbc, wbc);
    turb->timeMethod = c.timeMethod;
}

void Piso::start() {
    // The model exists before phi (main.cpp:71, 98): its first bounding and
    // viscosity see a zero face flux.
    if (turb) turb->start(std::vector<double>(m.nf, 0.0));
    // main.cpp:123: phi = flux(U, pressure options); histories start from the
    // current fields (main.cpp:157-158, history.h:7).
    flux(U, phi, c.pInterp, c.pGrad);
    initial = InitReport{};
    if (c.potentialInit) initializePotentialFlow();
    Uold.init(U);
    phiOld.init(phi);
    time = 0;
}

std::vector<double> Piso::temporalFluxCorrection() {  // main.cpp:19-53
    const double dt = Uold.dt;
    double wp, wo;
    temporalFluxWeights(c.timeMethod, Uold.levels, dt, Uold.previousDt, wp, wo);
    std::vector<double> fp, fo;
    flux(Uold.previous, fp, c.pInterp, c.pGrad);  // also sets the saved field's flux context
    if (wo != 0.0) flux(Uold.older, fo, c.pInterp, c.pGrad);
    std::vector<double> result(m.nf, 0.0);
    for (Index f = 0; f < m.nf; ++f) {
        if (m.boundary(f)) continue;
        const double e = phiOld.previous[f] - fp[f];
        double corr = wp * e;
        if (wo != 0.0) corr += wo * (phiOld.older[f] - fo[f]);
        const double fl = phiOld.previous[f];
        const double coupling = e == 0.0 ? 1.0 : fl == 0.0 ? 0.0 : std::max(0.0, 1.0 - std::abs(e) / std::abs(fl));
        result[f] += coupling * corr;
    }
    return result;
}

namespace {

double l2(const std::vector<V3>& v) {
    double s = 0;
    for (const V3& x : v) s += norm2(x);
    return std::sqrt(s);
}

void solveOn(const Mesh& m, LinearSystem& sys, const Ldu<double>& q, const SolverConfig& sc, std::vector<double>& x,
             SolveResult& res, std::vector<double>* hist) {
    Csr A;
    lduToCsr(m, q.diag, q.upper, q.lower, A);
    if (!sys.ready) sys.A = A;
    sys.update(A.diag, A.val);
    sys.prepare(sc);
    res = sys.solve(q.source, x, 0);
    if (hist) *hist = sys.hist;
}

}  // namespace

// main.cpp:127-150: the potential pPrime (homogeneous boundary conditions of
// p) solves laplacian(1, -1) with the source -div(phi), as many times as the
// initial non-orthogonal corrections ask (potentialEquation: its own linear
// system); then U -= grad(pPrime) and phi += faceFlux(potentialEquation,
// pPrime).  A solve that does not converge ends the run.
void Piso::initializePotentialFlow() {
    SField pp(m);
    pp.bc.resize(m.np);
    for (Index k = 0; k < m.np; ++k) {
        const int t = p.bc[k].type;
        pp.bc[k] = t == FixedValue ? BC<double>{FixedValue, 0.0}
                 : t == Symmetry ? BC<double>{Symmetry, 0.0} : BC<double>{ZeroGradient, 0.0};
    }
    const std::vector<double> one(m.nf, 1.0);
    std::vector<double> imbalance;
    divergenceFace(m, phi, imbalance);
    const int solves = c.pDiff == Orthogonal ? 1 : c.initialNonOrthCorr + 1;
    initial.solves = solves;
    LinearSystem potSys;
    std::vector<double> correction;
    SField snapshot(m);
    for (int k = 0; k < solves; ++k) {
        Ldu<double> peq; peq.init(m);
        Ldu<double> contrib; contrib.init(m);
        if (c.pDiff != Orthogonal) {
            std::vector<V3> g;
            gradient(pp, g, c.pGrad);
            addDiffusion<double>(contrib, one, pp, c.pGrad, c.pDiff, &g);
            std::vector<double> full, orthF;  // procedural_equation.cpp:77-87
            diffusionFlux(one, pp, g, c.pDiff, full);
            diffusionFlux(one, pp, g, Orthogonal, orthF);
            correction.assign(m.nf, 0.0);
            for (Index f = 0; f < m.nf; ++f) correction[f] = full[f] + (-1.0) * orthF[f];
        } else {
            addDiffusion<double>(contrib, one, pp, c.pGrad, c.pDiff, nullptr);
            correction.assign(m.nf, 0.0);
        }
        snapshot = pp;
        accumulate(peq, contrib, -(-1.0));
        for (Index i = 0; i < m.nc; ++i) peq.source[i] += (1.0 * m.V[i]) * (-1.0 * imbalance[i]);
        referenceIfUnanchored(m, peq, 0.0);
        SolveResult r;
        solveOn(m, potSys, peq, c.pre, pp.v, r, nullptr);
        initial.pIters += r.iterations;
        if (r.status != Converged) {
            initial.status = StepNumericalFailure;
            initial.done = 1;
            return;
        }
    }
    std::vector<V3> g;
    gradient(pp, g, c.pGrad);
    for (Index i = 0; i < m.nc; ++i) U.v[i] = U.v[i] - g[i];
    {
        SField value = snapshot;
        value.v = pp.v;
        std::vector<V3> zero(m.nc, V3{});
        std::vector<double> inc;
        diffusionFlux(one, value, zero, Orthogonal, inc);
        for (Index f = 0; f < m.nf; ++f) inc[f] += 1.0 * correction[f];
        for (Index f = 0; f < m.nf; ++f) phi[f] = phi[f] + (0.0 + -1.0 * inc[f]);
    }
    U.fluxCtx = phi;
    double l1 = 0, sc = 0, umax = 0;
    for (Index cidx = 0; cidx < m.nc; ++cidx) {
        double imb = 0;
        for (Index k = m.cfo[cidx]; k < m.cfo[cidx + 1]; ++k) {
            const Index f = m.cfi[k];
            const double out = m.own[f] == cidx ? phi[f] : -phi[f];
            imb += out;
            sc += std::abs(out);
        }
        l1 += std::abs(imb);
        umax = std::max(umax, std::sqrt(dot(U.v[cidx], U.v[cidx])));
    }
    initial.mass = l1 / std::max(sc, 1e-30);
    initial.uMax = umax;
    initial.done = 1;
}

StepReport Piso::step(double dt) {
    StepReport rep{};
    rep.dt = dt;
    if (!(dt > 0) || !std::isfinite(dt)) throw std::invalid_argument("history requires positive finite dt");
    if (initial.status != StepCompleted) return rep;  // runPiso ended in the initial projection
    time += dt;
    // main.cpp:162-167
    Uold.save(U, dt);
    phiOld.save(phi, dt);
    if (turb) turb->saveOld(dt);
    std::vector<double> tf = temporalFluxCorrection();
    for (auto& v : tf) v = c.rho * v;
    lastTemporalFlux = tf;
    const int pressureSolves = c.pDiff == Orthogonal ? 1 : c.nonOrthCorr + 1;
    bool converged = false;
    for (int iter = 0; iter < c.maxIterations; ++iter) {
        rep.passes = iter + 1;
        const std::vector<V3> Uiter = U.v;
        // Momentum predictor (main.cpp:174-181).
        Ldu<V3> mom; mom.init(m);
        {
            const TimeWeights w = timeWeights(c.timeMethod, Uold.levels, Uold.dt, Uold.previousDt);
            for (Index i = 0; i < m.nc; ++i) {  // procedural_equation.cpp:213-222
                const double cw = c.rho * m.V[i];
                mom.diag[i] += cw * w.a0;
                mom.source[i] += cw * w.a1 * Uold.previous.v[i];
                if (w.a2 != 0) mom.source[i] += cw * w.a2 * Uold.older.v[i];
            }
        }
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
        for (Index i = 0; i < m.nc; ++i) mom.source[i] += (1.0 * m.V[i]) * (-1.0 * gradP[i]);
        if (turb) {  // main.cpp:179-182: the explicit remainder of the eddy-viscosity stress
            std::vector<V3> divT;
            turb->stressDivergence(U, c.mInterp, c.mGrad, divT);
            for (Index i = 0; i < m.nc; ++i) mom.source[i] += (1.0 * m.V[i]) * divT[i];
        }
        rep.rU = relativeResidual(m, mom, U.v);  // main.cpp:184
        if (!std::isfinite(rep.rU)) { rep.status = StepNumericalFailure; return rep; }
        for (Index i = 0; i < m.nc; ++i) {  // relax, no rescaling (main.cpp:186)
            const double extra = mom.diag[i] * (1 / c.alphaU - 1);
            mom.diag[i] += extra;
            mom.source[i] += extra * Uiter[i];
        }
        const std::vector<double> aP = mom.diag;
        const SField rAU = calculatedRatio(m, m.V, aP);
        std::vector<double> inverseAP(m.nc);
        for (Index i = 0; i < m.nc; ++i) inverseAP[i] = 1.0 / aP[i];
        std::vector<double> rAUf;
        interpolate(rAU, rAUf, c.pCoefInterp, c.pCoefGrad);
        std::vector<V3> npr(m.nc);  // main.cpp:193
        for (Index i = 0; i < m.nc; ++i) npr[i] = mom.source[i] + m.V[i] * gradP[i];
        // Momentum solve, three right-hand sides (procedural_equation.cpp:357-370).
        int momStatus = Converged;
        double linU = 0;
        {
            Csr A;
            lduToCsr(m, mom.diag, mom.upper, mom.lower, A);
            if (!momSys.ready) momSys.A = A;
            momSys.update(A.diag, A.val);
            momSys.prepare(c.mom);
            for (int comp = 0; comp < 3; ++comp) {
                std::vector<double> b(m.nc), x(m.nc);
                for (Index i = 0; i < m.nc; ++i) { b[i] = mom.source[i][comp]; x[i] = U.v[i][comp]; }
                const SolveResult r = momSys.solve(b, x, comp);
                momHist[comp] = momSys.hist;
                for (Index i = 0; i < m.nc; ++i) U.v[i][comp] = x[i];
                rep.momIters[comp] = r.iterations;
                if (r.status == NumericalFailure) momStatus = NumericalFailure;
                else if (r.status != Converged && momStatus != NumericalFailure) momStatus = MaxIterations;
                linU = std::max(linU, r.relative);
            }
        }
        rep.linU = linU;
        if (momStatus == NumericalFailure) { rep.status = StepNumericalFailure; return rep; }
        // Correctors (main.cpp:202-249).
        bool pressureConverged = true;
        double pLin = 0.0, dP = 0.0, first = 0.0, coupling = 0.0;
        int used = 0;
        rep.pIters = 0;
        for (int corrector = 0; corrector < c.maxCorrectors; ++corrector) {
            std::vector<V3> AU;
            applyLdu(m, mom, U.v, AU);
            VField HbyA = U;  // U's boundary conditions
            for (Index i = 0; i < m.nc; ++i) HbyA.v[i] = inverseAP[i] * ((npr[i] - AU[i]) + aP[i] * U.v[i]);
            HbyA.fluxCtx = phi;
            const std::vector<double> previousPressure = p.v;
            std::vector<double> phiH;
            flux(HbyA, phiH, c.pInterp, c.pGrad);
            for (Index f = 0; f < m.nf; ++f)
                if (!m.boundary(f)) phiH[f] += rAUf[f] * tf[f];
            std::vector<double> imbalance;
            divergenceFace(m, phiH, imbalance);
            Ldu<double> peq;
            std::vector<double> correction;
            SField snapshot(m);
            for (int k = 0; k < pressureSolves; ++k) {  // main.cpp:216-225
                peq.init(m);
                Ldu<double> contrib; contrib.init(m);
                if (c.pDiff != Orthogonal) {
                    std::vector<V3> g;
                    gradient(p, g, c.pGrad);
                    addDiffusion<double>(contrib, rAUf, p, c.pGrad, c.pDiff, &g);
                    std::vector<double> full, orthF;  // procedural_equation.cpp:77-87
                    diffusionFlux(rAUf, p, g, c.pDiff, full);
                    diffusionFlux(rAUf, p, g, Orthogonal, orthF);
                    correction.assign(m.nf, 0.0);
                    for (Index f = 0; f < m.nf; ++f) correction[f] = full[f] + (-1.0) * orthF[f];
                } else {
                    addDiffusion<double>(contrib, rAUf, p, c.pGrad, c.pDiff, nullptr);
                    correction.assign(m.nf, 0.0);
                }
                snapshot = p;
                accumulate(peq, contrib, -(-1.0));
                for (Index i = 0; i < m.nc; ++i) peq.source[i] += (1.0 * m.V[i]) * (-1.0 * imbalance[i]);
                referenceIfUnanchored(m, peq, 0.0);
                SolveResult r;
                solveOn(m, pSys, peq, c.pre, p.v, r, &pHist);
                rep.pIters += r.iterations;
                if (r.status == NumericalFailure) { rep.status = StepNumericalFailure; return rep; }
                pressureConverged = pressureConverged && r.status == Converged;
                pLin = std::max(pLin, r.relative);
            }
            // phi = phiH + faceFlux(pEq, p) (main.cpp:228; procedural_equation.cpp:335-352)
            {
                SField value = snapshot;
                value.v = p.v;
                std::vector<V3> zero(m.nc, V3{});
                std::vector<double> inc;
                diffusionFlux(rAUf, value, zero, Orthogonal, inc);
                for (Index f = 0; f < m.nf; ++f) inc[f] += 1.0 * correction[f];
                phi.assign(m.nf, 0.0);
                for (Index f = 0; f < m.nf; ++f) phi[f] = phiH[f] + (0.0 + -1.0 * inc[f]);
            }
            std::vector<V3> gp;
            gradient(p, gp, c.pGrad);
            U.v = HbyA.v;  // main.cpp:230-232
            for (Index i = 0; i < m.nc; ++i) U.v[i] += -1.0 * (rAU.v[i] * gp[i]);
            U.fluxCtx = phi;
            {
                double v2 = 0, r2 = 0;
                for (Index i = 0; i < m.nc; ++i) {
                    const double d = p.v[i] - previousPressure[i];
                    v2 += d * d; r2 += p.v[i] * p.v[i];
                }
                dP = std::max(dP, std::sqrt(v2) / std::max(std::sqrt(r2), 1e-30));
            }
            // Coupling defect of the frozen momentum equation (main.cpp:234-239).
            applyLdu(m, mom, U.v, AU);
            std::vector<V3> ps(m.nc), defect(m.nc);
            for (Index i = 0; i < m.nc; ++i) {
                ps[i] = m.V[i] * gp[i];
                defect[i] = (npr[i] - AU[i]) - ps[i];
            }
            const double n0 = l2(npr), n1 = l2(AU), n2 = l2(ps), n3 = l2(defect);
            const double scale = n0 + n1 + n2;
            coupling = scale > 0.0 ? n3 / scale : n3;
            if (corrector == 0) first = coupling;
            used = corrector + 1;
            if (used >= c.nCorrectors && coupling <= c.couplingTol) break;
        }
        rep.correctors = used;
        rep.dP = dP;
        rep.rCouplingFirst = first;
        rep.rCoupling = coupling;
        rep.linP = pLin;
        // Turbulence transport at this time level (main.cpp:251-266).
        bool turbulenceConverged = true, turbulenceLinear = true;
        if (turb) {
            rep.kIters = rep.wIters = 0;
            for (int correction = 0; correction < tc.correctors; ++correction) {
                const TransportReport t = turb->solveTransport(U, phi, c.mGrad);
                rep.turbCorrections = correction + 1;
                rep.kIters += t.itK;
                rep.wIters += t.itOmega;
                rep.dTurb = std::max(t.dOmega, t.dK);
                rep.rTurb = std::max(t.rOmega, t.rK);
                if (!t.healthy || !std::isfinite(rep.dTurb) || !std::isfinite(rep.rTurb)) {
                    rep.status = StepNumericalFailure;
                    return rep;
                }
                turbulenceLinear = t.statusK == Converged && t.statusOmega == Converged;
                turbulenceConverged = turbulenceLinear && rep.dTurb <= turb->c.tolerance && rep.rTurb <= turb->c.tolerance;
                if (!turbulenceLinear || turbulenceConverged) break;
            }
            rep.turbLinear = turbulenceLinear;
            rep.turbConverged = turbulenceConverged;
        }
        // Diagnostics (main.cpp:268-292; fvm_execution.cpp:91-205).
        {
            double d2 = 0, c2 = 0;
            for (Index i = 0; i < m.nc; ++i) { d2 += norm2(U.v[i] - Uiter[i]); c2 += norm2(U.v[i]); }
            rep.dU = std::sqrt(d2) / std::max(std::sqrt(c2), 1e-30);
            double l1 = 0, sc = 0, weighted = 0, volume = 0, coMax = 0;
            for (Index cidx = 0; cidx < m.nc; ++cidx) {
                double imb = 0, sum = 0;
                for (Index k = m.cfo[cidx]; k < m.cfo[cidx + 1]; ++k) {
                    const Index f = m.cfi[k];
                    const double out = m.own[f] == cidx ? phi[f] : -phi[f];
                    imb += out;
                    sc += std::abs(out);
                    sum += std::abs(phi[f]);
                }
                l1 += std::abs(imb);
                const double co = 0.5 * dt * sum / m.V[cidx];
                coMax = std::max(coMax, co);
                weighted += co * m.V[cidx];
                volume += m.V[cidx];
            }
            rep.mass = l1 / std::max(sc, 1e-30);
            rep.coMax = coMax;
            rep.coMean = weighted / volume;
        }
        if (!(std::isfinite(rep.rU) && std::isfinite(rep.dU) && std::isfinite(rep.dP) && std::isfinite(coupling) &&
              std::isfinite(rep.mass))) {
            rep.status = StepNumericalFailure;
            return rep;
        }
        const bool linearConverged = momStatus == Converged && pressureConverged && turbulenceLinear;
        rep.linearOk = linearConverged;
        const bool accepted = linearConverged && rep.mass <= c.tolMass && coupling <= c.couplingTol;
        const bool settled = rep.rU <= c.tolMom && rep.dU <= c.tolU && rep.dP <= c.tolP && turbulenceConverged;
        converged = accepted && (c.maxIterations == 1 ? (tc.coupling == 1 || turbulenceConverged) : settled);
        rep.accepted = accepted;
        rep.converged = converged;
        if (!accepted) { rep.status = StepNotConverged; return rep; }
        if (converged) break;
    }
    rep.status = converged ? StepCompleted : StepNotConverged;
    return rep;
}

}  // namespace orc
