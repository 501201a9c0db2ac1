// Oracle transient PISO declarations.  Test infrastructure only.
#pragma once

#include <limits>

#include "rans.hpp"
#include "simple.hpp"

#include <memory>

namespace orc {

enum TimeMethodId : int { TimeEuler = 1, TimeBdf2 = 2 };
enum StepStatus : int { StepCompleted = 0, StepNotConverged = 1, StepNumericalFailure = 2 };

struct PisoConfig {
    double rho = 1.0, mu = 0.01;
    int mInterp = Linear, mGrad = GreenGauss, mConv = LinearUpwind, mDiff = Orthogonal;
    int mCoefInterp = Linear, mCoefGrad = GreenGauss;
    int pInterp = Linear, pGrad = GreenGauss, pDiff = Orthogonal;
    int pCoefInterp = Linear, pCoefGrad = GreenGauss;
    int timeMethod = TimeBdf2;
    int maxIterations = 1, nCorrectors = 2, maxCorrectors = 0;  // 0: equal to nCorrectors
    int nonOrthCorr = 1;
    double alphaU = 1.0;
    double couplingTol = std::numeric_limits<double>::max();
    double tolMass = 1e-8, tolU = 1e-7, tolMom = 1e-6, tolP = 1e-6;
    SolverConfig mom, pre;
    int potentialInit = 0, initialNonOrthCorr = -1;  // main.cpp:104-105 (negative: nonOrthCorr)
};

// The turbulence controls of runPiso (main.cpp:98-120): coupling 0 iterated,
// 1 segregated; correctors <= 0 takes the default (turbulenceMaxIterations
// for iterated coupling with one momentum pass, else 1).
struct PisoTurbulence {
    int coupling = 0, correctors = 0, maxIterations = 1000;
};

// What runPiso records after the initial projection (main.cpp:147-149).
struct InitReport {
    int done = 0, solves = 0, pIters = 0, status = StepCompleted;
    double mass = 0, uMax = 0;
};

struct StepReport {
    double dt = 0, coMax = 0, coMean = 0, mass = 0, dU = 0, rU = 0, dP = 0;
    double rCouplingFirst = 0, rCoupling = 0, linU = 0, linP = 0;
    int passes = 0, correctors = 0;
    int momIters[3] = {0, 0, 0};
    int pIters = 0, accepted = 0, converged = 0, linearOk = 0, status = StepNumericalFailure;
    double dTurb = 0, rTurb = 0;
    int kIters = 0, wIters = 0, turbLinear = 1, turbConverged = 1, turbCorrections = 0;
};

// Explicit time levels (include/babelsim/history.h:6-27).
template <class F> struct History {
    F previous, older;
    double dt = 0, previousDt = 0;
    int levels = 0;
    void init(const F& f) { previous = f; older = f; dt = previousDt = 0; levels = 0; }
    void save(const F& f, double step) {
        older = previous;
        previous = f;
        previousDt = dt;
        dt = step;
        ++levels;
    }
};

// ddt time-term weights (procedural_equation.cpp:198-224), Euler when fewer
// than two levels are saved (231-237).
struct TimeWeights { double a0, a1, a2; };
TimeWeights timeWeights(int method, int levels, double dt, double previousDt);
// Old-time flux weights of the ddtPhiCorr term (piso/main.cpp:28-33).
void temporalFluxWeights(int method, int levels, double dt, double previousDt, double& previousWeight,
                         double& olderWeight);

struct Piso {
    const Mesh& m;
    PisoConfig c;
    VField U;
    SField p, muEff;
    std::vector<double> phi;
    History<VField> Uold;
    History<std::vector<double>> phiOld;
    LinearSystem momSys, pSys;
    std::vector<double> momHist[3], pHist;
    std::vector<double> lastTemporalFlux;  // rho * temporalFluxCorrection of the last step
    double time = 0;
    Piso(const Mesh& mesh, const PisoConfig& cfg);
    void setBoundaries(const std::vector<BC<V3>>& ubc, const std::vector<BC<double>>& pbc);
    void start();
    void initializePotentialFlow();
    // Wilcox k-omega (main.cpp:98-121); before start().
    void enableKOmega(const KOmegaConfig& kw, const std::vector<BC<double>>& kbc, const std::vector<BC<double>>& wbc,
                      const PisoTurbulence& t);
    std::unique_ptr<KOmega> turb;
    PisoTurbulence tc;
    InitReport initial;
    StepReport step(double dt);
    std::vector<double> temporalFluxCorrection();
};

}  // namespace orc
