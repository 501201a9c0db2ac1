// Transient PISO time step resident in HBM: the order of
// src/physics/piso/main.cpp:161-300 (laminar) with every field, history,
// matrix and preconditioner kept on the device.  The host makes only the
// decisions the reference makes on the host: linear convergence, the
// corrector count and step acceptance.
#pragma once

#include "context.hpp"
#include "komega.hpp"
#include "perf.hpp"

#include <limits>

namespace bsim {

enum : int { kEuler = 1, kBdf2 = 2 };
enum : int { kStepCompleted = 0, kStepNotConverged = 1, kStepNumericalFailure = 2 };

struct PisoCfg {
    double rho = 1.0, mu = 0.01;
    int mInterp = 0, mGrad = 0, mConv = 1, mDiff = 0, mCoefInterp = 0, mCoefGrad = 0;
    int pInterp = 0, pGrad = 0, pDiff = 0, pCoefInterp = 0, pCoefGrad = 0;
    int timeMethod = kBdf2;
    int maxIterations = 1, nCorrectors = 2, maxCorrectors = 0, nonOrthCorr = 1;
    double alphaU = 1.0, couplingTol = std::numeric_limits<double>::max();
    double tolMass = 1e-8, tolU = 1e-7, tolMom = 1e-6, tolP = 1e-6;
    SolverCfg mom, pre;
    // initializePotentialFlow and initialNonOrthogonalCorrections
    // (main.cpp:104-105, 127-150); a negative count takes nonOrthCorr.
    int potentialInit = 0, initialNonOrthCorr = -1;
};

// The turbulence controls of runPiso (main.cpp:98-120): coupling 0 iterated,
// 1 segregated; correctors <= 0 takes the reference's default
// (turbulenceMaxIterations for iterated coupling with one momentum pass,
// else 1).
struct PisoTurbCfg {
    int coupling = 0, correctors = 0, maxIterations = 1000;
};

// What runPiso records after the initial projection (main.cpp:147-149), with
// its Krylov iterations and whether every solve converged.
struct InitRep {
    int done = 0, solves = 0, pIters = 0, status = kStepCompleted;
    double mass = 0, uMax = 0;
};

struct StepRep {
    double dt = 0, coMax = 0, coMean = 0, mass = 0, dU = 0, rU = 0, dP = 0;
    double rCouplingFirst = 0, rCoupling = 0, linU = 0, linP = 0;
    int passes = 0, correctors = 0;
    int momIters[3] = {0, 0, 0};
    int pIters = 0, accepted = 0, converged = 0, linearOk = 0, status = kStepNumericalFailure;
    // Turbulence transport at the time level (main.cpp:251-266): the last
    // correction's dTurb and rTurb, the Krylov iterations of all corrections.
    double dTurb = 0, rTurb = 0;
    int kIters = 0, wIters = 0, turbLinear = 1, turbConverged = 1, turbCorrections = 0;
};

// Per-phase wall time (seconds) accumulated over steps.
struct PisoTiming {
    double history = 0, assembly = 0, momentumSolve = 0, corrector = 0, pressureAssembly = 0, pressureSolve = 0,
           diagnostics = 0, turbulence = 0;
    int steps = 0;
};

class PisoSolver {
public:
    PisoSolver(Context& ctx, const PisoCfg& cfg, const int* ubct, const double* ubcv, const int* pbct,
               const double* pbcv);
    // Wilcox k-omega in transient form (main.cpp:98-121); before start().
    void enableKOmega(const KOmegaCfg& kw, const int* kbct, const double* kbcv, const int* wbct, const double* wbcv,
                      const PisoTurbCfg& t);
    KOmegaDevice* turbulence() { return turb.get(); }
    void setFields(const double* Ucaller, const double* pcaller);
    void getFields(double* Ucaller, double* pcaller, double* phicaller);
    void temporalFlux(double* caller);
    void start();
    InitRep initial;  // the potential-flow projection run by start(), when configured
    StepRep step(double dt);
    std::vector<double> momHist[3], pHist;
    PisoTiming timing;
    RunClock run;
    DeviceSystem& momentumSystem() { return *momSys; }
    DeviceSystem& pressureSystem() { return *pSys; }

private:
    FieldArg fieldU(const double* ctx) const { return {U.p, uType.p, uVal.p, ctx, nullptr}; }
    FieldArg fieldP() const { return {p.p, pType.p, pVal.p, nullptr, nullptr}; }
    FieldArg fieldRAU() const { return {rAU.p, zgType.p, zgVal.p, nullptr, rAUtr.p}; }
    void computeTemporalFlux();
    void initializePotentialFlow();
    FieldArg fieldPP() const { return {pPrime.p, ppType.p, ppVal.p, nullptr, nullptr}; }
    Context& C;
    PisoCfg c;
    int refCell = 0;
    bool inletOutlet = false;
    DBuf<int> uType, pType, zgType, ppType;
    DBuf<double> uVal, pVal, zgVal, ppVal, pPrime;
    DBuf<double> U, p, phi, Uiter, HbyA, gradU, gradP, work9, rAU, rAUtr, invAP, rAUf, npr, kfMom, mu, imbalance,
        pOld, corr, full, orthF, src, psrc, tf, fluxTmp, red;
    DBuf<double> faceWork;  // face-ordered assembly terms: 6 per interior face
    // Time levels (history.h:6-27): slot `cur` is the previous level, 1-cur the older.
    DBuf<double> Uh[2], phih[2], fluxh[2];  // fluxh: flux of the saved velocity
    int cur = 0, levels = 0;
    double dtCur = 0, dtPrev = 0;
    DBuf<double> bcomp[3], xcomp[3];
    std::unique_ptr<DeviceSystem> momSys, pSys;
    std::unique_ptr<KOmegaDevice> turb;
    PisoTurbCfg tc;
    bool started = false, stepped = false;
};

}  // namespace bsim
