// Steady SIMPLE outer iteration resident in HBM: the order of
// src/physics/simple/main.cpp:47-127 with every field, matrix and
// preconditioner kept on the device.  The host only makes the decisions the
// reference makes on the host (linear convergence, outer convergence).
#pragma once

#include "context.hpp"
#include "komega.hpp"
#include "perf.hpp"

namespace bsim {

struct SimpleCfg {
    double rho = 1.0, mu = 0.01;
    int mInterp = 0, mGrad = 0, mConv = 1, mDiff = 0, mCoefInterp = 0, mCoefGrad = 0;
    int pInterp = 0, pGrad = 0, pDiff = 0, pCoefInterp = 0, pCoefGrad = 0;
    double alphaU = 0.7, alphaP = 0.3;
    int nonOrthCorr = 1;
    double tolMass = 1e-8, tolU = 1e-7, tolMom = 1e-6, tolP = 1e-6;
    SolverCfg mom, pre;
};

struct IterRep {
    double rU = 0, dU = 0, dP = 0, mass = 0, linP = 0;
    int momIters[3] = {0, 0, 0};
    int pIters = 0, momStatus = 0, pStatus = 0, converged = 0, referenced = 0;
    double dTurb = 0, rTurb = 0, boundedOmega = 0;
    int kIters = 0, wIters = 0, turbLinear = 1, turbConverged = 1;
    int failed = 0;  // runSimple's numericalFailure: the iteration stopped where it was detected
};

// Per-phase wall time (seconds) accumulated over iterations.
struct SimpleTiming {
    double assembly = 0, momentumSolve = 0, rhieChow = 0, pressureAssembly = 0, pressureSolve = 0, correction = 0,
           convergence = 0, turbulence = 0;
    int iterations = 0;
};

class SimpleSolver {
public:
    SimpleSolver(Context& ctx, const SimpleCfg& cfg, const int* ubct, const double* ubcv, const int* pbct,
                 const double* pbcv);
    // Wilcox k-omega (src/physics/RANS/model.cpp:35-40); before start().
    void enableKOmega(const KOmegaCfg& kw, const int* kbct, const double* kbcv, const int* wbct, const double* wbcv);
    KOmegaDevice* turbulence() { return turb.get(); }
    void setFields(const double* Ucaller, const double* pcaller);
    void getFields(double* Ucaller, double* pcaller, double* phicaller);
    void start();
    IterRep iterate();
    std::vector<double> momHist[3], pHist;
    SimpleTiming timing;
    RunClock run;
    DeviceSystem& momentumSystem() { return *momSys; }
    DeviceSystem& pressureSystem() { return *pSys; }

private:
    FieldArg fieldU(const double* ctx) const;
    FieldArg fieldP() const;
    FieldArg fieldPP() const;
    FieldArg fieldRAU() const;
    double sync();
    Context& C;
    SimpleCfg c;
    int refCell = 0;
    DBuf<int> uType, pType, ppType, zgType;
    DBuf<double> uVal, pVal, ppVal, zgVal;
    DBuf<double> U, p, phi, phiH, Uprev, gradU, gradP, gradPP, work9, rAU, rAUtr, gtr, X, Xtr, Xf, gradX, rAUf, sn,
        kfMom, mu, kfP, divPhiH, pPrime, corr, src, psrc, ctxSeed, red;
    DBuf<double> faceWork;  // face-ordered assembly terms: 6 per interior face
    DBuf<double> bcomp[3], xcomp[3];  // momentum components, one array each
    std::unique_ptr<DeviceSystem> momSys, pSys;
    std::unique_ptr<KOmegaDevice> turb;
    bool started = false;
};

}  // namespace bsim
