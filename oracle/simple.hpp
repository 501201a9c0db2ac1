// Oracle SIMPLE driver declarations.  Test infrastructure only.
#pragma once

#include "internal.hpp"
#include "rans.hpp"

#include <memory>

namespace orc {

struct SimpleConfig {
    double rho = 1.0, mu = 0.01;
    int mInterp = Linear, mGrad = GreenGauss, mConv = LinearUpwind, mDiff = Orthogonal;
    int mCoefInterp = Linear, mCoefGrad = GreenGauss;
    int pInterp = Linear, pGrad = GreenGauss, pDiff = Orthogonal;
    int pCoefInterp = Linear, pCoefGrad = GreenGauss;
    double alphaU = 0.7, alphaP = 0.3;
    int nonOrthCorr = 1;
    double tolMass = 1e-8, tolU = 1e-7, tolMom = 1e-6, tolP = 1e-6;
    SolverConfig mom, pre;
    int turbulence = 0;  // 0 laminar, 1 Wilcox k-omega
    KOmegaConfig kw;
};

struct IterReport {
    double rU = 0, dU = 0, dP = 0, mass = 0, linP = 0;
    int momIters[3] = {0, 0, 0};
    int pIters = 0;
    int momStatus = 0, pStatus = 0;
    int converged = 0;
    int referenced = 0;
    double dTurb = 0, rTurb = 0, boundedOmega = 0;
    int kIters = 0, wIters = 0, turbLinear = 1, turbConverged = 1;
    int failed = 0;  // runSimple's numericalFailure (main.cpp:61-62, 67, 88, 106, 115-116)
};

template <class T> void accumulate(Ldu<T>& a, const Ldu<T>& b, double f);
template <class T> void applyLdu(const Mesh& m, const Ldu<T>& q, const std::vector<T>& x, std::vector<T>& y);
SField calculatedRatio(const Mesh& m, const std::vector<double>& num, const std::vector<double>& den);
template <class T> double relativeResidual(const Mesh& m, const Ldu<T>& q, const std::vector<T>& x);
bool referenceIfUnanchored(const Mesh& m, Ldu<double>& q, double value);

struct Simple {
    const Mesh& m;
    SimpleConfig c;
    VField U;
    SField p, muEff, pPrime;
    std::vector<double> phi;
    LinearSystem momSys, pSys;
    std::vector<double> momHist[3], pHist;
    std::unique_ptr<KOmega> turb;
    Simple(const Mesh& mesh, const SimpleConfig& cfg);
    void setBoundaries(const std::vector<BC<V3>>& ubc, const std::vector<BC<double>>& pbc);
    void start();
    IterReport iterate();
    void solveSystem(LinearSystem& sys, const Ldu<double>& q, const SolverConfig& sc, std::vector<double>& x,
                     SolveResult& res, std::vector<double>* hist);
};

}  // namespace orc
