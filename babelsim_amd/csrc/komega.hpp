// Wilcox (1988) k-omega on the device (src/physics/RANS/k_omega.cpp:25-157,
// model.cpp:67-71): eddy viscosity with calculated boundary traces, the
// explicit deviatoric stress remainder of the momentum equation, and the
// segregated omega-then-k transport solves with bounding.  Steady form for the
// SIMPLE driver; the PISO driver adds the k and omega histories and their ddt
// terms (k_omega.cpp:59-62, 82, 96).
#pragma once

#include "context.hpp"

namespace bsim {

struct TransportOpts {
    int interp = 0, grad = 0, conv = 1, diff = 0, coefInterp = 0, coefGrad = 0;
};

struct KOmegaCfg {
    double betaStar = 0.09, beta = 3.0 / 40.0, gamma = 5.0 / 9.0, sigmaK = 0.5, sigmaOmega = 0.5;
    double kMin = 1e-12, omegaMin = 1e-12, viscosityRatioMax = 1e5, relaxation = 0.7, tolerance = 1e-6;
    TransportOpts kOpts, wOpts;
    SolverCfg kSolver, wSolver;
};

struct TransportRep {
    double rOmega = 0, rK = 0, dOmega = 0, dK = 0, boundedOmega = 0;
    int itOmega = 0, itK = 0, statusOmega = 0, statusK = 0;
    bool healthy = true;
};

class KOmegaDevice {
public:
    KOmegaDevice(Context& ctx, const KOmegaCfg& cfg, double rho, double mu, const int* kbct, const double* kbcv,
                 const int* wbct, const double* wbcv);
    void setFields(const double* kCaller, const double* wCaller);
    void getFields(double* kCaller, double* wCaller, double* mutCaller);
    // The model is built before phi exists: the first bounding and viscosity
    // see a zero face flux (k_omega.cpp:50-53).
    void start();
    const double* muEff() const { return muEffC.p; }
    const double* muEffTrace() const { return muEffT.p; }
    // src += V div(muEff (grad(U)^T - 2/3 tr(grad U) I)); gradU is grad(U) of
    // the momentum gradient method, U carries its flux context.
    void addStressSource(FieldArg U, const double* gradU, int mInterp, int mGrad, double* src);
    TransportRep solveTransport(FieldArg U, const double* phi, int mGrad);
    double tolerance() const { return c.tolerance; }
    double relaxation() const { return c.relaxation; }
    // Transient form: timeMethod 1 Euler or 2 BDF2 (Euler below two saved levels).
    void setTransient(int timeMethod);
    void saveOld(double dt);  // history.h:6-27
    std::vector<double> kHist, wHist;

private:
    FieldArg fieldK() const { return {k.p, kType.p, kVal.p, ctxK.p, nullptr}; }
    FieldArg fieldW() const { return {w.p, wType.p, wVal.p, ctxW.p, nullptr}; }
    void updateViscosity();
    double boundOmega();
    struct TimeTerm {
        double a0 = 0, a1 = 0, a2 = 0;
        const double *prev = nullptr, *older = nullptr;
    };
    void assemble(DeviceSystem& sys, FieldArg x, double* ctx, const double* phi, const TransportOpts& o,
                  const double* diffusivity, const double* diffusivityTrace, const double* rate, const double* source,
                  const TimeTerm* tt);
    int timeMethod = 0;
    DBuf<double> kh[2], wh[2];  // slot hcur: previous level, 1 - hcur: older
    int hcur = 0, hlevels = 0;
    double hdt = 0, hdtPrev = 0;
    double residual(DeviceSystem& sys, const double* x);
    Context& C;
    KOmegaCfg c;
    double rho, mu;
    DBuf<int> kType, wType, zgType;
    DBuf<double> kVal, wVal, zgVal;
    DBuf<double> k, w, kPrev, wPrev, mut, muEffC, prod, rate, srcW, kd, wd, gradX, work, divT, gradR, red, bsrc;
    DBuf<double> mutT, muEffT, kdT, wdT, ctxK, ctxW, kf, qf;
    DBuf<double> rows[3], rowsT[3], gtr;
    std::unique_ptr<DeviceSystem> kSys, wSys;

public:
    const DeviceSystem* kSystem() const { return kSys.get(); }
    const DeviceSystem* omegaSystem() const { return wSys.get(); }
    std::uint64_t reads = 0;  // device reductions read by the host
};

}  // namespace bsim
