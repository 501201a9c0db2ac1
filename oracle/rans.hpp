// Oracle Wilcox (1988) k-omega declarations.  Test infrastructure only.
#pragma once

#include "internal.hpp"

namespace orc {

// Spatial methods of one transport equation (equation.<name>.* in the
// numerics dictionary).
struct TransportOptions {
    int interp = Linear, grad = GreenGauss, conv = LinearUpwind, diff = Orthogonal;
    int coefInterp = Linear, coefGrad = GreenGauss;
};

// src/physics/RANS/k_omega.cpp:26-56 constructor defaults.
struct KOmegaConfig {
    double betaStar = 0.09, beta = 3.0 / 40.0, gamma = 5.0 / 9.0, sigmaK = 0.5, sigmaOmega = 0.5;
    double kMin = 1e-12, omegaMin = 1e-12, viscosityRatioMax = 1e5;
    double relaxation = 0.7, tolerance = 1e-6;
    TransportOptions kOpts, wOpts;
    SolverConfig kSolver, wSolver;
};

struct TransportReport {
    double rOmega = 0, rK = 0, dOmega = 0, dK = 0;
    int itOmega = 0, itK = 0, statusOmega = 0, statusK = 0;
    double boundedOmegaCells = 0;
    bool healthy = true;
};

// 2 dev(symm(G)):dev(symm(G)) (k_omega.cpp:12-23).
double strainSquared(const T3& g);
// Vector-gradient boundary traces (fvm_execution.cpp:243-265).
void gradientTracesVec(const VField& U, const std::vector<T3>& g, std::vector<T3>& traces);
// div(T) of a cell tensor field with boundary traces, row by row
// (operators.cpp:1568-1588); result per cell, rows summed face by face.
void tensorDivergence(const Mesh& m, const std::vector<T3>& T, const std::vector<T3>& Ttrace, int interp, int gmethod,
                      std::vector<V3>& out);

struct KOmega {
    const Mesh& m;
    KOmegaConfig c;
    double rho, mu;
    SField k, omega;                 // unknowns with their boundary conditions
    std::vector<double> mut, mutTrace;  // calculated: cell values and boundary traces
    LinearSystem kSys, wSys;
    std::vector<double> kHist, wHist;
    KOmega(const Mesh& mesh, const KOmegaConfig& cfg, double density, double viscosity);
    void setBoundaries(const std::vector<BC<double>>& kbc, const std::vector<BC<double>>& wbc);
    void start(const std::vector<double>& phi);  // bound, then mut (k_omega.cpp:52-53)
    SField effectiveViscosity() const;            // mu + mut with traces
    // muEff (transpose(grad U) - 2/3 tr(grad U) I) (model.cpp:67-71) and its divergence.
    void stressDivergence(const VField& U, int mInterp, int mGrad, std::vector<V3>& out) const;
    TransportReport solveTransport(VField& U, const std::vector<double>& phi, int mGrad);
    void updateViscosity(const std::vector<double>& phi);
    void boundOmega(double& bounded);
    // Transient form (PISO): the k and omega histories (history.h:6-27) and
    // the ddt terms of k_omega.cpp:82, 96.  timeMethod 0 is the steady form.
    int timeMethod = 0;
    std::vector<double> kPrev, kOlder, wPrev, wOlder;
    double hdt = 0, hdtPrev = 0;
    int hlevels = 0;
    void saveOld(double dt);  // k_omega.cpp:59-62
};

}  // namespace orc
