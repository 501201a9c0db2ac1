// Face-loop operators of BabelSim's discretisation (src/discretization/
// operators.cpp) as row-gather HIP kernels: every cell reads its own faces, so
// nothing is scattered and no atomics are needed.
#pragma once

#include "dmesh.hpp"

namespace bsim {

enum BcType : int { kFixedValue = 0, kFixedGradient = 1, kZeroGradient = 2, kInletOutlet = 3, kSymmetry = 4 };
enum : int { kLinear = 0, kCorrected = 1 };
enum : int { kGreenGauss = 0, kLeastSquares = 1 };
enum : int { kUpwind = 0, kLinearUpwind = 1, kCentral = 2 };
enum : int { kOrthogonal = 0, kDiffCorrected = 1, kLimitedCorrected = 2 };

// Device-side error bits collected per operation.
enum OpError : int { kErrNoFluxContext = 1, kErrSingularLs = 2, kErrNegativeDiffusivity = 4, kErrNonFinite = 8 };

// A cell field with patch constraints (field.h:57-374): values, per-patch BC
// type/value, the inletOutlet flux context and, for calculated fields,
// per-face boundary traces (which then act as fixed values).
struct FieldArg {
    const double* v = nullptr;
    const int* bcType = nullptr;
    const double* bcVal = nullptr;
    const double* ctx = nullptr;
    const double* trace = nullptr;
};

namespace ops {

// Gradient (GG two-pass on non-orthogonal meshes, or LS).  `work` holds
// 3*ncomp*nc doubles (first GG pass).
void gradient(const MeshView& m, bool orthogonal, int ncomp, FieldArg x, int method, double* out, double* work,
              int* err, hipStream_t s);
// Cell-to-face interpolation (linear or gradient-corrected).
void interpolate(const MeshView& m, int ncomp, FieldArg x, int method, const double* grad, double* out, int* err,
                 hipStream_t s);
// Owner-velocity flux context for inletOutlet on first use (operators.cpp:977-987).
void seedFluxContext(const MeshView& m, const double* U, double* ctx, hipStream_t s);
// U_f . Sf on all faces, with optional corrected interpolation (operators.cpp:957-1002).
void flux(const MeshView& m, FieldArg U, const double* gradU, double* out, int* err, hipStream_t s);
// integratedNormalGradient / |Sf| (fvm_execution.cpp:522-543).
void normalGradient(const MeshView& m, FieldArg x, const double* grad, int dmethod, double* out, int* err,
                    hipStream_t s);
// k_f * integratedNormalGradient (operators.cpp:1004-1024).
void diffusionFlux(const MeshView& m, FieldArg x, const double* kf, const double* grad, int dmethod, double* out,
                   int* err, hipStream_t s);
// sum of signed face fluxes / V (operators.cpp:1055-1063).
void divergenceFace(const MeshView& m, const double* phi, double* out, hipStream_t s);
// Explicit laplacian with face or constant diffusivity (operators.cpp:1157-1213).
void explicitLaplacian(const MeshView& m, FieldArg x, const double* kf, double k, const double* grad, int dmethod,
                       double* out, int* err, hipStream_t s);
// Explicit convection div(phi x) (operators.cpp:1065-1131).
void explicitConvection(const MeshView& m, int ncomp, FieldArg x, const double* phi, int conv, int interp,
                        const double* grad, double* out, int* err, hipStream_t s);
// Implicit convection into CSR rows (operators.cpp:535-661).  Accumulates.
void assembleConvection(const MeshView& m, int ncomp, FieldArg x, const double* phi, int conv, int interp,
                        const double* grad, double scale, double* diag, double* off, double* src, int* err,
                        hipStream_t s);
// Implicit diffusion into CSR rows (operators.cpp:1287-1438).  Adds
// `factor` times the contribution computed on its own, like accumulate().
void assembleDiffusion(const MeshView& m, int ncomp, FieldArg x, const double* kf, const double* grad, int dmethod,
                       double factor, double* diag, double* off, double* src, int* err, hipStream_t s);
// assembleConvection (phi given) then assembleDiffusion (kf given) with the
// same results bit for bit, the gradient-dependent face terms formed once per
// face: faceWork holds 2 * ncomp * nfi doubles.
void assembleTransport(const MeshView& m, int ncomp, FieldArg x, const double* phi, int conv, int interp, double scale,
                       const double* kf, int dmethod, double factor, const double* grad, double* diag, double* off,
                       double* src, double* faceWork, int* err, hipStream_t s);
// assembleDiffusion with a corrected method on a scalar, plus the face
// correction of procedural_equation.cpp:77-87 (diffusionFlux with the method
// minus diffusionFlux orthogonal, every face) into corr, bit for bit as those
// calls; faceWork holds nfi doubles.
void assembleDiffusionCorrected(const MeshView& m, FieldArg x, const double* kf, const double* grad, int dmethod,
                                double factor, double* diag, double* off, double* src, double* corr, double* faceWork,
                                int* err, hipStream_t s);
// Computed-gradient boundary traces of a scalar (fvm_execution.cpp:221-235).
void gradientTraces(const MeshView& m, FieldArg x, const double* grad, double* traces, hipStream_t s);

}  // namespace ops
}  // namespace bsim
