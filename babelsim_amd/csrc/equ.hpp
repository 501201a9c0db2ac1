// The procedural equation seam (include/babelsim/equ.h, math.h reductions and
// the FvmExecution diagnostics) resident on a context: fields and equations
// live in HBM, every assembly call writes the equation's own matrix in the
// sliced row layout, and a solve hands that matrix to the equation's
// persistent linear system without a host round trip.
#pragma once

#include "context.hpp"

#include <memory>
#include <vector>

namespace bsim {

enum FieldLocation : int { kCellField = 0, kFaceField = 1 };

// Field<T> (include/babelsim/field.h:57-374) on a context: a cell field of 1 or
// 3 components with a boundary condition per patch (field.h:34-55) and an
// optional boundary flux context (Field::setBoundaryFlux, field.h:277-282), or
// a face field of one value per face.  Values are stored in the context's
// internal numbering.
struct DeviceField {
    DeviceField(Context& ctx, int location, int ncomp);
    Context& C;
    int location, ncomp;
    DBuf<double> v;
    DBuf<int> bct;
    DBuf<double> bcv;
    std::vector<int> hbct;
    DBuf<double> ctx;     // the boundary flux context: a copy of a face field's values
    bool hasCtx = false;
    std::size_t size() const;
    FieldArg arg() const { return {v.p, bct.p, bcv.p, hasCtx ? ctx.p : nullptr, nullptr}; }
    // Field::setBoundaryFlux (field.h:277-282) keeps a snapshot of the values.
    void setBoundaryFlux(const DeviceField& flux);
    // Per-patch types and ncomp values each (cell fields).
    void setBoundary(const int* types, const double* values);
    void setValues(const double* caller);    // caller numbering
    void getValues(double* caller) const;
    bool sameBoundary(const DeviceField& o) const { return hbct == o.hbct; }
};

// The operator options of an equation (include/babelsim/methods.h): the
// unknown's interpolation, gradient, convection and diffusion methods and the
// coefficient interpolation.
struct EquOptions {
    int interp = 0, grad = 0, conv = 1, diff = 0, coefInterp = 0, coefGrad = 0;
};

// equ::Equation<T> (src/discretization/procedural_equation.cpp:15-36): an
// integrated cell system A x = b bound to its unknown.  Assembly calls freeze
// their inputs at once; scalar laplacians also keep the frozen face
// coefficient, the unknown's boundary conditions and the explicit correction
// so that faceFlux can rebuild the assembled diffusion flux.
class DeviceEquation {
public:
    DeviceEquation(DeviceField& unknown, const EquOptions& options, const SolverCfg& linear);
    DeviceField& x;
    int ncomp;
    EquOptions opt;
    SolverCfg linear;
    DBuf<double> diag, off, src;   // sliced rows; src holds ncomp values per cell
    const DeviceField* boundaryFlux = nullptr;
    struct Snapshot {  // the frozen coefficient, the unknown's boundary and the correction
        DBuf<double> kf, corr, bcv, ctx;
        DBuf<int> bct;
        bool hasCtx = false;
    };
    struct Diffusion {
        std::shared_ptr<const Snapshot> snap;
        double factor;
    };
    std::vector<Diffusion> diffusion;

    void reset();                                                       // :124-126
    std::unique_ptr<DeviceEquation> copy() const;                       // :127-131
    void div(const DeviceField& phi, double factor);                    // :132-145
    void laplacian(const DeviceField* coefficient, double k, double multiplier);  // :146-193
    void reaction(const DeviceField* c, double value, double factor);   // :194-204
    void source(const DeviceField* value, const double* constant, double factor);  // :205-213
    void ddt(const DeviceField* capacity, double cap, const DeviceField& previous, double dt, int method,
             const DeviceField* older, double previousDt);              // :214-252
    void scale(double f);                                               // :259-266
    void add(const DeviceEquation& b, double f);                        // :267-280
    void relax(const DeviceField& previous, double alpha);              // :317-324
    void reference(int callerCell, double value);                      // :325-334
    bool referenceIfUnanchored(double value);                           // :335-350
    void diagonal(DeviceField& out) const;
    void rhs(DeviceField& out) const;
    void apply(const DeviceField& xin, DeviceField& out) const;         // :299-310
    void residual(const DeviceField& xin, DeviceField& out) const;      // :311-315
    double relativeResidual(const DeviceField& xin) const;              // :447-452
    void faceFlux(const DeviceField& solution, DeviceField& out) const; // :352-370
    // equ::solve (procedural_equation.cpp:371-390): the equation's persistent
    // system, the unknown written in place; cfg null: the equation's own.
    void solve(const SolverCfg* cfg, SolveRes* results);
    DeviceSystem* system() const { return sys.get(); }

private:
    void requireCell(const DeviceField& f, int ncompWanted) const;
    const double* gradientOf(const DeviceField& f, int method);
    std::unique_ptr<DeviceSystem> sys;
    DBuf<double> grad, work, kfBuf, comp[6], red;
};

// Reductions of src/discretization/field_math.cpp:30-62 and the diagnostics
// of src/discretization/fvm_execution.cpp:91-203, as deterministic two-stage
// device sums read once by the host.
std::vector<double> normsL2(const std::vector<const DeviceField*>& fields);
double relativeChange(const DeviceField& current, const DeviceField& previous);
double relativeMagnitude(const DeviceField& value, const DeviceField& reference);
void fluxBalance(const DeviceField& phi, double out[4]);  // l1, l2, maximum, relative
void courantNumber(const DeviceField& phi, double dt, double out[2]);  // maximum, mean
// math::add / subtract of a face increment (field_math.cpp:65-82): region 0
// all faces, 1 interior faces.
void faceIncrement(const DeviceField& increment, DeviceField& target, double factor, int region);

}  // namespace bsim
