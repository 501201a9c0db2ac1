// C ABI of the device fields, equations and reductions (include/bsim.h,
// "fields and equations").  A field or equation belongs to its context; the
// caller destroys equations before the fields they use and fields before the
// context.
#include "capi_types.hpp"
#include "equ.hpp"
#include "errors.hpp"

#include <cstring>

struct bsim_field {
    std::unique_ptr<bsim::DeviceField> f;
};
struct bsim_equ {
    std::unique_ptr<bsim::DeviceEquation> e;
};

using namespace bsim;

namespace {

DeviceField& fieldOf(const bsim_field* f) {
    if (!f || !f->f) throw std::invalid_argument("null field");
    return *f->f;
}
DeviceEquation& equOf(const bsim_equ* a) {
    if (!a || !a->e) throw std::invalid_argument("null equation");
    return *a->e;
}
EquOptions toOptions(const bsim_equ_methods* m) {
    EquOptions o;
    if (!m) return o;
    o.interp = m->interp; o.grad = m->grad; o.conv = m->conv; o.diff = m->diff;
    o.coefInterp = m->coef_interp; o.coefGrad = m->coef_grad;
    if (o.interp < 0 || o.interp > 1 || o.coefInterp < 0 || o.coefInterp > 1 || o.grad < 0 || o.grad > 1 ||
        o.coefGrad < 0 || o.coefGrad > 1 || o.conv < 0 || o.conv > 2 || o.diff < 0 || o.diff > 2)
        throw std::invalid_argument("unknown operator method");
    return o;
}
void results(DeviceEquation& a, const SolveRes* r, bsim_solve_result* out) {
    for (int k = 0; k < a.ncomp; ++k) out[k] = bsim_solve_result{r[k].status, r[k].iterations, r[k].initial, r[k].final_, r[k].relative};
}

}  // namespace

extern "C" {

int bsim_field_create(bsim_ctx* ctx, int location, int ncomp, const double* values, bsim_field** out) {
    return guarded([&] {
        if (!ctx || !out) throw std::invalid_argument("null argument");
        auto f = std::make_unique<bsim_field>();
        f->f = std::make_unique<DeviceField>(*ctx->c, location, ncomp);
        if (values) f->f->setValues(values);
        *out = f.release();
    });
}
void bsim_field_destroy(bsim_field* f) { delete f; }
int bsim_field_set(bsim_field* f, const double* values) {
    return guarded([&] {
        if (!values) throw std::invalid_argument("null argument");
        fieldOf(f).setValues(values);
    });
}
int bsim_field_get(const bsim_field* f, double* values) {
    return guarded([&] {
        if (!values) throw std::invalid_argument("null argument");
        fieldOf(f).getValues(values);
    });
}
int bsim_field_set_boundary(bsim_field* f, const int* bc_type, const double* bc_value) {
    return guarded([&] {
        if (!bc_type || !bc_value) throw std::invalid_argument("null argument");
        fieldOf(f).setBoundary(bc_type, bc_value);
    });
}
int bsim_field_set_flux(bsim_field* f, const bsim_field* face_flux) {
    return guarded([&] {
        fieldOf(f).setBoundaryFlux(fieldOf(face_flux));
    });
}

int bsim_equ_create(bsim_field* unknown, const bsim_equ_methods* methods, const bsim_solver_config* linear,
                    bsim_equ** out) {
    return guarded([&] {
        if (!out || !linear) throw std::invalid_argument("null argument");
        auto a = std::make_unique<bsim_equ>();
        a->e = std::make_unique<DeviceEquation>(fieldOf(unknown), toOptions(methods), toSolverCfg(linear));
        *out = a.release();
    });
}
void bsim_equ_destroy(bsim_equ* a) { delete a; }
int bsim_equ_reset(bsim_equ* a) { return guarded([&] { equOf(a).reset(); }); }
int bsim_equ_copy(const bsim_equ* a, bsim_equ** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null argument");
        auto b = std::make_unique<bsim_equ>();
        b->e = equOf(a).copy();
        *out = b.release();
    });
}
int bsim_equ_add(bsim_equ* a, const bsim_equ* b, double factor) {
    return guarded([&] { equOf(a).add(equOf(b), factor); });
}
int bsim_equ_scale(bsim_equ* a, double factor) { return guarded([&] { equOf(a).scale(factor); }); }
int bsim_equ_div(bsim_equ* a, const bsim_field* flux, double factor) {
    return guarded([&] { equOf(a).div(fieldOf(flux), factor); });
}
int bsim_equ_laplacian(bsim_equ* a, const bsim_field* coefficient, double k, double multiplier) {
    return guarded([&] { equOf(a).laplacian(coefficient ? &fieldOf(coefficient) : nullptr, k, multiplier); });
}
int bsim_equ_reaction(bsim_equ* a, const bsim_field* coefficient, double value, double factor) {
    return guarded([&] { equOf(a).reaction(coefficient ? &fieldOf(coefficient) : nullptr, value, factor); });
}
int bsim_equ_source(bsim_equ* a, const bsim_field* value, const double* constant, double factor) {
    return guarded([&] {
        if (!value && !constant) throw std::invalid_argument("null argument");
        equOf(a).source(value ? &fieldOf(value) : nullptr, constant, factor);
    });
}
int bsim_equ_ddt(bsim_equ* a, const bsim_field* capacity, double cap, const bsim_field* previous, double dt, int method,
                 const bsim_field* older, double previous_dt) {
    return guarded([&] {
        equOf(a).ddt(capacity ? &fieldOf(capacity) : nullptr, cap, fieldOf(previous), dt, method,
                     older ? &fieldOf(older) : nullptr, previous_dt);
    });
}
int bsim_equ_relax(bsim_equ* a, const bsim_field* previous, double alpha) {
    return guarded([&] { equOf(a).relax(fieldOf(previous), alpha); });
}
int bsim_equ_reference(bsim_equ* a, int cell, double value) {
    return guarded([&] { equOf(a).reference(cell, value); });
}
int bsim_equ_reference_if_unanchored(bsim_equ* a, double value, int* applied) {
    return guarded([&] {
        const bool r = equOf(a).referenceIfUnanchored(value);
        if (applied) *applied = r ? 1 : 0;
    });
}
int bsim_equ_diagonal(const bsim_equ* a, bsim_field* out) { return guarded([&] { equOf(a).diagonal(fieldOf(out)); }); }
int bsim_equ_rhs(const bsim_equ* a, bsim_field* out) { return guarded([&] { equOf(a).rhs(fieldOf(out)); }); }
int bsim_equ_apply(const bsim_equ* a, const bsim_field* x, bsim_field* out) {
    return guarded([&] { equOf(a).apply(fieldOf(x), fieldOf(out)); });
}
int bsim_equ_residual(const bsim_equ* a, const bsim_field* x, bsim_field* out) {
    return guarded([&] { equOf(a).residual(fieldOf(x), fieldOf(out)); });
}
int bsim_equ_face_flux(const bsim_equ* a, const bsim_field* solution, bsim_field* out) {
    return guarded([&] { equOf(a).faceFlux(fieldOf(solution), fieldOf(out)); });
}
int bsim_equ_relative_residual(const bsim_equ* a, const bsim_field* x, double* out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null argument");
        *out = equOf(a).relativeResidual(fieldOf(x));
    });
}
int bsim_equ_solve(bsim_equ* a, const bsim_solver_config* cfg, bsim_solve_result* out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null argument");
        DeviceEquation& e = equOf(a);
        SolverCfg c;
        if (cfg) c = toSolverCfg(cfg);
        SolveRes r[3];
        e.solve(cfg ? &c : nullptr, r);
        results(e, r, out);
    });
}
int bsim_equ_performance(const bsim_equ* a, bsim_equation_counters* out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null argument");
        const DeviceEquation& e = equOf(a);
        std::memset(out, 0, sizeof(*out));
        if (e.system()) *out = equationCounters(*e.system(), "");
    });
}

int bsim_math_norms(const bsim_field* const* fields, int n, double* out) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (!fields || !out))) throw std::invalid_argument("null argument");
        std::vector<const DeviceField*> v;
        for (int k = 0; k < n; ++k) v.push_back(&fieldOf(fields[k]));
        const std::vector<double> r = normsL2(v);
        std::memcpy(out, r.data(), sizeof(double) * r.size());
    });
}
int bsim_math_relative_change(const bsim_field* current, const bsim_field* previous, double* out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null argument");
        *out = relativeChange(fieldOf(current), fieldOf(previous));
    });
}
int bsim_math_relative_magnitude(const bsim_field* value, const bsim_field* reference, double* out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null argument");
        *out = relativeMagnitude(fieldOf(value), fieldOf(reference));
    });
}
int bsim_math_flux_balance(const bsim_field* flux, double* out4) {
    return guarded([&] {
        if (!out4) throw std::invalid_argument("null argument");
        fluxBalance(fieldOf(flux), out4);
    });
}
int bsim_math_courant(const bsim_field* flux, double dt, double* out2) {
    return guarded([&] {
        if (!out2) throw std::invalid_argument("null argument");
        courantNumber(fieldOf(flux), dt, out2);
    });
}
int bsim_math_face_add(const bsim_field* increment, bsim_field* target, double factor, int region) {
    return guarded([&] { faceIncrement(fieldOf(increment), fieldOf(target), factor, region); });
}

}  // extern "C"
