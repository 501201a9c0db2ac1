// Driver for the reference's own discretisation layer (src/core/mesh.cpp and
// src/discretization/operators.cpp, compiled from /root/reference by
// oracle/Makefile into oracle/_ref/).  It only converts plain arrays to the
// reference's Mesh/Field types and calls the reference functions; no reference
// code is reproduced here.  Test infrastructure only.
#include "babelsim/operators.h"
#include "internal/field_access.h"
#include "internal/mesh_access.h"

#include <cstring>
#include <exception>
#include <string>

using namespace babelsim;

namespace {
thread_local std::string g_error;
template <class F> int guard(F&& f) {
    try { f(); return 0; }
    catch (const std::exception& e) { g_error = e.what(); return 1; }
    catch (...) { g_error = "unknown"; return 1; }
}
inline void set(double& v, const double* p) { v = p[0]; }
inline void set(Vec3& v, const double* p) { v = {p[0], p[1], p[2]}; }
template <class T> constexpr int nc();
template <> constexpr int nc<double>() { return 1; }
template <> constexpr int nc<Vec3>() { return 3; }

template <class T> Field<T> makeField(const Mesh& m, const double* x, const int* bct, const double* bcv,
                                      const double* ctx) {
    Field<T> f(m, FieldLocation::Cell, "x");
    T* d = detail::fieldData(f);
    for (Index i = 0; i < m.cellCount(); ++i) set(d[i], x + nc<T>() * i);
    for (Index p = 0; p < m.patchCount(); ++p) {
        BoundaryCondition<T> c;
        c.type = static_cast<BoundaryType>(bct[p]);
        set(c.value, bcv + nc<T>() * p);
        f.setBoundary(p, c);
    }
    if (ctx) {
        ScalarField phi(m, FieldLocation::Face, "phi");
        std::memcpy(detail::fieldData(phi), ctx, sizeof(double) * m.faceCount());
        f.setBoundaryFlux(phi);
    }
    return f;
}
void put(const double* v, std::size_t n, double* out) { std::memcpy(out, v, sizeof(double) * n); }
void put(const Vec3* v, std::size_t n, double* out) {
    for (std::size_t i = 0; i < n; ++i) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
}
void put(const Tensor3* v, std::size_t n, double* out) {
    for (std::size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) { out[9 * i + 3 * c] = v[i][c].x; out[9 * i + 3 * c + 1] = v[i][c].y; out[9 * i + 3 * c + 2] = v[i][c].z; }
}
ScalarField faceField(const Mesh& m, const double* v) {
    ScalarField f(m, FieldLocation::Face, "k");
    std::memcpy(detail::fieldData(f), v, sizeof(double) * m.faceCount());
    return f;
}
}  // namespace

extern "C" {
const char* ref_last_error(void) { return g_error.c_str(); }

void* ref_mesh_new(int nv, const double* vertices, int nf, const int* fpo, const int* fpi, const int* owner,
                   const int* neighbour, const int* patch, int npatch) {
    Mesh* out = nullptr;
    guard([&] {
        std::vector<Vec3> vx(nv);
        for (int i = 0; i < nv; ++i) vx[i] = {vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]};
        std::vector<PolyhedralFaceSpec> faces(nf);
        for (int f = 0; f < nf; ++f) {
            faces[f].vertices.assign(fpi + fpo[f], fpi + fpo[f + 1]);
            faces[f].owner = owner[f]; faces[f].neighbour = neighbour[f]; faces[f].patch = patch[f];
        }
        std::vector<PatchSpec> patches(npatch);
        for (int p = 0; p < npatch; ++p) patches[p].name = "patch" + std::to_string(p);
        out = new Mesh(Mesh::polyhedral(std::move(vx), std::move(faces), std::move(patches)));
    });
    return out;
}
void ref_mesh_free(void* m) { delete static_cast<Mesh*>(m); }
void ref_mesh_sizes(void* mp, int* ncells, int* nfaces, int* orthogonal) {
    const Mesh& m = *static_cast<Mesh*>(mp);
    *ncells = m.cellCount(); *nfaces = m.faceCount(); *orthogonal = m.orthogonalGeometry() ? 1 : 0;
}
void ref_mesh_cells(void* mp, double* centres, double* volumes) {
    const Mesh& m = *static_cast<Mesh*>(mp);
    const auto& d = detail::meshData(m);
    put(d.cell_centres.data(), d.cell_centres.size(), centres);
    put(d.cell_volumes.data(), d.cell_volumes.size(), volumes);
}
void ref_mesh_faces(void* mp, double* cf, double* sf, double* af, double* nrm, double* w, double* orth, double* kf,
                    double* skew) {
    const Mesh& m = *static_cast<Mesh*>(mp);
    const auto& d = detail::meshData(m);
    const std::size_t n = d.face_owner.size();
    put(d.face_centres.data(), n, cf); put(d.face_area_vectors.data(), n, sf); put(d.face_areas.data(), n, af);
    put(d.face_normals.data(), n, nrm); put(d.face_owner_weights.data(), n, w);
    put(d.face_orthogonal_coefficients.data(), n, orth); put(d.face_non_orthogonal.data(), n, kf);
    put(d.face_skewness.data(), n, skew);
}

int ref_gradient(void* mp, int ncomp, const double* x, const int* bct, const double* bcv, const double* ctx,
                 int method, double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mp);
        if (ncomp == 1) {
            auto f = makeField<double>(m, x, bct, bcv, ctx);
            VectorField g(m, FieldLocation::Cell, "g");
            gradient(f, g, static_cast<GradientMethod>(method));
            put(detail::fieldData(static_cast<const VectorField&>(g)), m.cellCount(), out);
        } else {
            auto f = makeField<Vec3>(m, x, bct, bcv, ctx);
            TensorField g(m, FieldLocation::Cell, "g");
            gradient(f, g, static_cast<GradientMethod>(method));
            put(detail::fieldData(static_cast<const TensorField&>(g)), m.cellCount(), out);
        }
    });
}
int ref_interpolate(void* mp, int ncomp, const double* x, const int* bct, const double* bcv, const double* ctx,
                    int method, int gmethod, double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mp);
        if (ncomp == 1) {
            auto f = makeField<double>(m, x, bct, bcv, ctx);
            ScalarField o(m, FieldLocation::Face, "o");
            interpolate(f, o, static_cast<InterpolationMethod>(method), static_cast<GradientMethod>(gmethod));
            put(detail::fieldData(static_cast<const ScalarField&>(o)), m.faceCount(), out);
        } else {
            auto f = makeField<Vec3>(m, x, bct, bcv, ctx);
            VectorField o(m, FieldLocation::Face, "o");
            interpolate(f, o, static_cast<InterpolationMethod>(method), static_cast<GradientMethod>(gmethod));
            put(detail::fieldData(static_cast<const VectorField&>(o)), m.faceCount(), out);
        }
    });
}
int ref_flux(void* mp, const double* U, const int* bct, const double* bcv, const double* ctx, int method,
             int gmethod, double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mp);
        auto f = makeField<Vec3>(m, U, bct, bcv, ctx);
        ScalarField o(m, FieldLocation::Face, "o");
        flux(f, o, static_cast<InterpolationMethod>(method), static_cast<GradientMethod>(gmethod));
        put(detail::fieldData(static_cast<const ScalarField&>(o)), m.faceCount(), out);
    });
}
int ref_normal_gradient(void* mp, const double* x, const int* bct, const double* bcv, int gmethod, int dmethod,
                        double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mp);
        auto f = makeField<double>(m, x, bct, bcv, nullptr);
        VectorField g(m, FieldLocation::Cell, "g");
        gradient(f, g, static_cast<GradientMethod>(gmethod));
        for (Index face = 0; face < m.faceCount(); ++face)
            out[face] = integratedNormalGradient(f, g, face, static_cast<DiffusionMethod>(dmethod)) / m.faceArea(face);
    });
}
int ref_divergence_face(void* mp, const double* phi, double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mp);
        ScalarField p = faceField(m, phi);
        ScalarField o(m, FieldLocation::Cell, "o");
        divergence(p, o);
        put(detail::fieldData(static_cast<const ScalarField&>(o)), m.cellCount(), out);
    });
}
int ref_explicit_laplacian(void* mp, const double* x, const int* bct, const double* bcv, const double* kf, double k,
                           int gmethod, int dmethod, double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mp);
        auto f = makeField<double>(m, x, bct, bcv, nullptr);
        ScalarField o(m, FieldLocation::Cell, "o");
        if (kf) {
            ScalarField kk = faceField(m, kf);
            laplacian(kk, f, o, static_cast<GradientMethod>(gmethod), static_cast<DiffusionMethod>(dmethod));
        } else {
            laplacian(k, f, o, static_cast<GradientMethod>(gmethod), static_cast<DiffusionMethod>(dmethod));
        }
        put(detail::fieldData(static_cast<const ScalarField&>(o)), m.cellCount(), out);
    });
}
int ref_explicit_convection(void* mp, int ncomp, const double* x, const int* bct, const double* bcv,
                            const double* phi, int conv, int interp, int gmethod, double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mp);
        ScalarField ph = faceField(m, phi);
        if (ncomp == 1) {
            auto f = makeField<double>(m, x, bct, bcv, phi);
            ScalarField o(m, FieldLocation::Cell, "o");
            convection(ph, f, o, static_cast<ConvectionMethod>(conv), static_cast<InterpolationMethod>(interp),
                       static_cast<GradientMethod>(gmethod));
            put(detail::fieldData(static_cast<const ScalarField&>(o)), m.cellCount(), out);
        } else {
            auto f = makeField<Vec3>(m, x, bct, bcv, phi);
            VectorField o(m, FieldLocation::Cell, "o");
            convection(ph, f, o, static_cast<ConvectionMethod>(conv), static_cast<InterpolationMethod>(interp),
                       static_cast<GradientMethod>(gmethod));
            put(detail::fieldData(static_cast<const VectorField&>(o)), m.cellCount(), out);
        }
    });
}
}  // extern "C"
template <class T> void putEq(const DiscreteEquation<T>& q, const Mesh& m, double* diag, double* upper,
                              double* lower, double* source) {
    for (Index i = 0; i < m.cellCount(); ++i) diag[i] += q.diagonal[i];
    for (Index f = 0; f < m.faceCount(); ++f) { upper[f] += q.upper[f]; lower[f] += q.lower[f]; }
    std::vector<double> s(static_cast<std::size_t>(m.cellCount()) * nc<T>());
    put(q.source.data(), q.source.size(), s.data());
    for (std::size_t i = 0; i < s.size(); ++i) source[i] += s[i];
}
extern "C" {
int ref_assemble_convection(void* mp, int ncomp, const double* x, const int* bct, const double* bcv,
                            const double* phi, int conv, int interp, int gmethod, double scale, double* diag,
                            double* upper, double* lower, double* source) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mp);
        ScalarField ph = faceField(m, phi);
        if (ncomp == 1) {
            auto f = makeField<double>(m, x, bct, bcv, phi);
            ScalarDiscreteEquation q(m);
            addConvection(q, ph, f, static_cast<ConvectionMethod>(conv), static_cast<InterpolationMethod>(interp),
                          static_cast<GradientMethod>(gmethod), scale);
            putEq(q, m, diag, upper, lower, source);
        } else {
            auto f = makeField<Vec3>(m, x, bct, bcv, phi);
            VectorDiscreteEquation q(m);
            addConvection(q, ph, f, static_cast<ConvectionMethod>(conv), static_cast<InterpolationMethod>(interp),
                          static_cast<GradientMethod>(gmethod), scale);
            putEq(q, m, diag, upper, lower, source);
        }
    });
}
int ref_assemble_diffusion(void* mp, int ncomp, const double* x, const int* bct, const double* bcv,
                           const double* ctx, const double* kf, int gmethod, int dmethod, double* diag,
                           double* upper, double* lower, double* source) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mp);
        ScalarField k = faceField(m, kf);
        if (ncomp == 1) {
            auto f = makeField<double>(m, x, bct, bcv, ctx);
            ScalarDiscreteEquation q(m);
            addDiffusion(q, k, f, static_cast<GradientMethod>(gmethod), static_cast<DiffusionMethod>(dmethod));
            putEq(q, m, diag, upper, lower, source);
        } else {
            auto f = makeField<Vec3>(m, x, bct, bcv, ctx);
            VectorDiscreteEquation q(m);
            addDiffusion(q, k, f, static_cast<GradientMethod>(gmethod), static_cast<DiffusionMethod>(dmethod));
            putEq(q, m, diag, upper, lower, source);
        }
    });
}
}  // extern "C"
