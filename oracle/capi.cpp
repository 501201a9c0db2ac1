// Oracle C ABI wrappers (test infrastructure only).
#include "oracle.h"

#include <omp.h>

#include "piso.hpp"

#include <cstring>
#include <memory>
#include <exception>
#include <string>

using namespace orc;

namespace {

thread_local std::string g_error;

template <class F> int guard(F&& f) {
    try { f(); return 0; }
    catch (const std::exception& e) { g_error = e.what(); return 1; }
    catch (...) { g_error = "unknown error"; return 1; }
}

SolverConfig toConfig(const orc_solver_config* c) {
    SolverConfig s;
    s.solver = c->solver; s.pc = c->pc; s.atol = c->atol; s.rtol = c->rtol; s.maxIter = c->max_iter;
    s.warmStart = c->warm_start; s.rebuildInterval = c->rebuild_interval; s.reuseGrowth = c->reuse_growth;
    s.amgMaxLevels = c->amg_max_levels; s.amgCoarseSize = c->amg_coarse_size;
    s.amgPreSweeps = c->amg_pre_sweeps; s.amgPostSweeps = c->amg_post_sweeps;
    s.amgOmega = c->amg_omega; s.amgStrength = c->amg_strength; s.amgCorrection = c->amg_correction;
    return s;
}

template <class T> Field<T> makeField(const Mesh& m, const double* x, const int* bct, const double* bcv,
                                      const double* ctx) {
    Field<T> f(m);
    constexpr int k = ncomp<T>();
    for (Index i = 0; i < m.nc; ++i) for (int c = 0; c < k; ++c) setComp(f.v[i], c, x[i * k + c]);
    for (Index p = 0; p < m.np; ++p) {
        f.bc[p].type = bct[p];
        for (int c = 0; c < k; ++c) setComp(f.bc[p].value, c, bcv[p * k + c]);
    }
    if (ctx) f.fluxCtx.assign(ctx, ctx + m.nf);
    return f;
}

template <class T> void store(const std::vector<T>& v, double* out) {
    constexpr int k = ncomp<T>();
    for (std::size_t i = 0; i < v.size(); ++i) for (int c = 0; c < k; ++c) out[i * k + c] = comp(v[i], c);
}
void storeG(const std::vector<V3>& v, double* out) { store(v, out); }
void storeG(const std::vector<T3>& v, double* out) {
    for (std::size_t i = 0; i < v.size(); ++i)
        for (int c = 0; c < 3; ++c) for (int d = 0; d < 3; ++d) out[i * 9 + c * 3 + d] = v[i].r[c][d];
}

template <class T> void storeLdu(const Mesh& m, const Ldu<T>& q, double* diag, double* upper, double* lower,
                                 double* source) {
    constexpr int k = ncomp<T>();
    for (Index i = 0; i < m.nc; ++i) {
        diag[i] += q.diag[i];
        for (int c = 0; c < k; ++c) source[i * k + c] += comp(q.source[i], c);
    }
    for (Index f = 0; f < m.nf; ++f) { upper[f] += q.upper[f]; lower[f] += q.lower[f]; }
}

Csr makeCsr(int n, const int* ptr, const int* col, const double* diag, const double* val) {
    Csr A;
    A.n = n;
    A.ptr.assign(ptr, ptr + n + 1);
    A.col.assign(col, col + ptr[n]);
    A.diag.assign(diag, diag + n);
    A.val.assign(val, val + ptr[n]);
    return A;
}

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_error.c_str(); }
int orc_threads(void) { return omp_get_max_threads(); }

void* orc_mesh_new(int nv, const double* vertices, int nf, const int* fpo, const int* fpi, const int* owner,
                   const int* neighbour, const int* patch, int npatch) {
    Mesh* out = nullptr;
    guard([&] {
        std::vector<V3> vx(nv);
        for (int i = 0; i < nv; ++i) vx[i] = {vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]};
        std::vector<Index> o(fpo, fpo + nf + 1), r(fpi, fpi + fpo[nf]);
        out = new Mesh(buildMesh(vx, o, r, std::vector<Index>(owner, owner + nf),
                                 std::vector<Index>(neighbour, neighbour + nf), std::vector<Index>(patch, patch + nf),
                                 npatch));
    });
    return out;
}
void orc_mesh_free(void* mesh) { delete static_cast<Mesh*>(mesh); }
void orc_mesh_sizes(void* mesh, int* nc, int* nf, int* orthogonal, int* nfp) {
    const Mesh& m = *static_cast<Mesh*>(mesh);
    *nc = m.nc; *nf = m.nf; *orthogonal = m.orthogonal ? 1 : 0; *nfp = static_cast<int>(m.fpi.size());
}
void orc_mesh_cells(void* mesh, double* centres, double* volumes) {
    const Mesh& m = *static_cast<Mesh*>(mesh);
    store(m.C, centres);
    std::memcpy(volumes, m.V.data(), sizeof(double) * m.nc);
}
void orc_mesh_faces(void* mesh, double* cf, double* sf, double* af, double* nrm, double* w, double* orth,
                    double* kf, double* skew) {
    const Mesh& m = *static_cast<Mesh*>(mesh);
    store(m.Cf, cf); store(m.Sf, sf); store(m.nrm, nrm); store(m.kf, kf); store(m.skew, skew);
    std::memcpy(af, m.Af.data(), sizeof(double) * m.nf);
    std::memcpy(w, m.w.data(), sizeof(double) * m.nf);
    std::memcpy(orth, m.orth.data(), sizeof(double) * m.nf);
}
void orc_mesh_rings(void* mesh, int* fpo, int* fpi) {
    const Mesh& m = *static_cast<Mesh*>(mesh);
    std::memcpy(fpo, m.fpo.data(), sizeof(int) * m.fpo.size());
    std::memcpy(fpi, m.fpi.data(), sizeof(int) * m.fpi.size());
}

int orc_gradient(void* mesh, int ncomp, const double* x, const int* bct, const double* bcv, const double* ctx,
                 int method, double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        if (ncomp == 1) { auto f = makeField<double>(m, x, bct, bcv, ctx); std::vector<V3> g; gradient(f, g, method); storeG(g, out); }
        else { auto f = makeField<V3>(m, x, bct, bcv, ctx); std::vector<T3> g; gradient(f, g, method); storeG(g, out); }
    });
}
int orc_gradient_traces(void* mesh, const double* x, const int* bct, const double* bcv, const double* grad,
                        double* traces) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        auto f = makeField<double>(m, x, bct, bcv, nullptr);
        std::vector<V3> g(m.nc), t;
        for (Index i = 0; i < m.nc; ++i) g[i] = {grad[3 * i], grad[3 * i + 1], grad[3 * i + 2]};
        gradientTraces(f, g, t);
        store(t, traces);
    });
}
int orc_interpolate(void* mesh, int ncomp, const double* x, const int* bct, const double* bcv, const double* ctx,
                    int method, int gmethod, double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        if (ncomp == 1) { auto f = makeField<double>(m, x, bct, bcv, ctx); std::vector<double> o; interpolate(f, o, method, gmethod); store(o, out); }
        else { auto f = makeField<V3>(m, x, bct, bcv, ctx); std::vector<V3> o; interpolate(f, o, method, gmethod); store(o, out); }
    });
}
int orc_flux(void* mesh, const double* U, const int* bct, const double* bcv, const double* ctx, int method,
             int gmethod, double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        auto f = makeField<V3>(m, U, bct, bcv, ctx);
        std::vector<double> o;
        flux(f, o, method, gmethod);
        store(o, out);
    });
}
int orc_normal_gradient(void* mesh, const double* x, const int* bct, const double* bcv, int gmethod, int dmethod,
                        double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        auto f = makeField<double>(m, x, bct, bcv, nullptr);
        std::vector<double> o;
        normalGradient(f, nullptr, gmethod, dmethod, o);
        store(o, out);
    });
}
int orc_diffusion_flux(void* mesh, const double* x, const int* bct, const double* bcv, const double* ctx,
                       const double* kf, int gmethod, int dmethod, double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        auto f = makeField<double>(m, x, bct, bcv, ctx);
        std::vector<V3> g;
        gradient(f, g, gmethod);
        std::vector<double> o;
        diffusionFlux(std::vector<double>(kf, kf + m.nf), f, g, dmethod, o);
        store(o, out);
    });
}
int orc_divergence_face(void* mesh, const double* phi, double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        std::vector<double> o;
        divergenceFace(m, std::vector<double>(phi, phi + m.nf), o);
        store(o, out);
    });
}
int orc_explicit_laplacian(void* mesh, const double* x, const int* bct, const double* bcv, const double* kf,
                           double k, int gmethod, int dmethod, double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        auto f = makeField<double>(m, x, bct, bcv, nullptr);
        std::vector<double> kv, o;
        if (kf) kv.assign(kf, kf + m.nf);
        explicitLaplacian(f, kf ? &kv : nullptr, k, gmethod, dmethod, o);
        store(o, out);
    });
}
int orc_explicit_convection(void* mesh, int ncomp, const double* x, const int* bct, const double* bcv,
                            const double* phi, int conv, int interp, int gmethod, double* out) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        std::vector<double> ph(phi, phi + m.nf);
        if (ncomp == 1) { auto f = makeField<double>(m, x, bct, bcv, phi); std::vector<double> o; explicitConvection(ph, f, o, conv, interp, gmethod); store(o, out); }
        else { auto f = makeField<V3>(m, x, bct, bcv, phi); std::vector<V3> o; explicitConvection(ph, f, o, conv, interp, gmethod); store(o, out); }
    });
}
int orc_assemble_convection(void* mesh, int ncomp, const double* x, const int* bct, const double* bcv,
                            const double* phi, int conv, int interp, int gmethod, double scale, double* diag,
                            double* upper, double* lower, double* source) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        std::vector<double> ph(phi, phi + m.nf);
        if (ncomp == 1) {
            auto f = makeField<double>(m, x, bct, bcv, phi); Ldu<double> q; q.init(m);
            addConvection(q, ph, f, conv, interp, gmethod, scale); storeLdu(m, q, diag, upper, lower, source);
        } else {
            auto f = makeField<V3>(m, x, bct, bcv, phi); Ldu<V3> q; q.init(m);
            addConvection(q, ph, f, conv, interp, gmethod, scale); storeLdu(m, q, diag, upper, lower, source);
        }
    });
}
int orc_assemble_diffusion(void* mesh, int ncomp, const double* x, const int* bct, const double* bcv,
                           const double* ctx, const double* kf, int gmethod, int dmethod, double* diag,
                           double* upper, double* lower, double* source) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        std::vector<double> k(kf, kf + m.nf);
        if (ncomp == 1) {
            auto f = makeField<double>(m, x, bct, bcv, ctx); Ldu<double> q; q.init(m);
            addDiffusion<double>(q, k, f, gmethod, dmethod, nullptr); storeLdu(m, q, diag, upper, lower, source);
        } else {
            auto f = makeField<V3>(m, x, bct, bcv, ctx); Ldu<V3> q; q.init(m);
            addDiffusion<V3>(q, k, f, gmethod, dmethod, nullptr); storeLdu(m, q, diag, upper, lower, source);
        }
    });
}

int orc_csr_solve(int n, const int* ptr, const int* col, const double* diag, const double* val, const double* b,
                  double* x, const orc_solver_config* cfg, orc_solve_result* res, double* hist, int cap,
                  int* hist_len) {
    return guard([&] {
        LinearSystem sys;
        sys.A = makeCsr(n, ptr, col, diag, val);
        sys.update(sys.A.diag, sys.A.val);
        sys.prepare(toConfig(cfg));
        std::vector<double> bv(b, b + n), xv(x, x + n);
        const SolveResult r = sys.solve(bv, xv, 0);
        std::memcpy(x, xv.data(), sizeof(double) * n);
        res->status = r.status; res->iterations = r.iterations; res->initial = r.initial;
        res->final_residual = r.final_; res->relative = r.relative;
        const int len = static_cast<int>(sys.hist.size());
        if (hist_len) *hist_len = len;
        if (hist) std::memcpy(hist, sys.hist.data(), sizeof(double) * std::min(len, cap));
    });
}

int orc_block_ilu(int n, const int* ptr, const int* col, const double* diag, const double* val, double* fd,
                  double* fv) {
    return guard([&] {
        const Csr A = makeCsr(n, ptr, col, diag, val);
        std::vector<double> d, v;
        blockIluFactor(A, d, v);
        std::memcpy(fd, d.data(), sizeof(double) * n);
        if (!v.empty()) std::memcpy(fv, v.data(), sizeof(double) * v.size());
    });
}
int orc_whole_ilu(int n, const int* ptr, const int* col, const double* diag, const double* val, int chol,
                  int* fptr, int* fcol, double* f, int cap, int* nf, int* failed) {
    return guard([&] {
        const Csr A = makeCsr(n, ptr, col, diag, val);
        WholeIlu w;
        w.analyse(A, chol != 0);
        w.factor(A);
        *nf = static_cast<int>(w.col.size());
        *failed = w.fail ? 1 : 0;
        if (*nf > cap) throw std::invalid_argument("factor capacity too small");
        std::memcpy(fptr, w.ptr.data(), sizeof(int) * (n + 1));
        if (*nf) {
            std::memcpy(fcol, w.col.data(), sizeof(int) * *nf);
            std::memcpy(f, w.f.data(), sizeof(double) * *nf);
        }
    });
}
int orc_block_ilu_apply(int n, const int* ptr, const int* col, const double* diag, const double* fd,
                        const double* fv, const double* r, double* z) {
    return guard([&] {
        const Csr A = makeCsr(n, ptr, col, diag, fv);
        blockIluApply(A, std::vector<double>(fd, fd + n), A.val, r, z);
    });
}

struct OrcSystem {
    const Mesh* mesh;
    LinearSystem sys;
};
void* orc_system_new(void* mesh) {
    OrcSystem* out = nullptr;
    guard([&] { out = new OrcSystem{static_cast<const Mesh*>(mesh), {}}; });
    return out;
}
void orc_system_free(void* s) { delete static_cast<OrcSystem*>(s); }
int orc_system_solve_ldu(void* sp, const double* diag, const double* upper, const double* lower, int nrhs,
                         const double* source, double* x, const orc_solver_config* cfg, orc_solve_result* results) {
    return guard([&] {
        if (nrhs != 1 && nrhs != 3) throw std::invalid_argument("linear systems have one or three right-hand sides");
        OrcSystem& S = *static_cast<OrcSystem*>(sp);
        const Mesh& m = *S.mesh;
        std::vector<double> d(diag, diag + m.nc), up(upper, upper + m.nf), lo(lower, lower + m.nf);
        Csr A;
        lduToCsr(m, d, up, lo, A);
        if (!S.sys.ready) S.sys.A = A;
        S.sys.update(A.diag, A.val);
        S.sys.prepare(toConfig(cfg));
        for (int comp = 0; comp < nrhs; ++comp) {
            std::vector<double> b(m.nc), xv(m.nc);
            for (Index i = 0; i < m.nc; ++i) { b[i] = source[i * nrhs + comp]; xv[i] = x[i * nrhs + comp]; }
            const SolveResult r = S.sys.solve(b, xv, comp);
            for (Index i = 0; i < m.nc; ++i) x[i * nrhs + comp] = xv[i];
            results[comp].status = r.status; results[comp].iterations = r.iterations; results[comp].initial = r.initial;
            results[comp].final_residual = r.final_; results[comp].relative = r.relative;
        }
    });
}
int orc_system_counters(void* sp, uint64_t* o, int* last_status, double* last3) {
    return guard([&] {
        const LinearSystem& s = static_cast<OrcSystem*>(sp)->sys;
        const uint64_t v[8] = {s.matrixUpdates, s.rhsOnly, s.pcSetups, s.pcReuses, s.retries, s.solves,
                               s.krylovIterations, s.trueResidualChecks};
        std::memcpy(o, v, sizeof(v));
        if (last_status) *last_status = s.lastStatus;
        if (last3) { last3[0] = s.lastInitial; last3[1] = s.lastFinal; last3[2] = s.lastRelative; }
    });
}

int orc_amg_hierarchy(int n, const int* ptr, const int* col, const double* diag, const double* val,
                      const orc_solver_config* cfg, int* nlevels, int* sizes, int* agg0, double* coarse_diag1) {
    return guard([&] {
        Amg amg;
        amg.setup(makeCsr(n, ptr, col, diag, val), toConfig(cfg));
        *nlevels = static_cast<int>(amg.levels.size());
        for (std::size_t l = 0; l < amg.levels.size(); ++l) sizes[l] = amg.levels[l].A.n;
        if (agg0 && !amg.levels[0].agg.empty()) std::memcpy(agg0, amg.levels[0].agg.data(), sizeof(int) * n);
        if (coarse_diag1 && amg.levels.size() > 1)
            std::memcpy(coarse_diag1, amg.levels[1].A.diag.data(), sizeof(double) * amg.levels[1].A.n);
    });
}

namespace {
SimpleConfig simpleConfig(const orc_simple_config* cfg) {
    SimpleConfig c;
    c.rho = cfg->rho; c.mu = cfg->mu;
    c.mInterp = cfg->m_interp; c.mGrad = cfg->m_grad; c.mConv = cfg->m_conv; c.mDiff = cfg->m_diff;
    c.mCoefInterp = cfg->m_coef_interp; c.mCoefGrad = cfg->m_coef_grad;
    c.pInterp = cfg->p_interp; c.pGrad = cfg->p_grad; c.pDiff = cfg->p_diff;
    c.pCoefInterp = cfg->p_coef_interp; c.pCoefGrad = cfg->p_coef_grad;
    c.alphaU = cfg->alpha_u; c.alphaP = cfg->alpha_p; c.nonOrthCorr = cfg->non_orth_corr;
    c.tolMass = cfg->tol_mass; c.tolU = cfg->tol_u; c.tolMom = cfg->tol_mom; c.tolP = cfg->tol_p;
    c.mom = toConfig(&cfg->mom); c.pre = toConfig(&cfg->pre);
    return c;
}
KOmegaConfig komegaConfig(const orc_komega_config* k) {
    KOmegaConfig c;
    c.betaStar = k->beta_star; c.beta = k->beta; c.gamma = k->gamma; c.sigmaK = k->sigma_k;
    c.sigmaOmega = k->sigma_omega; c.kMin = k->k_min; c.omegaMin = k->omega_min;
    c.viscosityRatioMax = k->viscosity_ratio_max; c.relaxation = k->relaxation; c.tolerance = k->tolerance;
    c.kOpts = {k->k_interp, k->k_grad, k->k_conv, k->k_diff, k->k_coef_interp, k->k_coef_grad};
    c.wOpts = {k->w_interp, k->w_grad, k->w_conv, k->w_diff, k->w_coef_interp, k->w_coef_grad};
    c.kSolver = toConfig(&k->k_solver); c.wSolver = toConfig(&k->w_solver);
    return c;
}
std::vector<BC<double>> scalarBcs(const Mesh& m, const int* t, const double* v) {
    std::vector<BC<double>> b(m.np);
    for (Index p = 0; p < m.np; ++p) { b[p].type = t[p]; b[p].value = v[p]; }
    return b;
}
}  // namespace

void* orc_simple_new_komega(void* mesh, const orc_simple_config* cfg, const orc_komega_config* kw, const int* ubct,
                            const double* ubcv, const int* pbct, const double* pbcv, const int* kbct,
                            const double* kbcv, const int* wbct, const double* wbcv) {
    Simple* s = nullptr;
    guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        SimpleConfig c = simpleConfig(cfg);
        c.turbulence = 1;
        c.kw = komegaConfig(kw);
        std::unique_ptr<Simple> owned(new Simple(m, c));
        std::vector<BC<V3>> ub(m.np);
        for (Index p = 0; p < m.np; ++p) { ub[p].type = ubct[p]; ub[p].value = {ubcv[3 * p], ubcv[3 * p + 1], ubcv[3 * p + 2]}; }
        owned->setBoundaries(ub, scalarBcs(m, pbct, pbcv));
        owned->turb->setBoundaries(scalarBcs(m, kbct, kbcv), scalarBcs(m, wbct, wbcv));
        s = owned.release();
    });
    return s;
}
int orc_simple_set_turbulence(void* sp, const double* k, const double* omega) {
    return guard([&] {
        Simple& s = *static_cast<Simple*>(sp);
        if (!s.turb) throw std::invalid_argument("laminar SIMPLE has no turbulence fields");
        for (Index i = 0; i < s.m.nc; ++i) {
            if (k) s.turb->k.v[i] = k[i];
            if (omega) s.turb->omega.v[i] = omega[i];
        }
    });
}
int orc_simple_get_turbulence(void* sp, double* k, double* omega, double* mut) {
    return guard([&] {
        Simple& s = *static_cast<Simple*>(sp);
        if (!s.turb) throw std::invalid_argument("laminar SIMPLE has no turbulence fields");
        if (k) store(s.turb->k.v, k);
        if (omega) store(s.turb->omega.v, omega);
        if (mut) store(s.turb->mut, mut);
    });
}
int orc_komega_transport(void* mesh, const orc_komega_config* kw, double rho, double mu, int m_grad,
                         const double* U, const int* ubct, const double* ubcv, const double* phi,
                         const double* k0, const int* kbct, const double* kbcv, const double* w0, const int* wbct,
                         const double* wbcv, double* k, double* omega, double* mut, double* out4) {
    return guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        KOmega model(m, komegaConfig(kw), rho, mu);
        model.setBoundaries(scalarBcs(m, kbct, kbcv), scalarBcs(m, wbct, wbcv));
        model.k.v.assign(k0, k0 + m.nc);
        model.omega.v.assign(w0, w0 + m.nc);
        const std::vector<double> ph(phi, phi + m.nf);
        model.start(ph);
        VField Uf = makeField<V3>(m, U, ubct, ubcv, phi);
        const TransportReport r = model.solveTransport(Uf, ph, m_grad);
        if (!r.healthy) throw std::runtime_error("transport solve failed");
        store(model.k.v, k);
        store(model.omega.v, omega);
        store(model.mut, mut);
        out4[0] = r.rOmega; out4[1] = r.rK; out4[2] = r.dOmega; out4[3] = r.dK;
    });
}
int orc_strain_squared(const double* g9, double* out) {
    return guard([&] {
        T3 g;
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) g.r[i][j] = g9[3 * i + j];
        *out = strainSquared(g);
    });
}
int orc_stress_divergence(void* mesh, const double* U, const int* bct, const double* bcv, const double* ctx,
                          const double* coef, const double* coef_trace, int interp, int gmethod, double* out) {
    return guard([&] {
        // coef (cell values and boundary traces) * (transpose(grad U) - 2/3 tr(grad U) I), then div.
        const Mesh& m = *static_cast<Mesh*>(mesh);
        VField Uf = makeField<V3>(m, U, bct, bcv, ctx);
        std::vector<T3> g, gt;
        gradient(Uf, g, gmethod);
        gradientTracesVec(Uf, g, gt);
        auto stress = [](double mu_, const T3& G) {
            const double tr = G.r[0][0] + G.r[1][1] + G.r[2][2];
            T3 s;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) s.r[i][j] = mu_ * (G.r[j][i] - (2.0 / 3.0) * (i == j ? tr : 0.0));
            return s;
        };
        std::vector<T3> T(m.nc), Tt(m.nf);
        for (Index i = 0; i < m.nc; ++i) T[i] = stress(coef[i], g[i]);
        for (Index f = 0; f < m.nf; ++f) if (m.boundary(f)) Tt[f] = stress(coef_trace[f], gt[f]);
        std::vector<V3> d;
        tensorDivergence(m, T, Tt, interp, gmethod, d);
        store(d, out);
    });
}

void* orc_simple_new(void* mesh, const orc_simple_config* cfg, const int* ubct, const double* ubcv,
                     const int* pbct, const double* pbcv) {
    Simple* s = nullptr;
    guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        SimpleConfig c;
        c.rho = cfg->rho; c.mu = cfg->mu;
        c.mInterp = cfg->m_interp; c.mGrad = cfg->m_grad; c.mConv = cfg->m_conv; c.mDiff = cfg->m_diff;
        c.mCoefInterp = cfg->m_coef_interp; c.mCoefGrad = cfg->m_coef_grad;
        c.pInterp = cfg->p_interp; c.pGrad = cfg->p_grad; c.pDiff = cfg->p_diff;
        c.pCoefInterp = cfg->p_coef_interp; c.pCoefGrad = cfg->p_coef_grad;
        c.alphaU = cfg->alpha_u; c.alphaP = cfg->alpha_p; c.nonOrthCorr = cfg->non_orth_corr;
        c.tolMass = cfg->tol_mass; c.tolU = cfg->tol_u; c.tolMom = cfg->tol_mom; c.tolP = cfg->tol_p;
        c.mom = toConfig(&cfg->mom); c.pre = toConfig(&cfg->pre);
        s = new Simple(m, c);
        std::vector<BC<V3>> ub(m.np);
        std::vector<BC<double>> pb(m.np);
        for (Index p = 0; p < m.np; ++p) {
            ub[p].type = ubct[p]; ub[p].value = {ubcv[3 * p], ubcv[3 * p + 1], ubcv[3 * p + 2]};
            pb[p].type = pbct[p]; pb[p].value = pbcv[p];
        }
        s->setBoundaries(ub, pb);
    });
    return s;
}
void orc_simple_free(void* s) { delete static_cast<Simple*>(s); }
int orc_simple_set_fields(void* sp, const double* U, const double* p) {
    return guard([&] {
        Simple& s = *static_cast<Simple*>(sp);
        for (Index i = 0; i < s.m.nc; ++i) { s.U.v[i] = {U[3 * i], U[3 * i + 1], U[3 * i + 2]}; s.p.v[i] = p[i]; }
    });
}
int orc_simple_start(void* sp) { return guard([&] { static_cast<Simple*>(sp)->start(); }); }
int orc_simple_iterate(void* sp, int n, orc_iter_report* reports) {
    return guard([&] {
        Simple& s = *static_cast<Simple*>(sp);
        for (int k = 0; k < n; ++k) {
            const IterReport r = s.iterate();
            orc_iter_report& o = reports[k];
            o.rU = r.rU; o.dU = r.dU; o.dP = r.dP; o.mass = r.mass; o.linP = r.linP;
            for (int c = 0; c < 3; ++c) o.mom_iters[c] = r.momIters[c];
            o.p_iters = r.pIters; o.mom_status = r.momStatus; o.p_status = r.pStatus;
            o.converged = r.converged; o.referenced = r.referenced;
            o.d_turb = r.dTurb; o.r_turb = r.rTurb; o.bounded_omega = r.boundedOmega;
            o.k_iters = r.kIters; o.w_iters = r.wIters; o.turb_linear = r.turbLinear;
            o.turb_converged = r.turbConverged;
            o.numerical_failure = r.failed;
            if (r.failed) {
                for (int j = k + 1; j < n; ++j) {
                    reports[j] = orc_iter_report{};
                    reports[j].numerical_failure = 1;
                }
                break;
            }
        }
    });
}
int orc_simple_get_fields(void* sp, double* U, double* p, double* phi) {
    return guard([&] {
        Simple& s = *static_cast<Simple*>(sp);
        if (U) store(s.U.v, U);
        if (p) store(s.p.v, p);
        if (phi) store(s.phi, phi);
    });
}
namespace {
void systemCounters(const LinearSystem& s, uint64_t* o) {
    const uint64_t v[8] = {s.matrixUpdates, s.rhsOnly, s.pcSetups, s.pcReuses, s.retries, s.solves,
                           s.krylovIterations, s.trueResidualChecks};
    std::memcpy(o, v, sizeof(v));
}
}  // namespace
int orc_simple_counters(void* sp, int which, uint64_t* out8) {
    return guard([&] {
        Simple& s = *static_cast<Simple*>(sp);
        if (which != 0 && which != 1) throw std::invalid_argument("system index is 0 (momentum) or 1 (pressure)");
        systemCounters(which == 0 ? s.momSys : s.pSys, out8);
    });
}
int orc_piso_counters(void* sp, int which, uint64_t* out8) {
    return guard([&] {
        Piso& s = *static_cast<Piso*>(sp);
        if (which != 0 && which != 1) throw std::invalid_argument("system index is 0 (momentum) or 1 (pressure)");
        systemCounters(which == 0 ? s.momSys : s.pSys, out8);
    });
}
int orc_simple_history(void* sp, int which, double* out, int cap, int* len) {
    return guard([&] {
        Simple& s = *static_cast<Simple*>(sp);
        if (which > 3 && !s.turb) throw std::invalid_argument("laminar SIMPLE has no turbulence histories");
        const std::vector<double>& h = which < 3 ? s.momHist[which] : which == 3 ? s.pHist
                                     : which == 4 ? s.turb->kHist : s.turb->wHist;
        *len = static_cast<int>(h.size());
        std::memcpy(out, h.data(), sizeof(double) * std::min<int>(cap, static_cast<int>(h.size())));
    });
}

void* orc_piso_new(void* mesh, const orc_piso_config* cfg, const int* ubct, const double* ubcv, const int* pbct,
                   const double* pbcv) {
    Piso* s = nullptr;
    guard([&] {
        const Mesh& m = *static_cast<Mesh*>(mesh);
        PisoConfig c;
        c.rho = cfg->rho; c.mu = cfg->mu;
        c.mInterp = cfg->m_interp; c.mGrad = cfg->m_grad; c.mConv = cfg->m_conv; c.mDiff = cfg->m_diff;
        c.mCoefInterp = cfg->m_coef_interp; c.mCoefGrad = cfg->m_coef_grad;
        c.pInterp = cfg->p_interp; c.pGrad = cfg->p_grad; c.pDiff = cfg->p_diff;
        c.pCoefInterp = cfg->p_coef_interp; c.pCoefGrad = cfg->p_coef_grad;
        c.timeMethod = cfg->time_method;
        c.maxIterations = cfg->max_iterations; c.nCorrectors = cfg->n_correctors;
        c.maxCorrectors = cfg->max_correctors; c.nonOrthCorr = cfg->non_orth_corr;
        c.alphaU = cfg->alpha_u; c.couplingTol = cfg->coupling_tol;
        c.tolMass = cfg->tol_mass; c.tolU = cfg->tol_u; c.tolMom = cfg->tol_mom; c.tolP = cfg->tol_p;
        c.mom = toConfig(&cfg->mom); c.pre = toConfig(&cfg->pre);
        c.potentialInit = cfg->initialize_potential_flow;
        c.initialNonOrthCorr = cfg->initial_non_orth_corr;
        std::unique_ptr<Piso> owned(new Piso(m, c));
        std::vector<BC<V3>> ub(m.np);
        std::vector<BC<double>> pb(m.np);
        for (Index p = 0; p < m.np; ++p) {
            ub[p].type = ubct[p]; ub[p].value = {ubcv[3 * p], ubcv[3 * p + 1], ubcv[3 * p + 2]};
            pb[p].type = pbct[p]; pb[p].value = pbcv[p];
        }
        owned->setBoundaries(ub, pb);
        s = owned.release();
    });
    return s;
}
void orc_piso_free(void* s) { delete static_cast<Piso*>(s); }
int orc_piso_enable_komega(void* sp, const orc_komega_config* kw, const int* kbct, const double* kbcv,
                           const int* wbct, const double* wbcv, const orc_piso_turbulence* t) {
    return guard([&] {
        Piso& s = *static_cast<Piso*>(sp);
        if (!kw || !t || !kbct || !kbcv || !wbct || !wbcv) throw std::invalid_argument("null argument");
        s.enableKOmega(komegaConfig(kw), scalarBcs(s.m, kbct, kbcv), scalarBcs(s.m, wbct, wbcv),
                       PisoTurbulence{t->coupling, t->correctors, t->max_iterations});
    });
}
int orc_piso_set_turbulence(void* sp, const double* k, const double* omega) {
    return guard([&] {
        Piso& s = *static_cast<Piso*>(sp);
        if (!s.turb) throw std::invalid_argument("laminar PISO has no turbulence fields");
        for (Index i = 0; i < s.m.nc; ++i) {
            if (k) s.turb->k.v[i] = k[i];
            if (omega) s.turb->omega.v[i] = omega[i];
        }
    });
}
int orc_piso_get_turbulence(void* sp, double* k, double* omega, double* mut) {
    return guard([&] {
        Piso& s = *static_cast<Piso*>(sp);
        if (!s.turb) throw std::invalid_argument("laminar PISO has no turbulence fields");
        if (k) store(s.turb->k.v, k);
        if (omega) store(s.turb->omega.v, omega);
        if (mut) store(s.turb->mut, mut);
    });
}
int orc_piso_set_fields(void* sp, const double* U, const double* p) {
    return guard([&] {
        Piso& s = *static_cast<Piso*>(sp);
        for (Index i = 0; i < s.m.nc; ++i) {
            if (U) s.U.v[i] = {U[3 * i], U[3 * i + 1], U[3 * i + 2]};
            if (p) s.p.v[i] = p[i];
        }
    });
}
int orc_piso_start(void* sp) { return guard([&] { static_cast<Piso*>(sp)->start(); }); }
int orc_piso_initial_report(void* sp, orc_piso_initial* o) {
    return guard([&] {
        const InitReport& r = static_cast<Piso*>(sp)->initial;
        *o = orc_piso_initial{r.mass, r.uMax, r.solves, r.pIters, r.done, r.status};
    });
}
int orc_piso_step(void* sp, double dt, orc_step_report* o) {
    return guard([&] {
        const StepReport r = static_cast<Piso*>(sp)->step(dt);
        o->dt = r.dt; o->co_max = r.coMax; o->co_mean = r.coMean; o->mass = r.mass; o->dU = r.dU; o->rU = r.rU;
        o->dP = r.dP; o->r_coupling_first = r.rCouplingFirst; o->r_coupling = r.rCoupling; o->lin_u = r.linU;
        o->lin_p = r.linP; o->passes = r.passes; o->correctors = r.correctors;
        for (int c = 0; c < 3; ++c) o->mom_iters[c] = r.momIters[c];
        o->p_iters = r.pIters; o->accepted = r.accepted; o->converged = r.converged; o->linear_ok = r.linearOk;
        o->status = r.status;
        o->d_turb = r.dTurb; o->r_turb = r.rTurb; o->k_iters = r.kIters; o->w_iters = r.wIters;
        o->turb_linear = r.turbLinear; o->turb_converged = r.turbConverged; o->turb_corrections = r.turbCorrections;
    });
}
int orc_piso_get_fields(void* sp, double* U, double* p, double* phi) {
    return guard([&] {
        Piso& s = *static_cast<Piso*>(sp);
        if (U) store(s.U.v, U);
        if (p) store(s.p.v, p);
        if (phi) store(s.phi, phi);
    });
}
int orc_piso_temporal_flux(void* sp, double* out) {
    return guard([&] {
        Piso& s = *static_cast<Piso*>(sp);
        if (s.lastTemporalFlux.empty()) throw std::invalid_argument("no step taken");
        store(s.lastTemporalFlux, out);
    });
}
int orc_piso_history(void* sp, int which, double* out, int cap, int* len) {
    return guard([&] {
        Piso& s = *static_cast<Piso*>(sp);
        const std::vector<double>& h = which < 3 ? s.momHist[which] : s.pHist;
        *len = static_cast<int>(h.size());
        std::memcpy(out, h.data(), sizeof(double) * std::min<int>(cap, static_cast<int>(h.size())));
    });
}
int orc_time_weights(int method, int levels, double dt, double previous_dt, double* out5) {
    return guard([&] {
        const TimeWeights w = timeWeights(method, levels, dt, previous_dt);
        out5[0] = w.a0; out5[1] = w.a1; out5[2] = w.a2;
        temporalFluxWeights(method, levels, dt, previous_dt, out5[3], out5[4]);
    });
}

}  // extern "C"
