// C ABI of the device library (include/bsim.h): contexts, operators, linear
// systems and the SIMPLE driver.  Host arrays are in the caller's numbering.
#include "capi_types.hpp"
#include "errors.hpp"

#include <cstring>

struct bsim_system {
    bsim_ctx* ctx;
    std::unique_ptr<bsim::DeviceSystem> sys;
    std::vector<double> hist[3];  // per right-hand side of the last solve
    int nrhs = 1;
};
struct bsim_simple {
    bsim_ctx* ctx;
    std::unique_ptr<bsim::SimpleSolver> s;
};
struct bsim_piso {
    bsim_ctx* ctx;
    std::unique_ptr<bsim::PisoSolver> s;
};

using namespace bsim;

namespace {

struct OpField {
    DBuf<double> v, bcv, ctx;
    DBuf<int> bct;
    FieldArg arg;
};

OpField makeField(Context& C, int ncomp, const double* x, const int* bct, const double* bcv, const double* ctx) {
    if (ncomp != 1 && ncomp != 3) throw std::invalid_argument("fields have 1 or 3 components");
    OpField f;
    hipStream_t s = C.stream;
    const int np = C.mesh->np;
    for (int p = 0; p < np; ++p)
        if (bct[p] < 0 || bct[p] > 4) throw std::invalid_argument("invalid boundary type");
    f.v.upload(C.cellsIn(x, ncomp), s);
    f.bct.upload(bct, np, s);
    f.bcv.upload(bcv, static_cast<std::size_t>(np) * ncomp, s);
    if (ctx) f.ctx.upload(C.facesIn(ctx, 1), s);
    f.arg = FieldArg{f.v.p, f.bct.p, f.bcv.p, ctx ? f.ctx.p : nullptr, nullptr};
    return f;
}

void syncAndCheck(Context& C) {
    BSIM_HIP(hipStreamSynchronize(C.stream));
    C.checkErrors();
}

// CSR off-diagonal arrays -> LDU upper/lower in caller face numbering (accumulating).
void lduOut(Context& C, const std::vector<double>& dg, const std::vector<double>& off, double* diag, double* upper,
            double* lower) {
    const DeviceMesh& M = *C.mesh;
    for (int i = 0; i < M.nc; ++i) diag[M.cellNewToOld[i]] += dg[i];
    for (int i = 0; i < M.nc; ++i)
        for (int k = M.hmptr[i]; k < M.hmptr[i + 1]; ++k) {
            const int f = M.hmface[k];
            const double v = off[M.csrToSliced[k]];
            if (M.hown[f] == i) upper[M.faceNewToOld[f]] += v;
            else lower[M.faceNewToOld[f]] += v;
        }
}

}  // namespace

namespace bsim {

SimpleCfg toSimpleCfg(const bsim_simple_config& cfg) {
    SimpleCfg c;
    c.rho = cfg.rho; c.mu = cfg.mu;
    c.mInterp = cfg.m_interp; c.mGrad = cfg.m_grad; c.mConv = cfg.m_conv; c.mDiff = cfg.m_diff;
    c.mCoefInterp = cfg.m_coef_interp; c.mCoefGrad = cfg.m_coef_grad;
    c.pInterp = cfg.p_interp; c.pGrad = cfg.p_grad; c.pDiff = cfg.p_diff;
    c.pCoefInterp = cfg.p_coef_interp; c.pCoefGrad = cfg.p_coef_grad;
    c.alphaU = cfg.alpha_u; c.alphaP = cfg.alpha_p; c.nonOrthCorr = cfg.non_orth_corr;
    c.tolMass = cfg.tol_mass; c.tolU = cfg.tol_u; c.tolMom = cfg.tol_mom; c.tolP = cfg.tol_p;
    c.mom = toSolverCfg(&cfg.mom);
    c.pre = toSolverCfg(&cfg.pre);
    if (c.nonOrthCorr < 0 || c.nonOrthCorr > 20) throw std::invalid_argument("nonOrthogonalCorrections must lie in [0, 20]");
    return c;
}

KOmegaCfg toKOmegaCfg(const bsim_komega_config& k) {
    KOmegaCfg c;
    c.betaStar = k.beta_star; c.beta = k.beta; c.gamma = k.gamma; c.sigmaK = k.sigma_k;
    c.sigmaOmega = k.sigma_omega; c.kMin = k.k_min; c.omegaMin = k.omega_min;
    c.viscosityRatioMax = k.viscosity_ratio_max; c.relaxation = k.relaxation; c.tolerance = k.tolerance;
    c.kOpts = {k.k_interp, k.k_grad, k.k_conv, k.k_diff, k.k_coef_interp, k.k_coef_grad};
    c.wOpts = {k.w_interp, k.w_grad, k.w_conv, k.w_diff, k.w_coef_interp, k.w_coef_grad};
    c.kSolver = toSolverCfg(&k.k_solver);
    c.wSolver = toSolverCfg(&k.w_solver);
    return c;
}

PisoCfg toPisoCfg(const bsim_piso_config& cfg) {
    PisoCfg c;
    c.rho = cfg.rho; c.mu = cfg.mu;
    c.mInterp = cfg.m_interp; c.mGrad = cfg.m_grad; c.mConv = cfg.m_conv; c.mDiff = cfg.m_diff;
    c.mCoefInterp = cfg.m_coef_interp; c.mCoefGrad = cfg.m_coef_grad;
    c.pInterp = cfg.p_interp; c.pGrad = cfg.p_grad; c.pDiff = cfg.p_diff;
    c.pCoefInterp = cfg.p_coef_interp; c.pCoefGrad = cfg.p_coef_grad;
    c.timeMethod = cfg.time_method;
    c.maxIterations = cfg.max_iterations; c.nCorrectors = cfg.n_correctors;
    c.maxCorrectors = cfg.max_correctors; c.nonOrthCorr = cfg.non_orth_corr;
    c.alphaU = cfg.alpha_u; c.couplingTol = cfg.coupling_tol;
    c.tolMass = cfg.tol_mass; c.tolU = cfg.tol_u; c.tolMom = cfg.tol_mom; c.tolP = cfg.tol_p;
    c.mom = toSolverCfg(&cfg.mom);
    c.pre = toSolverCfg(&cfg.pre);
    c.potentialInit = cfg.initialize_potential_flow;
    c.initialNonOrthCorr = cfg.initial_non_orth_corr;
    return c;
}

}  // namespace bsim

extern "C" {

int bsim_set_graph_replay(int enable) {
    return guarded([&] { setGraphReplay(enable != 0); });
}

int bsim_ctx_create(const bsim_mesh* mesh, int renumber, int device, bsim_ctx** out) {
    return guarded([&] {
        if (!mesh) throw std::invalid_argument("mesh is null");
        auto* c = new bsim_ctx;
        try { c->c = std::make_unique<Context>(mesh->d, renumber != 0, device); }
        catch (...) { delete c; throw; }
        *out = c;
    });
}

void bsim_ctx_destroy(bsim_ctx* ctx) { delete ctx; }

int bsim_ctx_sizes(bsim_ctx* ctx, int* nc, int* nf, int* nfi, int* nnz, int* orthogonal) {
    return guarded([&] {
        const DeviceMesh& M = *ctx->c->mesh;
        if (nc) *nc = M.nc;
        if (nf) *nf = M.nf;
        if (nfi) *nfi = M.nfi;
        if (nnz) *nnz = M.nnzOff;
        if (orthogonal) *orthogonal = M.orthogonal ? 1 : 0;
    });
}

int bsim_ctx_geometry(bsim_ctx* ctx, double* C_, double* V, double* Cf, double* Sf, double* Af, double* nrm, double* w,
                      double* orth, double* kf, double* skew) {
    return guarded([&] {
        Context& C = *ctx->c;
        const DeviceMesh& M = *C.mesh;
        hipStream_t s = C.stream;
        if (C_) C.cellsOut(M.C.toHost(s), C_, 3);
        if (V) C.cellsOut(M.V.toHost(s), V, 1);
        if (Cf) C.facesOut(M.Cf.toHost(s), Cf, 3);
        if (Sf) C.facesOut(M.Sf.toHost(s), Sf, 3);
        if (Af) C.facesOut(M.Af.toHost(s), Af, 1);
        if (nrm) C.facesOut(M.nrm.toHost(s), nrm, 3);
        if (w) C.facesOut(M.w.toHost(s), w, 1);
        if (orth) C.facesOut(M.orth.toHost(s), orth, 1);
        if (kf) C.facesOut(M.kf.toHost(s), kf, 3);
        if (skew) C.facesOut(M.skew.toHost(s), skew, 3);
    });
}

int bsim_op_gradient(bsim_ctx* ctx, int ncomp, const double* x, const int* bct, const double* bcv,
                     const double* flux_ctx, int method, double* out) {
    return guarded([&] {
        Context& C = *ctx->c;
        const DeviceMesh& M = *C.mesh;
        OpField f = makeField(C, ncomp, x, bct, bcv, flux_ctx);
        DBuf<double> g(3 * ncomp * static_cast<std::size_t>(M.nc)), work(3 * ncomp * static_cast<std::size_t>(M.nc));
        ops::gradient(M.view(), M.orthogonal, ncomp, f.arg, method, g.p, work.p, C.err.p, C.stream);
        syncAndCheck(C);
        C.cellsOut(g.toHost(C.stream), out, 3 * ncomp);
    });
}

int bsim_op_interpolate(bsim_ctx* ctx, int ncomp, const double* x, const int* bct, const double* bcv,
                        const double* flux_ctx, int method, int gmethod, double* out) {
    return guarded([&] {
        Context& C = *ctx->c;
        const DeviceMesh& M = *C.mesh;
        OpField f = makeField(C, ncomp, x, bct, bcv, flux_ctx);
        DBuf<double> g(3 * ncomp * static_cast<std::size_t>(M.nc)), work(3 * ncomp * static_cast<std::size_t>(M.nc));
        DBuf<double> o(static_cast<std::size_t>(ncomp) * M.nf);
        if (method == kCorrected) ops::gradient(M.view(), M.orthogonal, ncomp, f.arg, gmethod, g.p, work.p, C.err.p, C.stream);
        ops::interpolate(M.view(), ncomp, f.arg, method, g.p, o.p, C.err.p, C.stream);
        syncAndCheck(C);
        C.facesOut(o.toHost(C.stream), out, ncomp);
    });
}

int bsim_op_flux(bsim_ctx* ctx, const double* U, const int* bct, const double* bcv, const double* flux_ctx, int method,
                 int gmethod, double* out) {
    return guarded([&] {
        Context& C = *ctx->c;
        const DeviceMesh& M = *C.mesh;
        if (method != kLinear && method != kCorrected) throw std::invalid_argument("unsupported flux interpolation method");
        OpField f = makeField(C, 3, U, bct, bcv, flux_ctx);
        DBuf<double> seed;
        if (!flux_ctx) {  // operators.cpp:977-987: seed directions from owner velocity
            seed.alloc(M.nf);
            ops::seedFluxContext(M.view(), f.v.p, seed.p, C.stream);
            f.arg.ctx = seed.p;
        }
        DBuf<double> g(9 * static_cast<std::size_t>(M.nc)), work(9 * static_cast<std::size_t>(M.nc)), o(M.nf);
        const double* gp = nullptr;
        if (method == kCorrected) { ops::gradient(M.view(), M.orthogonal, 3, f.arg, gmethod, g.p, work.p, C.err.p, C.stream); gp = g.p; }
        ops::flux(M.view(), f.arg, gp, o.p, C.err.p, C.stream);
        syncAndCheck(C);
        C.facesOut(o.toHost(C.stream), out, 1);
    });
}

int bsim_op_normal_gradient(bsim_ctx* ctx, const double* x, const int* bct, const double* bcv, int gmethod,
                            int dmethod, double* out) {
    return guarded([&] {
        Context& C = *ctx->c;
        const DeviceMesh& M = *C.mesh;
        OpField f = makeField(C, 1, x, bct, bcv, nullptr);
        DBuf<double> g(3 * static_cast<std::size_t>(M.nc)), work(3 * static_cast<std::size_t>(M.nc)), o(M.nf);
        ops::gradient(M.view(), M.orthogonal, 1, f.arg, gmethod, g.p, work.p, C.err.p, C.stream);
        ops::normalGradient(M.view(), f.arg, g.p, dmethod, o.p, C.err.p, C.stream);
        syncAndCheck(C);
        C.facesOut(o.toHost(C.stream), out, 1);
    });
}

int bsim_op_divergence_face(bsim_ctx* ctx, const double* phi, double* out) {
    return guarded([&] {
        Context& C = *ctx->c;
        const DeviceMesh& M = *C.mesh;
        DBuf<double> ph, o(M.nc);
        ph.upload(C.facesIn(phi, 1), C.stream);
        ops::divergenceFace(M.view(), ph.p, o.p, C.stream);
        syncAndCheck(C);
        C.cellsOut(o.toHost(C.stream), out, 1);
    });
}

int bsim_op_explicit_laplacian(bsim_ctx* ctx, const double* x, const int* bct, const double* bcv, const double* kf,
                               double k, int gmethod, int dmethod, double* out) {
    return guarded([&] {
        Context& C = *ctx->c;
        const DeviceMesh& M = *C.mesh;
        OpField f = makeField(C, 1, x, bct, bcv, nullptr);
        DBuf<double> g(3 * static_cast<std::size_t>(M.nc)), work(3 * static_cast<std::size_t>(M.nc)), o(M.nc), kd;
        if (kf) kd.upload(C.facesIn(kf, 1), C.stream);
        if (dmethod != kOrthogonal) ops::gradient(M.view(), M.orthogonal, 1, f.arg, gmethod, g.p, work.p, C.err.p, C.stream);
        ops::explicitLaplacian(M.view(), f.arg, kf ? kd.p : nullptr, k, g.p, dmethod, o.p, C.err.p, C.stream);
        syncAndCheck(C);
        C.cellsOut(o.toHost(C.stream), out, 1);
    });
}

int bsim_op_explicit_convection(bsim_ctx* ctx, int ncomp, const double* x, const int* bct, const double* bcv,
                                const double* phi, int conv, int interp, int gmethod, double* out) {
    return guarded([&] {
        Context& C = *ctx->c;
        const DeviceMesh& M = *C.mesh;
        OpField f = makeField(C, ncomp, x, bct, bcv, phi);
        DBuf<double> g(3 * ncomp * static_cast<std::size_t>(M.nc)), work(3 * ncomp * static_cast<std::size_t>(M.nc));
        DBuf<double> o(static_cast<std::size_t>(ncomp) * M.nc);
        if (conv == kLinearUpwind || (conv == kCentral && interp == kCorrected))
            ops::gradient(M.view(), M.orthogonal, ncomp, f.arg, gmethod, g.p, work.p, C.err.p, C.stream);
        ops::explicitConvection(M.view(), ncomp, f.arg, f.ctx.p, conv, interp, g.p, o.p, C.err.p, C.stream);
        syncAndCheck(C);
        C.cellsOut(o.toHost(C.stream), out, ncomp);
    });
}

int bsim_op_assemble_convection(bsim_ctx* ctx, int ncomp, const double* x, const int* bct, const double* bcv,
                                const double* phi, int conv, int interp, int gmethod, double scale, double* diag,
                                double* upper, double* lower, double* source) {
    return guarded([&] {
        Context& C = *ctx->c;
        const DeviceMesh& M = *C.mesh;
        OpField f = makeField(C, ncomp, x, bct, bcv, phi);
        DBuf<double> g(3 * ncomp * static_cast<std::size_t>(M.nc)), work(3 * ncomp * static_cast<std::size_t>(M.nc));
        DBuf<double> dg(M.nc), off(std::max(M.nnzSliced, 1)), src(static_cast<std::size_t>(ncomp) * M.nc);
        dg.zero(C.stream); off.zero(C.stream); src.zero(C.stream);
        if (conv == kLinearUpwind || (conv == kCentral && interp == kCorrected))
            ops::gradient(M.view(), M.orthogonal, ncomp, f.arg, gmethod, g.p, work.p, C.err.p, C.stream);
        ops::assembleConvection(M.view(), ncomp, f.arg, f.ctx.p, conv, interp, g.p, scale, dg.p, off.p, src.p, C.err.p,
                                C.stream);
        syncAndCheck(C);
        lduOut(C, dg.toHost(C.stream), off.toHost(C.stream), diag, upper, lower);
        std::vector<double> s(static_cast<std::size_t>(ncomp) * M.nc, 0.0);
        C.cellsOut(src.toHost(C.stream), s.data(), ncomp);
        for (std::size_t i = 0; i < s.size(); ++i) source[i] += s[i];
    });
}

int bsim_op_assemble_diffusion(bsim_ctx* ctx, int ncomp, const double* x, const int* bct, const double* bcv,
                               const double* flux_ctx, const double* kf, int gmethod, int dmethod, double* diag,
                               double* upper, double* lower, double* source) {
    return guarded([&] {
        Context& C = *ctx->c;
        const DeviceMesh& M = *C.mesh;
        OpField f = makeField(C, ncomp, x, bct, bcv, flux_ctx);
        DBuf<double> g(3 * ncomp * static_cast<std::size_t>(M.nc)), work(3 * ncomp * static_cast<std::size_t>(M.nc));
        DBuf<double> dg(M.nc), off(std::max(M.nnzSliced, 1)), src(static_cast<std::size_t>(ncomp) * M.nc), kd;
        dg.zero(C.stream); off.zero(C.stream); src.zero(C.stream);
        kd.upload(C.facesIn(kf, 1), C.stream);
        if (dmethod != kOrthogonal) ops::gradient(M.view(), M.orthogonal, ncomp, f.arg, gmethod, g.p, work.p, C.err.p, C.stream);
        ops::assembleDiffusion(M.view(), ncomp, f.arg, kd.p, g.p, dmethod, 1.0, dg.p, off.p, src.p, C.err.p, C.stream);
        syncAndCheck(C);
        lduOut(C, dg.toHost(C.stream), off.toHost(C.stream), diag, upper, lower);
        std::vector<double> s(static_cast<std::size_t>(ncomp) * M.nc, 0.0);
        C.cellsOut(src.toHost(C.stream), s.data(), ncomp);
        for (std::size_t i = 0; i < s.size(); ++i) source[i] += s[i];
    });
}

int bsim_system_create(bsim_ctx* ctx, bsim_system** out) {
    return guarded([&] {
        const DeviceMesh& M = *ctx->c->mesh;
        auto* s = new bsim_system;
        s->ctx = ctx;
        s->sys = std::make_unique<DeviceSystem>(ctx->c->stream, meshPattern(M));
        *out = s;
    });
}

int bsim_system_solve_ldu(bsim_system* sys, const double* diag, const double* upper, const double* lower, int nrhs,
                          const double* source, double* x, const bsim_solver_config* cfg, bsim_solve_result* results) {
    return guarded([&] {
        if (nrhs != 1 && nrhs != 3) throw std::invalid_argument("linear systems have one or three right-hand sides");
        Context& C = *sys->ctx->c;
        const DeviceMesh& M = *C.mesh;
        hipStream_t s = C.stream;
        const SolverCfg sc = toSolverCfg(cfg);
        DeviceSystem& D = *sys->sys;
        // LDU -> CSR values (petsc_assembly_plan.cpp:62-87, row owner holds upper).
        std::vector<double> dg = C.cellsIn(diag, 1), off(M.nnzSliced, 0.0);
        for (int i = 0; i < M.nc; ++i)
            for (int k = M.hmptr[i]; k < M.hmptr[i + 1]; ++k) {
                const int f = M.hmface[k];
                off[M.csrToSliced[k]] = M.hown[f] == i ? upper[M.faceNewToOld[f]] : lower[M.faceNewToOld[f]];
            }
        BSIM_HIP(hipMemcpyAsync(D.candDiag(), dg.data(), sizeof(double) * M.nc, hipMemcpyHostToDevice, s));
        if (M.nnzSliced) BSIM_HIP(hipMemcpyAsync(D.candVal(), off.data(), sizeof(double) * M.nnzSliced, hipMemcpyHostToDevice, s));
        D.update();
        D.prepare(sc);
        const std::vector<double> b = C.cellsIn(source, nrhs), x0 = C.cellsIn(x, nrhs);
        std::vector<double> bc(M.nc), xc(M.nc), xout(x0.size());
        DBuf<double> db[3], dx[3];
        for (int comp = 0; comp < nrhs; ++comp) {
            for (int i = 0; i < M.nc; ++i) { bc[i] = b[static_cast<std::size_t>(i) * nrhs + comp]; xc[i] = x0[static_cast<std::size_t>(i) * nrhs + comp]; }
            db[comp].upload(bc, s);
            dx[comp].upload(xc, s);
        }
        SolveRes r[3];
        if (nrhs == 1) {
            r[0] = D.solve(db[0].p, dx[0].p, 0);
            sys->hist[0] = D.hist;
        } else {
            const double* const b3[3] = {db[0].p, db[1].p, db[2].p};
            double* const x3[3] = {dx[0].p, dx[1].p, dx[2].p};
            D.solve3(b3, x3, r, sys->hist);
        }
        for (int comp = 0; comp < nrhs; ++comp) {
            const std::vector<double> h = dx[comp].toHost(s);
            for (int i = 0; i < M.nc; ++i) xout[static_cast<std::size_t>(i) * nrhs + comp] = h[i];
            results[comp] = bsim_solve_result{r[comp].status, r[comp].iterations, r[comp].initial, r[comp].final_,
                                              r[comp].relative};
        }
        sys->nrhs = nrhs;
        C.cellsOut(xout, x, nrhs);
    });
}

int bsim_system_history(bsim_system* sys, double* out, int cap, int* len) {
    return guarded([&] {
        const auto& h = sys->hist[sys->nrhs - 1];
        *len = static_cast<int>(h.size());
        if (out) std::memcpy(out, h.data(), sizeof(double) * std::min<std::size_t>(cap, h.size()));
    });
}

int bsim_system_counters(bsim_system* sys, uint64_t* o) {
    return guarded([&] {
        const auto& c = sys->sys->counters;
        const uint64_t v[10] = {c.matrixUpdates, c.rhsOnly, c.pcSetups, c.pcReuses, c.retries, c.solves,
                                c.krylovIterations, c.trueResidualChecks, c.spmvs, c.pcApplications};
        std::memcpy(o, v, sizeof(v));
    });
}

}  // extern "C"

namespace {
// Names of petsc_backend.cpp:30-60 (kspName, pcName, solveStatusName).
const char* kspName(int solver) {
    return solver == 0 ? "cg" : solver == 1 ? "bcgs" : solver == 2 ? "gmres" : solver == 3 ? "fgmres" : "unknown";
}
const char* pcName(int pc) {
    return pc == 0 ? "none" : pc == 1 ? "jacobi" : pc == 2 ? "amg" : pc == 3 ? "bjacobi" : pc == 4 ? "icc" : pc == 5 ? "asm" : "unknown";
}
const char* statusName(int st) {
    return st == kConverged ? "converged" : st == kMaxIterations ? "maxIterations" : st == kNumericalFailure ? "numericalFailure" : "";
}
void copyName(char* dst, std::size_t cap, const char* src) {
    std::strncpy(dst, src, cap - 1);
    dst[cap - 1] = '\0';
}
}  // namespace

bsim_equation_counters bsim::equationCounters(const DeviceSystem& D, const char* identity) {
    bsim_equation_counters o{};
    const SystemCounters& c = D.counters;
    copyName(o.identity, sizeof(o.identity), identity);
    copyName(o.ksp_type, sizeof(o.ksp_type), kspName(D.cfg.solver));
    copyName(o.pc_type, sizeof(o.pc_type), pcName(D.cfg.pc));
    copyName(o.last_status, sizeof(o.last_status), statusName(c.lastStatus));
    o.matrix_pattern_builds = 1;
    o.matrix_value_updates = c.matrixUpdates;
    o.rhs_only_solves = c.rhsOnly;
    o.linear_solves = c.solves;
    o.krylov_iterations = c.krylovIterations;
    o.preconditioner_setups = c.pcSetups;
    o.preconditioner_reuses = c.pcReuses;
    o.preconditioner_retries = c.retries;
    o.true_residual_checks = c.trueResidualChecks;
    o.sparse_matvecs = c.spmvs;
    o.preconditioner_applications = c.pcApplications;
    o.last_initial_residual = c.lastInitial;
    o.last_final_residual = c.lastFinal;
    o.last_relative_residual = c.lastRelative;
    o.pattern_build_seconds = c.patternSeconds;
    o.matrix_update_seconds = c.matrixUpdateSeconds;
    o.preconditioner_seconds = c.setupSeconds;
    o.rhs_seconds = c.rhsSeconds;
    o.solve_seconds = c.solveSeconds;
    o.residual_check_seconds = c.residualSeconds;
    return o;
}

namespace {
// Sums the systems into the run record (petsc_backend.cpp:236-286 keeps the
// totals and the per-equation snapshots side by side).
void addSystem(bsim_performance& t, const bsim_equation_counters& e) {
    t.linear_solves += e.linear_solves;
    t.krylov_iterations += e.krylov_iterations;
    t.sparse_matvecs += e.sparse_matvecs;
    t.equation_assemblies += e.matrix_value_updates + e.rhs_only_solves;
    t.preconditioner_setups += e.preconditioner_setups;
    t.preconditioner_applications += e.preconditioner_applications;
    t.preconditioner_seconds += e.preconditioner_seconds;
    t.linear_solve_seconds += e.solve_seconds;
    t.matrix_pattern_builds += e.matrix_pattern_builds;
    t.matrix_value_updates += e.matrix_value_updates;
    t.rhs_only_solves += e.rhs_only_solves;
    t.true_residual_checks += e.true_residual_checks;
    t.matrix_update_seconds += e.matrix_update_seconds;
    t.rhs_seconds += e.rhs_seconds;
    t.residual_check_seconds += e.residual_check_seconds;
    t.pattern_build_seconds += e.pattern_build_seconds;
    t.preconditioner_reuses += e.preconditioner_reuses;
    t.preconditioner_retries += e.preconditioner_retries;
}
void runPerformance(const std::vector<std::pair<const DeviceSystem*, const char*>>& systems, const RunClock& run,
                    std::uint64_t reductions, double assemblySeconds, bsim_performance* total,
                    bsim_equation_counters* out, int cap, int* count) {
    bsim_performance t{};
    int k = 0;
    for (const auto& s : systems) {
        const bsim_equation_counters e = equationCounters(*s.first, s.second);
        addSystem(t, e);
        if (out && k < cap) out[k] = e;
        ++k;
    }
    t.global_reductions = reductions;
    t.elapsed_seconds = run.elapsed();
    t.assembly_seconds = assemblySeconds;
    t.reported_iteration_intervals = run.intervals;
    t.iteration_compute_seconds = run.window;
    if (total) *total = t;
    if (count) *count = k;
}
}  // namespace
extern "C" {

int bsim_system_performance(bsim_system* sys, bsim_equation_counters* out) {
    return guarded([&] {
        if (!sys || !out) throw std::invalid_argument("null argument");
        *out = equationCounters(*sys->sys, "");
    });
}

void bsim_system_destroy(bsim_system* sys) { delete sys; }

namespace {
struct StandaloneCsr {
    hipStream_t stream = nullptr;
    DBuf<int> ptr, col;
    std::unique_ptr<DeviceSystem> sys;
    StandaloneCsr(int device, int n, const int* p, const int* c, const double* d, const double* v) {
        int count = 0;
        BSIM_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw std::invalid_argument("device index is out of range");
        if (n <= 0) throw std::invalid_argument("matrix is empty");
        BSIM_HIP(hipSetDevice(device));
        {
            std::lock_guard<std::recursive_mutex> rl(pool::runtimeLock());
            BSIM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        }
        std::vector<int> hp(p, p + n + 1), hc(c, c + p[n]);
        for (int k = 0; k < p[n]; ++k)
            if (hc[k] < 0 || hc[k] >= n) throw std::invalid_argument("column index out of range");
        ptr.upload(hp, stream);
        col.upload(hc, stream);
        Pattern P;
        P.n = n;
        P.nnzStore = p[n];
        P.K = 0;
        P.ptr = ptr.p;
        P.col = col.p;
        P.sliceOff = nullptr;
        P.hptr = hp;
        P.hcol = hc;
        sys = std::make_unique<DeviceSystem>(stream, P);
        BSIM_HIP(hipMemcpyAsync(sys->candDiag(), d, sizeof(double) * n, hipMemcpyHostToDevice, stream));
        if (p[n]) BSIM_HIP(hipMemcpyAsync(sys->candVal(), v, sizeof(double) * p[n], hipMemcpyHostToDevice, stream));
        sys->update();
    }
    ~StandaloneCsr() {
        sys.reset();
        ptr.release();
        col.release();
        if (stream) {
            std::lock_guard<std::recursive_mutex> rl(pool::runtimeLock());
            (void)hipStreamDestroy(stream);
        }
    }
};
}  // namespace

int bsim_csr_solve(int device, int n, const int* ptr, const int* col, const double* diag, const double* val,
                   const double* b, double* x, const bsim_solver_config* cfg, bsim_solve_result* res, double* hist,
                   int cap, int* len) {
    return guarded([&] {
        StandaloneCsr S(device, n, ptr, col, diag, val);
        S.sys->prepare(toSolverCfg(cfg));
        DBuf<double> db, dx;
        db.upload(b, n, S.stream);
        dx.upload(x, n, S.stream);
        const SolveRes r = S.sys->solve(db.p, dx.p, 0);
        dx.download(x, n, S.stream);
        BSIM_HIP(hipStreamSynchronize(S.stream));
        *res = bsim_solve_result{r.status, r.iterations, r.initial, r.final_, r.relative};
        if (len) *len = static_cast<int>(S.sys->hist.size());
        if (hist) std::memcpy(hist, S.sys->hist.data(), sizeof(double) * std::min<std::size_t>(cap, S.sys->hist.size()));
    });
}

int bsim_csr_spmv(int device, int n, const int* ptr, const int* col, const double* diag, const double* val,
                  const double* x, double* y) {
    return guarded([&] {
        StandaloneCsr S(device, n, ptr, col, diag, val);
        DBuf<double> dx, dy(n);
        dx.upload(x, n, S.stream);
        S.sys->spmv(dx.p, dy.p);
        dy.download(y, n, S.stream);
        BSIM_HIP(hipStreamSynchronize(S.stream));
    });
}

int bsim_amg_hierarchy(int device, int n, const int* ptr, const int* col, const double* diag, const double* val,
                       const bsim_solver_config* cfg, int* nlevels, int* sizes, int* agg0, double* coarse_diag1) {
    return guarded([&] {
        StandaloneCsr S(device, n, ptr, col, diag, val);
        SolverCfg c = toSolverCfg(cfg);
        c.pc = 2;
        AmgDevice amg(S.stream);
        amg.setup(S.sys->view(), S.sys->pat, c);
        const std::vector<int> sz = amg.sizes();
        *nlevels = static_cast<int>(sz.size());
        std::memcpy(sizes, sz.data(), sizeof(int) * sz.size());
        const std::vector<int> a0 = amg.aggregates0();
        if (agg0 && !a0.empty()) std::memcpy(agg0, a0.data(), sizeof(int) * a0.size());
        if (coarse_diag1 && sz.size() > 1) {
            const std::vector<double> d1 = amg.coarseDiag(1);
            std::memcpy(coarse_diag1, d1.data(), sizeof(double) * d1.size());
        }
    });
}

int bsim_simple_create(bsim_ctx* ctx, const bsim_simple_config* cfg, const int* ubct, const double* ubcv,
                       const int* pbct, const double* pbcv, bsim_simple** out) {
    return guarded([&] {
        if (!ctx || !cfg || !out) throw std::invalid_argument("null argument");
        const SimpleCfg c = toSimpleCfg(*cfg);
        auto* s = new bsim_simple;
        s->ctx = ctx;
        try { s->s = std::make_unique<SimpleSolver>(*ctx->c, c, ubct, ubcv, pbct, pbcv); }
        catch (...) { delete s; throw; }
        *out = s;
    });
}

int bsim_simple_set_fields(bsim_simple* s, const double* U, const double* p) {
    return guarded([&] { s->s->setFields(U, p); });
}
int bsim_simple_enable_komega(bsim_simple* s, const bsim_komega_config* k, const int* kbct, const double* kbcv,
                              const int* wbct, const double* wbcv) {
    return guarded([&] {
        if (!k || !kbct || !kbcv || !wbct || !wbcv) throw std::invalid_argument("null argument");
        s->s->enableKOmega(toKOmegaCfg(*k), kbct, kbcv, wbct, wbcv);
    });
}
int bsim_simple_set_turbulence(bsim_simple* s, const double* k, const double* omega) {
    return guarded([&] {
        if (!s->s->turbulence()) throw std::invalid_argument("laminar SIMPLE has no turbulence fields");
        s->s->turbulence()->setFields(k, omega);
    });
}
int bsim_simple_get_turbulence(bsim_simple* s, double* k, double* omega, double* mut) {
    return guarded([&] {
        if (!s->s->turbulence()) throw std::invalid_argument("laminar SIMPLE has no turbulence fields");
        s->s->turbulence()->getFields(k, omega, mut);
    });
}
int bsim_simple_start(bsim_simple* s) { return guarded([&] { s->s->start(); }); }
int bsim_simple_iterate(bsim_simple* s, int n, bsim_iter_report* reports) {
    return guarded([&] {
        for (int k = 0; k < n; ++k) {
            const IterRep r = s->s->iterate();
            bsim_iter_report& o = reports[k];
            o.numerical_failure = r.failed;
            o.rU = r.rU; o.dU = r.dU; o.dP = r.dP; o.mass = r.mass; o.linP = r.linP;
            for (int c = 0; c < 3; ++c) o.mom_iters[c] = r.momIters[c];
            o.p_iters = r.pIters; o.mom_status = r.momStatus; o.p_status = r.pStatus;
            o.converged = r.converged; o.referenced = r.referenced;
            o.d_turb = r.dTurb; o.r_turb = r.rTurb; o.bounded_omega = r.boundedOmega;
            o.k_iters = r.kIters; o.w_iters = r.wIters; o.turb_linear = r.turbLinear;
            o.turb_converged = r.turbConverged;
            if (r.failed) {
                for (int j = k + 1; j < n; ++j) {
                    reports[j] = bsim_iter_report{};
                    reports[j].numerical_failure = 1;
                }
                break;
            }
        }
    });
}
int bsim_simple_get_fields(bsim_simple* s, double* U, double* p, double* phi) {
    return guarded([&] { s->s->getFields(U, p, phi); });
}
int bsim_simple_history(bsim_simple* s, int which, double* out, int cap, int* len) {
    return guarded([&] {
        if (which < 0 || which > 5) throw std::invalid_argument("history index is 0..5");
        if (which > 3 && !s->s->turbulence()) throw std::invalid_argument("laminar SIMPLE has no turbulence histories");
        const std::vector<double>& h = which < 3 ? s->s->momHist[which] : which == 3 ? s->s->pHist
                                     : which == 4 ? s->s->turbulence()->kHist : s->s->turbulence()->wHist;
        *len = static_cast<int>(h.size());
        if (out) std::memcpy(out, h.data(), sizeof(double) * std::min<std::size_t>(cap, h.size()));
    });
}
int bsim_simple_timing(bsim_simple* s, double* o) {
    return guarded([&] {
        const SimpleTiming& t = s->s->timing;
        const double v[9] = {t.assembly, t.momentumSolve, t.rhieChow, t.pressureAssembly, t.pressureSolve,
                             t.correction, t.convergence, static_cast<double>(t.iterations), t.turbulence};
        std::memcpy(o, v, sizeof(v));
    });
}
int bsim_simple_performance(bsim_simple* s, bsim_performance* total, bsim_equation_counters* systems, int cap,
                            int* count) {
    return guarded([&] {
        if (!s) throw std::invalid_argument("null argument");
        SimpleSolver& S = *s->s;
        std::vector<std::pair<const DeviceSystem*, const char*>> sys = {
            {&S.momentumSystem(), "vector:momentum:U"}, {&S.pressureSystem(), "scalar:pressureCorrection:pPrime"}};
        std::uint64_t reads = S.run.reductions;
        if (const KOmegaDevice* t = S.turbulence()) {
            if (t->kSystem()) sys.push_back({t->kSystem(), "scalar:kTransport:k"});
            if (t->omegaSystem()) sys.push_back({t->omegaSystem(), "scalar:omegaTransport:omega"});
            reads += t->reads;
        }
        runPerformance(sys, S.run, reads, S.timing.assembly + S.timing.pressureAssembly, total, systems, cap, count);
    });
}
int bsim_simple_time_spmv(bsim_simple* s, int which, int reps, int cold, double* seconds, int* n, int* nnz) {
    return guarded([&] {
        DeviceSystem& D = which == 0 ? s->s->momentumSystem() : s->s->pressureSystem();
        *seconds = D.timeSpmv(reps, cold != 0);
        *n = D.n;
        *nnz = D.nnz;
    });
}
int bsim_simple_probe_start(bsim_simple* s, int capacity) {
    return guarded([&] { s->s->pressureSystem().probe.start(capacity); });
}
int bsim_simple_probe_stop(bsim_simple* s, int* launches, double* seconds, double* bytes_per_launch) {
    return guarded([&] {
        KernelProbe& P = s->s->pressureSystem().probe;
        *seconds = P.stop(*launches);
        *bytes_per_launch = P.bytesPerLaunch;
    });
}
void bsim_simple_destroy(bsim_simple* s) { delete s; }

int bsim_piso_create(bsim_ctx* ctx, const bsim_piso_config* cfg, const int* ubct, const double* ubcv, const int* pbct,
                     const double* pbcv, bsim_piso** out) {
    return guarded([&] {
        if (!ctx || !cfg || !ubct || !ubcv || !pbct || !pbcv || !out) throw std::invalid_argument("null argument");
        const PisoCfg c = toPisoCfg(*cfg);
        auto* s = new bsim_piso;
        s->ctx = ctx;
        try { s->s = std::make_unique<PisoSolver>(*ctx->c, c, ubct, ubcv, pbct, pbcv); }
        catch (...) { delete s; throw; }
        *out = s;
    });
}
int bsim_piso_set_fields(bsim_piso* s, const double* U, const double* p) {
    return guarded([&] { s->s->setFields(U, p); });
}
int bsim_piso_start(bsim_piso* s) { return guarded([&] { s->s->start(); }); }
int bsim_piso_initial_report(bsim_piso* s, bsim_piso_initial* o) {
    return guarded([&] {
        if (!s || !o) throw std::invalid_argument("null argument");
        const InitRep& r = s->s->initial;
        *o = bsim_piso_initial{r.mass, r.uMax, r.solves, r.pIters, r.done, r.status};
    });
}
int bsim_piso_step(bsim_piso* s, double dt, bsim_step_report* o) {
    return guarded([&] {
        if (!o) throw std::invalid_argument("null report");
        const StepRep r = s->s->step(dt);
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
int bsim_piso_enable_komega(bsim_piso* s, const bsim_komega_config* k, const int* kbct, const double* kbcv,
                            const int* wbct, const double* wbcv, const bsim_piso_turbulence* t) {
    return guarded([&] {
        if (!s || !k || !kbct || !kbcv || !wbct || !wbcv || !t) throw std::invalid_argument("null argument");
        s->s->enableKOmega(toKOmegaCfg(*k), kbct, kbcv, wbct, wbcv,
                           PisoTurbCfg{t->coupling, t->correctors, t->max_iterations});
    });
}
int bsim_piso_set_turbulence(bsim_piso* s, const double* k, const double* omega) {
    return guarded([&] {
        if (!s->s->turbulence()) throw std::invalid_argument("laminar PISO has no turbulence fields");
        s->s->turbulence()->setFields(k, omega);
    });
}
int bsim_piso_get_turbulence(bsim_piso* s, double* k, double* omega, double* mut) {
    return guarded([&] {
        if (!s->s->turbulence()) throw std::invalid_argument("laminar PISO has no turbulence fields");
        s->s->turbulence()->getFields(k, omega, mut);
    });
}
int bsim_piso_get_fields(bsim_piso* s, double* U, double* p, double* phi) {
    return guarded([&] { s->s->getFields(U, p, phi); });
}
int bsim_piso_temporal_flux(bsim_piso* s, double* out) {
    return guarded([&] { s->s->temporalFlux(out); });
}
int bsim_piso_history(bsim_piso* s, int which, double* out, int cap, int* len) {
    return guarded([&] {
        if (which < 0 || which > 3) throw std::invalid_argument("history index is 0..3");
        const std::vector<double>& h = which < 3 ? s->s->momHist[which] : s->s->pHist;
        *len = static_cast<int>(h.size());
        if (out) std::memcpy(out, h.data(), sizeof(double) * std::min<std::size_t>(cap, h.size()));
    });
}
int bsim_piso_timing(bsim_piso* s, double* o) {
    return guarded([&] {
        const PisoTiming& t = s->s->timing;
        const double v[8] = {t.history, t.assembly, t.momentumSolve, t.corrector, t.pressureAssembly,
                             t.pressureSolve, t.diagnostics, static_cast<double>(t.steps)};
        std::memcpy(o, v, sizeof(v));
    });
}
int bsim_piso_performance(bsim_piso* s, bsim_performance* total, bsim_equation_counters* systems, int cap,
                          int* count) {
    return guarded([&] {
        if (!s) throw std::invalid_argument("null argument");
        PisoSolver& S = *s->s;
        std::vector<std::pair<const DeviceSystem*, const char*>> sys = {
            {&S.momentumSystem(), "vector:momentum:U"}, {&S.pressureSystem(), "scalar:pressure:p"}};
        std::uint64_t reads = S.run.reductions;
        if (const KOmegaDevice* t = S.turbulence()) {
            if (t->kSystem()) sys.push_back({t->kSystem(), "scalar:kTransport:k"});
            if (t->omegaSystem()) sys.push_back({t->omegaSystem(), "scalar:omegaTransport:omega"});
            reads += t->reads;
        }
        runPerformance(sys, S.run, reads, S.timing.assembly + S.timing.pressureAssembly, total, systems, cap, count);
    });
}
int bsim_piso_probe_start(bsim_piso* s, int capacity) {
    return guarded([&] { s->s->pressureSystem().probe.start(capacity); });
}
int bsim_piso_probe_stop(bsim_piso* s, int* launches, double* seconds, double* bytes_per_launch) {
    return guarded([&] {
        KernelProbe& P = s->s->pressureSystem().probe;
        *seconds = P.stop(*launches);
        *bytes_per_launch = P.bytesPerLaunch;
    });
}
void bsim_piso_destroy(bsim_piso* s) { delete s; }

}  // extern "C"
