// Serving runtime (see serve.hpp): admission by an HBM budget, continuous
// admission as memory and slots free up, one worker thread per slot driving
// its own context stream, resident per-mesh contexts with LRU eviction, and
// the bsim_server_* C ABI.
#include "serve.hpp"

#include "errors.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace bsim {

namespace {
constexpr long long kMiB = 1LL << 20;
}

// Context: face geometry (owner/neighbour/patch, Cf, Sf, n, k, skew, |Sf|, w,
// orth: ~156 B per face), sliced half-face maps (~16 B per interior face) and
// cell arrays.  Solvers: the SIMPLE/PISO fields, two double-buffered systems
// with their Krylov vectors and the AMG hierarchy.  The figures are upper
// bounds of what the measured allocations of cavity and channel meshes gave
// (at most 1.94 GB for a 1M-cell steady request, against 2.49 GB reserved);
// a case's context reservation is replaced by its measured size once built.
long long Server::contextEstimate(int nc, int nf) {
    return 200LL * nf + 160LL * nc + 2 * kMiB;
}
long long Server::solverEstimate(int kind, int nc, int nf) {
    const long long perCell = kind == kRequestTransient ? 2200 : 2000;
    // The fixed part: two systems' reduction partials (6 x 32768 doubles
    // each), Krylov scalars and the solver's own reduction buffers.
    return perCell * nc + 208LL * nf + 10 * kMiB;  // 48 B per face: the face-ordered assembly terms
}

Server::Server(int dev, int slots, long long budgetBytes) : device(dev), budget(budgetBytes) {
    int count = 0;
    BSIM_HIP(hipGetDeviceCount(&count));
    if (dev < 0 || dev >= count) throw std::invalid_argument("device index is out of range");
    if (slots < 1 || slots > 64) throw std::invalid_argument("slots must lie in [1, 64]");
    if (budgetBytes <= 0) throw std::invalid_argument("memory budget must be positive");
    st.budget = budget;
    t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < slots; ++k) threads.emplace_back([this, k] { worker(k); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu);
        stopping = true;
        for (ServeRequest* r : queue) {
            r->status = kServeCancelled;
            r->state = 2;
        }
        queue.clear();
    }
    cv.notify_all();
    doneCv.notify_all();
    for (std::thread& t : threads) t.join();
    (void)hipSetDevice(device);
    cases.clear();  // resident contexts
}

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

ServeCase& Server::caseOf(int id) {
    std::lock_guard<std::mutex> lk(mu);
    if (id < 0 || id >= static_cast<int>(cases.size())) throw std::invalid_argument("unknown case id");
    return *cases[id];
}

int Server::addCase(const MeshDesc& m) {
    auto c = std::make_unique<ServeCase>();
    c->mesh = m;
    c->ncells = m.ncells();
    c->nfaces = m.nfaces();
    if (c->ncells <= 0) throw std::invalid_argument("mesh has no cells");
    c->ctxBytes = contextEstimate(c->ncells, c->nfaces);
    std::lock_guard<std::mutex> lk(mu);
    cases.push_back(std::move(c));
    return static_cast<int>(cases.size()) - 1;
}

int Server::submit(std::unique_ptr<ServeRequest> r, double arrival) {
    std::lock_guard<std::mutex> lk(mu);
    if (stopping) throw std::logic_error("server is closing");
    if (r->caseId < 0 || r->caseId >= static_cast<int>(cases.size())) throw std::invalid_argument("unknown case id");
    const double t = now();
    r->release = std::max(arrival, t);
    r->id = nextId++;
    ServeRequest* raw = r.get();
    requests.emplace(raw->id, std::move(r));
    // Release order, submission order among equal release times.
    auto it = std::upper_bound(queue.begin(), queue.end(), raw->release,
                               [](double v, const ServeRequest* q) { return v < q->release; });
    queue.insert(it, raw);
    cv.notify_all();
    return raw->id;
}

bool Server::evictOne(int keepCase, std::vector<std::unique_ptr<Context>>& evicted) {
    ServeCase* best = nullptr;
    std::size_t bestIdx = 0;
    unsigned long long bestUse = ~0ULL;
    for (int ci = 0; ci < static_cast<int>(cases.size()); ++ci) {
        ServeCase& c = *cases[ci];
        for (std::size_t k = 0; k < c.idle.size(); ++k) {
            // The case about to run keeps one idle context to reuse.
            if (ci == keepCase && c.idle.size() == 1) continue;
            if (c.idle[k].lastUse < bestUse) { bestUse = c.idle[k].lastUse; best = &c; bestIdx = k; }
        }
    }
    if (!best) return false;
    evicted.push_back(std::move(best->idle[bestIdx].ctx));
    best->idle.erase(best->idle.begin() + static_cast<std::ptrdiff_t>(bestIdx));
    st.reserved -= best->ctxBytes;
    ++st.evictions;
    --st.residentIdle;
    return true;
}

void Server::worker(int slot) {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        if (stopping) return;
        if (queue.empty()) { cv.wait(lk); continue; }
        ServeRequest* r = queue.front();
        const double t = now();
        if (r->release > t) {
            cv.wait_for(lk, std::chrono::duration<double>(r->release - t));
            continue;
        }
        ServeCase& sc = *cases[r->caseId];
        const long long solver = solverEstimate(r->kind, sc.ncells, sc.nfaces);
        if (solver + sc.ctxBytes > budget) {  // can never fit: answer at once
            queue.pop_front();
            r->status = kServeRejected;
            r->error = "request needs more device memory than the whole budget";
            r->tAdmit = r->tFinish = t;
            r->state = 2;
            ++st.rejected;
            doneCv.notify_all();
            continue;
        }
        const bool reuse = !sc.idle.empty();
        const long long need = solver + (reuse ? 0 : sc.ctxBytes);
        std::vector<std::unique_ptr<Context>> evicted;
        while (st.reserved + need > budget && evictOne(r->caseId, evicted)) {}
        if (st.reserved + need > budget) {
            // The head waits for a running request to return its memory;
            // later arrivals do not overtake it.
            lk.unlock();
            evicted.clear();
            lk.lock();
            cv.wait(lk);
            continue;
        }
        queue.pop_front();
        std::unique_ptr<Context> ctx;
        long long ctxAccount;  // what the reservation holds for this request's context
        if (reuse) {
            ctx = std::move(sc.idle.back().ctx);
            sc.idle.pop_back();
            --st.residentIdle;
            ++st.contextReuses;
            r->reusedContext = 1;
            ctxAccount = sc.ctxBytes;
        } else {
            ctxAccount = sc.ctxBytes;
            st.reserved += ctxAccount;
        }
        st.reserved += solver;
        r->reserved = solver + ctxAccount;
        r->state = 1;
        r->slot = slot;
        r->tAdmit = t;
        ++st.running;
        st.peakRunning = std::max(st.peakRunning, st.running);
        st.peakReserved = std::max(st.peakReserved, st.reserved);
        lk.unlock();
        evicted.clear();  // frees outside the lock
        long long ctxMeasured = -1;
        try {
            if (!ctx) {
                const long long before = t_deviceBytes;
                ctx = std::make_unique<Context>(sc.mesh, true, device);
                ctxMeasured = t_deviceBytes - before;
            }
            run(*r, ctx);
        } catch (const std::exception& e) {
            r->status = kServeError;
            r->error = e.what();
            ctx.reset();  // a context that saw a failure is not kept
        } catch (...) {
            r->status = kServeError;
            r->error = "unknown error";
            ctx.reset();
        }
        const double tf = now();
        lk.lock();
        st.reserved -= solver;
        if (ctxMeasured >= 0) {
            ++st.contextsCreated;
            if (!sc.measured) {
                sc.measured = true;
                sc.ctxBytes = ctxMeasured;
            }
        }
        if (ctx) {
            // Resident contexts are accounted at the case's (measured) size.
            st.reserved += sc.ctxBytes - ctxAccount;
            sc.idle.push_back({std::move(ctx), ++useTick});
            ++st.residentIdle;
        } else {
            st.reserved -= ctxAccount;
        }
        r->tFinish = tf;
        r->state = 2;
        --st.running;
        ++st.completed;
        cv.notify_all();
        doneCv.notify_all();
    }
}

void Server::run(ServeRequest& r, std::unique_ptr<Context>& ctx) {
    BSIM_HIP(hipSetDevice(device));
    // The solver's buffers return to the block cache stream-ordered on the
    // context's stream: the next request on this mesh takes the same blocks
    // back, so its solver, fields and work vectors are built in resident
    // memory without runtime allocations.
    pool::StreamScope scope(ctx->stream);
    const long long base = t_deviceBytes;
    long long peak = 0;
    r.tStart = now();
    const double* U0 = r.U0.empty() ? nullptr : r.U0.data();
    const double* p0 = r.p0.empty() ? nullptr : r.p0.data();
    const int nc = ctx->mesh->nc, nf = ctx->mesh->nf;
    r.U.assign(3 * static_cast<std::size_t>(nc), 0.0);
    r.p.assign(nc, 0.0);
    r.phi.assign(nf, 0.0);
    if (r.kind == kRequestSteady) {
        SimpleSolver s(*ctx, r.simple, r.ubct.data(), r.ubcv.data(), r.pbct.data(), r.pbcv.data());
        if (U0 || p0) s.setFields(U0, p0);
        s.start();
        r.status = kServeNotConverged;
        for (int it = 1; it <= r.maxIterations; ++it) {
            const IterRep rep = s.iterate();
            peak = std::max(peak, t_deviceBytes - base);
            r.iterations = it;
            r.rU = rep.rU; r.dU = rep.dU; r.dP = rep.dP; r.mass = rep.mass;
            if (rep.failed) { r.status = kServeNumericalFailure; break; }  // runSimple's numericalFailure
            if (rep.converged) { r.status = kServeConverged; break; }
        }
        s.getFields(r.U.data(), r.p.data(), r.phi.data());
    } else {
        PisoSolver s(*ctx, r.piso, r.ubct.data(), r.ubcv.data(), r.pbct.data(), r.pbcv.data());
        if (U0 || p0) s.setFields(U0, p0);
        s.start();
        r.status = kServeConverged;
        for (int k = 1; k <= r.maxIterations; ++k) {
            const StepRep rep = s.step(r.dt);
            peak = std::max(peak, t_deviceBytes - base);
            r.iterations = k;
            r.rU = rep.rU; r.dU = rep.dU; r.dP = rep.dP; r.mass = rep.mass;
            if (rep.status == kStepNotConverged) { r.status = kServeNotConverged; break; }
            if (rep.status != kStepCompleted) { r.status = kServeNumericalFailure; break; }
        }
        s.getFields(r.U.data(), r.p.data(), r.phi.data());
    }
    r.measured = peak;
}

const ServeRequest& Server::wait(int id) {
    std::unique_lock<std::mutex> lk(mu);
    auto it = requests.find(id);
    if (it == requests.end()) throw std::invalid_argument("unknown request id");
    ServeRequest* r = it->second.get();
    doneCv.wait(lk, [&] { return r->state == 2; });
    return *r;
}

bool Server::done(int id) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = requests.find(id);
    if (it == requests.end()) throw std::invalid_argument("unknown request id");
    return it->second->state == 2;
}

void Server::release(int id) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = requests.find(id);
    if (it == requests.end()) throw std::invalid_argument("unknown request id");
    if (it->second->state != 2) throw std::logic_error("request is still queued or running");
    requests.erase(it);
}

ServeStats Server::stats() {
    std::lock_guard<std::mutex> lk(mu);
    return st;
}

}  // namespace bsim

struct bsim_server {
    std::unique_ptr<bsim::Server> s;
};

using namespace bsim;

namespace {

void copyBcs(int np, const int* ubct, const double* ubcv, const int* pbct, const double* pbcv, ServeRequest& r) {
    if (!ubct || !ubcv || !pbct || !pbcv) throw std::invalid_argument("null boundary conditions");
    r.ubct.assign(ubct, ubct + np);
    r.ubcv.assign(ubcv, ubcv + 3 * np);
    r.pbct.assign(pbct, pbct + np);
    r.pbcv.assign(pbcv, pbcv + np);
}

}  // namespace

extern "C" {

int bsim_server_create(int device, int slots, double memory_budget, bsim_server** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null argument");
        if (!(memory_budget > 0) || !std::isfinite(memory_budget)) throw std::invalid_argument("memory budget must be positive");
        auto* s = new bsim_server;
        try { s->s = std::make_unique<Server>(device, slots, static_cast<long long>(memory_budget)); }
        catch (...) { delete s; throw; }
        *out = s;
    });
}

int bsim_server_add_case(bsim_server* s, const bsim_mesh* mesh, int* case_id) {
    return guarded([&] {
        if (!s || !mesh || !case_id) throw std::invalid_argument("null argument");
        *case_id = s->s->addCase(mesh->d);
    });
}

int bsim_server_submit(bsim_server* s, const bsim_request* q, int* id) {
    return guarded([&] {
        if (!s || !q || !id) throw std::invalid_argument("null argument");
        ServeCase& sc = s->s->caseOf(q->case_id);
        auto r = std::make_unique<ServeRequest>();
        r->kind = q->kind;
        r->caseId = q->case_id;
        r->maxIterations = q->max_iterations;
        r->dt = q->dt;
        if (r->maxIterations < 1) throw std::invalid_argument("max_iterations must be positive");
        copyBcs(sc.mesh.npatch, q->u_bc_type, q->u_bc_value, q->p_bc_type, q->p_bc_value, *r);
        for (int k = 0; k < sc.mesh.npatch; ++k)
            if (r->ubct[k] < 0 || r->ubct[k] > 4 || r->pbct[k] < 0 || r->pbct[k] > 4)
                throw std::invalid_argument("invalid boundary type");
        if (q->kind == BSIM_REQ_STEADY) {
            if (!q->simple) throw std::invalid_argument("steady request needs a SIMPLE configuration");
            r->simple = toSimpleCfg(*q->simple);
        } else if (q->kind == BSIM_REQ_TRANSIENT) {
            if (!q->piso) throw std::invalid_argument("transient request needs a PISO configuration");
            if (!(q->dt > 0) || !std::isfinite(q->dt)) throw std::invalid_argument("time step must be positive");
            r->piso = toPisoCfg(*q->piso);
        } else {
            throw std::invalid_argument("unknown request kind");
        }
        if (q->U0) r->U0.assign(q->U0, q->U0 + 3 * static_cast<std::size_t>(sc.ncells));
        if (q->p0) r->p0.assign(q->p0, q->p0 + sc.ncells);
        *id = s->s->submit(std::move(r), q->arrival);
    });
}

int bsim_server_wait(bsim_server* s, int id, bsim_request_result* o) {
    return guarded([&] {
        if (!s || !o) throw std::invalid_argument("null argument");
        const ServeRequest& r = s->s->wait(id);
        o->status = r.status; o->iterations = r.iterations; o->slot = r.slot; o->reused_context = r.reusedContext;
        o->released = r.release; o->admitted = r.tAdmit; o->started = r.tStart; o->finished = r.tFinish;
        o->rU = r.rU; o->dU = r.dU; o->dP = r.dP; o->mass = r.mass;
        o->reserved_bytes = static_cast<double>(r.reserved);
        o->measured_bytes = static_cast<double>(r.measured);
        std::strncpy(o->message, r.error.c_str(), sizeof(o->message) - 1);
        o->message[sizeof(o->message) - 1] = '\0';
    });
}

int bsim_server_done(bsim_server* s, int id, int* done) {
    return guarded([&] {
        if (!s || !done) throw std::invalid_argument("null argument");
        *done = s->s->done(id) ? 1 : 0;
    });
}

int bsim_server_fields(bsim_server* s, int id, double* U, double* p, double* phi) {
    return guarded([&] {
        if (!s) throw std::invalid_argument("null argument");
        const ServeRequest& r = s->s->wait(id);
        if (r.U.empty()) throw std::invalid_argument("request produced no fields");
        if (U) std::memcpy(U, r.U.data(), sizeof(double) * r.U.size());
        if (p) std::memcpy(p, r.p.data(), sizeof(double) * r.p.size());
        if (phi) std::memcpy(phi, r.phi.data(), sizeof(double) * r.phi.size());
    });
}

int bsim_server_release(bsim_server* s, int id) {
    return guarded([&] {
        if (!s) throw std::invalid_argument("null argument");
        s->s->release(id);
    });
}

int bsim_server_stats(bsim_server* s, double* o) {
    return guarded([&] {
        if (!s || !o) throw std::invalid_argument("null argument");
        const ServeStats t = s->s->stats();
        const double v[11] = {static_cast<double>(t.budget), static_cast<double>(t.peakReserved),
                              static_cast<double>(t.reserved), static_cast<double>(t.peakRunning),
                              static_cast<double>(t.running), static_cast<double>(t.completed),
                              static_cast<double>(t.rejected), static_cast<double>(t.contextsCreated),
                              static_cast<double>(t.contextReuses), static_cast<double>(t.evictions),
                              static_cast<double>(t.residentIdle)};
        std::memcpy(o, v, sizeof(v));
    });
}

double bsim_server_clock(bsim_server* s) { return s ? s->s->now() : 0.0; }

int bsim_server_estimate(int kind, int ncells, int nfaces, double* context_bytes, double* solver_bytes) {
    return guarded([&] {
        if (!context_bytes || !solver_bytes) throw std::invalid_argument("null argument");
        if (kind != BSIM_REQ_STEADY && kind != BSIM_REQ_TRANSIENT) throw std::invalid_argument("unknown request kind");
        if (ncells <= 0 || nfaces <= 0) throw std::invalid_argument("mesh sizes must be positive");
        *context_bytes = static_cast<double>(Server::contextEstimate(ncells, nfaces));
        *solver_bytes = static_cast<double>(Server::solverEstimate(kind, ncells, nfaces));
    });
}

void bsim_server_destroy(bsim_server* s) { delete s; }

int bsim_device_pool_stats(int device, double* o) {
    return guarded([&] {
        if (!o) throw std::invalid_argument("null argument");
        const pool::Stats t = pool::deviceStats(device);
        const double v[5] = {t.cachedBytes, t.liveBytes, t.hits, t.misses, t.released};
        std::memcpy(o, v, sizeof(v));
    });
}

int bsim_device_pool_trim(int device) {
    return guarded([&] { pool::trim(device); });
}

}  // extern "C"
