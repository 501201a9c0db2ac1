// Serving runtime: many steady (SIMPLE) and transient (PISO) solves on one
// device, admitted by an HBM budget as they arrive and run concurrently, each
// on its own context and stream.  There is no counterpart in the reference,
// where one RunTime is active per thread (src/runtime/runtime.cpp:15, 48-51)
// and a Case owns one mesh (src/io/case.cpp:280-298); each request here runs
// exactly the SimpleSolver / PisoSolver path of a solo run, so its result does
// not depend on what else is on the device.
#pragma once

#include "capi_types.hpp"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace bsim {

enum : int { kRequestSteady = 0, kRequestTransient = 1 };
enum : int {
    kServeConverged = 0,     // steady: converged; transient: every step completed
    kServeNotConverged = 1,  // steady: iteration cap reached; transient: a step not converged
    kServeNumericalFailure = 2,
    kServeRejected = 3,      // larger than the whole memory budget
    kServeError = 4,         // an exception (message kept with the request)
    kServeCancelled = 5,     // still queued when the server closed
};

// A registered mesh: requests name it, and its device contexts (mesh and
// geometry in HBM) stay resident between requests until memory is needed.
struct ServeCase {
    MeshDesc mesh;
    int ncells = 0, nfaces = 0;
    long long ctxBytes = 0;      // reserved per context: estimate, then the measured size
    bool measured = false;
    struct Idle {
        std::unique_ptr<Context> ctx;
        unsigned long long lastUse;
    };
    std::vector<Idle> idle;
};

struct ServeRequest {
    int id = 0, kind = kRequestSteady, caseId = 0;
    SimpleCfg simple;
    PisoCfg piso;
    std::vector<int> ubct, pbct;
    std::vector<double> ubcv, pbcv, U0, p0;
    int maxIterations = 1;  // steady: outer-iteration cap; transient: time steps
    double dt = 0.0;
    double release = 0.0;   // server-clock seconds at which it becomes eligible
    // Outcome.
    int state = 0;  // 0 queued, 1 running, 2 done
    int status = kServeConverged, iterations = 0, slot = -1, reusedContext = 0;
    double tAdmit = 0.0, tStart = 0.0, tFinish = 0.0;
    double rU = 0.0, dU = 0.0, dP = 0.0, mass = 0.0;
    long long reserved = 0, measured = 0;
    std::vector<double> U, p, phi;
    std::string error;
};

struct ServeStats {
    long long budget = 0, peakReserved = 0, reserved = 0;
    int peakRunning = 0, running = 0, completed = 0, rejected = 0, contextsCreated = 0, contextReuses = 0,
        evictions = 0, residentIdle = 0;
};

class Server {
public:
    Server(int device, int slots, long long budgetBytes);
    ~Server();
    int addCase(const MeshDesc& mesh);
    int submit(std::unique_ptr<ServeRequest> r, double arrival);
    // Blocks until the request is done; the record stays until release().
    const ServeRequest& wait(int id);
    bool done(int id);
    void release(int id);
    ServeStats stats();
    double now() const;
    ServeCase& caseOf(int id);

    // Admission estimates (bytes of device memory per context and per solver).
    static long long contextEstimate(int ncells, int nfaces);
    static long long solverEstimate(int kind, int ncells, int nfaces);

private:
    void worker(int slot);
    void run(ServeRequest& r, std::unique_ptr<Context>& ctx);
    bool evictOne(int keepCase, std::vector<std::unique_ptr<Context>>& evicted);

    int device;
    long long budget;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv, doneCv;
    bool stopping = false;
    std::vector<std::thread> threads;
    std::vector<std::unique_ptr<ServeCase>> cases;
    std::map<int, std::unique_ptr<ServeRequest>> requests;
    std::deque<ServeRequest*> queue;  // by release time, then submission order
    unsigned long long useTick = 0;
    int nextId = 1;
    ServeStats st;
};

}  // namespace bsim
