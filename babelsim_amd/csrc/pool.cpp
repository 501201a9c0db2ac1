// Block cache of device and pinned host memory (see pool.hpp).
#include "pool.hpp"

#include "device_common.hpp"

#include <algorithm>
#include <map>
#include <unordered_map>
#include <vector>

namespace bsim::pool {

namespace {

constexpr std::size_t kGrain = 256;
constexpr int kMaxDevices = 64;
constexpr int kHost = kMaxDevices;  // index of the pinned-host cache
constexpr std::size_t kHostCap = std::size_t(256) << 20;

std::size_t roundUp(std::size_t b) { return (std::max<std::size_t>(b, 1) + kGrain - 1) / kGrain * kGrain; }

struct Block {
    void* p;
    hipStream_t stream;
    hipEvent_t done;  // recorded on `stream` when the block was freed
    int eventDevice;  // the device `done` belongs to
};
struct Cache {
    std::map<std::size_t, std::vector<Block>> free;
    std::size_t cached = 0, live = 0, cap = 0;
    double hits = 0, misses = 0, released = 0;
};
struct Live {
    int cache;  // device index, or kHost
    std::size_t bytes;
};

std::mutex g_mu;  // the caches, the live map and the spare events
Cache g_cache[kMaxDevices + 1];
std::unordered_map<void*, Live> g_live;
std::vector<hipEvent_t> g_events[kMaxDevices];  // spare events per device

thread_local hipStream_t t_stream = nullptr;
thread_local bool t_scoped = false;

int currentDevice() {
    int d = 0;
    BSIM_HIP(hipGetDevice(&d));
    if (d < 0 || d >= kMaxDevices) throw std::runtime_error("device index outside the block cache");
    return d;
}

// A runtime call whose failure the cache tolerates (it then takes another
// path) leaves no error behind for the caller's next launch check: HIP keeps
// the last failure per thread until hipGetLastError reads it.
hipError_t tolerated(hipError_t e) {
    if (e != hipSuccess) (void)hipGetLastError();
    return e;
}

// Event queries, records and waits, like allocation and release, run under
// runtimeLock(): the runtime refuses them while another thread captures a
// stream into a graph.
std::size_t deviceCap(int d) {
    Cache& c = g_cache[d];
    if (c.cap == 0) {
        std::size_t freeB = 0, total = 0;
        std::lock_guard<std::recursive_mutex> rl(runtimeLock());
        c.cap = tolerated(hipMemGetInfo(&freeB, &total)) == hipSuccess ? total / 4 : std::size_t(8) << 30;
    }
    return c.cap;
}

// A cached block of `bytes` that the calling thread may use now: a device
// block freed on its own stream, or any block whose free has completed (host
// code writes pinned memory directly, outside stream order, so pinned blocks
// always wait for their event).  Called under g_mu.
bool takeCached(int idx, std::size_t bytes, void*& out) {
    Cache& c = g_cache[idx];
    auto it = c.free.find(bytes);
    if (it == c.free.end()) return false;
    std::vector<Block>& v = it->second;
    for (std::size_t k = v.size(); k-- > 0;) {
        const bool ordered = idx != kHost && t_scoped && v[k].stream == t_stream;
        if (!ordered) {
            std::lock_guard<std::recursive_mutex> rl(runtimeLock());
            if (tolerated(hipEventQuery(v[k].done)) != hipSuccess) continue;
        }
        out = v[k].p;
        g_events[v[k].eventDevice].push_back(v[k].done);
        v.erase(v.begin() + static_cast<std::ptrdiff_t>(k));
        c.cached -= bytes;
        c.live += bytes;
        c.hits += 1;
        g_live[out] = {idx, bytes};
        return true;
    }
    return false;
}

void returnToRuntime(int idx, void* p) {
    std::lock_guard<std::recursive_mutex> rl(runtimeLock());
    if (idx == kHost) (void)tolerated(hipHostFree(p));
    else (void)tolerated(hipFree(p));
}

void trimCache(int idx) {
    std::vector<Block> blocks;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        Cache& c = g_cache[idx];
        for (auto& kv : c.free)
            for (Block& b : kv.second) blocks.push_back(b);
        c.released += static_cast<double>(blocks.size());
        c.free.clear();
        c.cached = 0;
    }
    for (Block& b : blocks) {
        {
            std::lock_guard<std::recursive_mutex> rl(runtimeLock());
            (void)tolerated(hipEventSynchronize(b.done));
            (void)tolerated(hipEventDestroy(b.done));
        }
        returnToRuntime(idx, b.p);
    }
}

// Caches p (idx, bytes) when a scope names its stream and the cap allows;
// otherwise returns it to the runtime.
void release(void* p) noexcept {
    if (!p) return;
    int idx = 0;
    std::size_t bytes = 0;
    bool cache = false;
    try {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_live.find(p);
        if (it == g_live.end()) return;  // not ours
        idx = it->second.cache;
        bytes = it->second.bytes;
        g_live.erase(it);
        Cache& c = g_cache[idx];
        c.live -= bytes;
        const std::size_t cap = idx == kHost ? kHostCap : deviceCap(idx);
        if (t_scoped && c.cached + bytes <= cap) {
            const int ed = currentDevice();
            hipEvent_t ev = nullptr;
            if (!g_events[ed].empty()) {
                ev = g_events[ed].back();
                g_events[ed].pop_back();
            } else {
                std::lock_guard<std::recursive_mutex> rl(runtimeLock());
                if (tolerated(hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) ev = nullptr;
            }
            bool recorded = false;
            if (ev) {
                std::lock_guard<std::recursive_mutex> rl(runtimeLock());
                recorded = tolerated(hipEventRecord(ev, t_stream)) == hipSuccess;
            }
            if (recorded) {
                c.free[bytes].push_back({p, t_stream, ev, ed});
                c.cached += bytes;
                cache = true;
            } else if (ev) {
                g_events[ed].push_back(ev);
            }
        }
        if (!cache) c.released += 1;
    } catch (...) {
        cache = false;
    }
    if (!cache) returnToRuntime(idx, p);
}

}  // namespace

std::recursive_mutex& runtimeLock() {
    static std::recursive_mutex m;
    return m;
}

StreamScope::StreamScope(hipStream_t stream) : prev(t_stream), prevScoped(t_scoped) {
    t_stream = stream;
    t_scoped = true;
}
StreamScope::~StreamScope() {
    t_stream = prev;
    t_scoped = prevScoped;
}

void* deviceAlloc(std::size_t bytes) {
    const std::size_t b = roundUp(bytes);
    const int d = currentDevice();
    {
        std::lock_guard<std::mutex> lk(g_mu);
        void* p = nullptr;
        if (takeCached(d, b, p)) return p;
        g_cache[d].misses += 1;
    }
    void* p = nullptr;
    hipError_t e;
    {
        std::lock_guard<std::recursive_mutex> rl(runtimeLock());
        e = hipMalloc(&p, b);
    }
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        trimCache(d);
        std::lock_guard<std::recursive_mutex> rl(runtimeLock());
        e = hipMalloc(&p, b);
    }
    BSIM_HIP(e);
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[d].live += b;
    g_live[p] = {d, b};
    return p;
}

void deviceFree(void* p) noexcept { release(p); }

void* hostAlloc(std::size_t bytes) {
    const std::size_t b = roundUp(bytes);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        void* p = nullptr;
        if (takeCached(kHost, b, p)) return p;
        g_cache[kHost].misses += 1;
    }
    void* p = nullptr;
    {
        std::lock_guard<std::recursive_mutex> rl(runtimeLock());
        BSIM_HIP(hipHostMalloc(&p, b, hipHostMallocDefault));
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[kHost].live += b;
    g_live[p] = {kHost, b};
    return p;
}

void hostFree(void* p) noexcept { release(p); }

Stats deviceStats(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device index is out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    const Cache& c = g_cache[device];
    return {static_cast<double>(c.cached), static_cast<double>(c.live), c.hits, c.misses, c.released};
}

void trim(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device index is out of range");
    trimCache(device);
}

}  // namespace bsim::pool
