// Device and pinned-host block cache of the library, and the lock that keeps
// stream capture apart from memory release.
//
// Device buffers (DBuf) and pinned host mirrors freed inside a StreamScope go
// to a per-device cache of blocks keyed by size (rounded up to 256 bytes),
// each with an event recorded on the scope's stream.  A later allocation of
// the same size takes a cached block when it runs on that same stream
// (stream order makes the reuse safe) or when the block's event has
// completed.  A request that follows another on the same mesh asks for the
// same sizes, so in steady serving the solvers, fields and work vectors of a
// case are rebuilt in memory that stays resident, and no hipMalloc or hipFree
// runs.  Outside a scope, memory goes straight back to the runtime.  The
// cache holds at most a quarter of the device's memory; beyond that, and when
// an allocation fails, blocks return to the runtime.
//
// hipFree and hipHostFree synchronise the device, which must not reach a
// stream another thread is capturing into a graph.  runtimeLock() is held
// around every call that creates or destroys a runtime object (memory,
// streams, events, graphs) and around each graph capture and instantiation,
// so those never overlap across threads.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

namespace bsim::pool {

std::recursive_mutex& runtimeLock();

// Frees on this thread while the scope lives are stream-ordered on `stream`.
class StreamScope {
public:
    explicit StreamScope(hipStream_t stream);
    ~StreamScope();
    StreamScope(const StreamScope&) = delete;
    StreamScope& operator=(const StreamScope&) = delete;

private:
    hipStream_t prev;
    bool prevScoped;
};

// Device memory on the current device; throws on failure.
void* deviceAlloc(std::size_t bytes);
void deviceFree(void* p) noexcept;
// Pinned host memory (hipHostMalloc); throws on failure.
void* hostAlloc(std::size_t bytes);
void hostFree(void* p) noexcept;

struct Stats {
    double cachedBytes = 0, liveBytes = 0, hits = 0, misses = 0, released = 0;
};
Stats deviceStats(int device);
// Returns every cached block of the device to the runtime.
void trim(int device);

}  // namespace bsim::pool
