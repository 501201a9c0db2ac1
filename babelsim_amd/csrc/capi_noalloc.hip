// bsim_csr_solve_noalloc: the stand-alone CSR solve of bsim_csr_solve on
// caller-owned device arrays, for callers that cannot allocate device memory.
//
// The Krylov vectors, reduction partials and device scalars live in one
// file-scope device array, `scratch`, that every call on a device shares.
// Host code never passes the array's name to a kernel: its device address is
// looked up once per device (hipGetSymbolAddress, with that device current)
// and only that address is used.  Calls on one device are serialised by a
// per-device lock and return after their stream has drained, so the next
// call finds the array idle; the array keeps the previous call's contents,
// and everything a solve reads is written by it first (vectors, partials) or
// cleared at its start (scalars, flags, the argument-check flag).
#include "../../include/bsim.h"
#include "context.hpp"
#include "errors.hpp"
#include "linsolve.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace bsim {
namespace {

constexpr int kNoAllocMaxRows = BSIM_NOALLOC_MAX_ROWS;
constexpr std::size_t kCheckDoubles = 32;  // the argument-check flag, 256-byte region
// DeviceSystem::workspaceDoubles(kNoAllocMaxRows): two 256-byte scalar
// regions, 6 x 32768 reduction partials and 12 vectors of kNoAllocMaxRows.
constexpr std::size_t kWorkspaceDoubles =
    2 * 32 + 6 * static_cast<std::size_t>(kMaxReduceBlocks) + 12 * static_cast<std::size_t>(kNoAllocMaxRows);
constexpr std::size_t kScratchDoubles = kCheckDoubles + kWorkspaceDoubles;  // 6,488,160 doubles = 49.5 MiB

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_deviceLock[kMaxDevices];
double* g_scratchAddress[kMaxDevices] = {};  // per device; written under that device's lock

// Device address of `scratch` on the current device (the caller holds the
// device's lock).
double* scratchAddress(int device) {
    if (!g_scratchAddress[device]) {
        void* ptr = nullptr;
        BSIM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw std::runtime_error("device address of the solve scratch array is null");
        g_scratchAddress[device] = static_cast<double*>(ptr);
    }
    return g_scratchAddress[device];
}

// Row pointers non-decreasing within [0, nnz] and every column in [0, n):
// what bsim_csr_solve checks on the host before uploading.  Only entries
// inside a row range that passed its own check are read.
__global__ void __launch_bounds__(kBlock) kCheckCsr(int n, int nnz, const int* ptr, const int* col, int* bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = ptr[i], b = ptr[i + 1];
    if (a < 0 || b < a || b > nnz) {
        atomicOr(bad, 1);
        return;
    }
    for (int k = a; k < b; ++k)
        if (col[k] < 0 || col[k] >= n) {
            atomicOr(bad, 2);
            return;
        }
}

}  // namespace
}  // namespace bsim

using namespace bsim;

extern "C" int bsim_csr_solve_noalloc(int device, void* stream, int n, const int* ptr, const int* col,
                                      const double* diag, const double* val, const double* b, double* x,
                                      const bsim_solver_config* cfg, bsim_solve_result* res, double* hist, int cap,
                                      int* len) {
    return guarded([&] {
        if (!ptr || !diag || !b || !x || !cfg || !res) throw std::invalid_argument("null argument");
        int count = 0;
        BSIM_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count || device >= kMaxDevices)
            throw std::invalid_argument("device index is out of range");
        if (n <= 0) throw std::invalid_argument("matrix is empty");
        if (n > kNoAllocMaxRows)
            throw std::invalid_argument("matrix has " + std::to_string(n) + " rows; the allocation-free solve holds at most " +
                                        std::to_string(kNoAllocMaxRows));
        const SolverCfg c = toSolverCfg(cfg);
        if (c.pc >= 2 || c.precision != 0)
            throw std::invalid_argument("the allocation-free solve runs FP64 CG or BiCGSTAB with no preconditioner "
                                        "or Jacobi only");
        if (DeviceSystem::workspaceDoubles(n) > kWorkspaceDoubles)
            throw std::logic_error("solve workspace layout exceeds the scratch array");

        std::lock_guard<std::mutex> lock(g_deviceLock[device]);
        BSIM_HIP(hipSetDevice(device));
        hipStream_t st = static_cast<hipStream_t>(stream);
        double* base = scratchAddress(device);

        int ends[2] = {-1, -1};
        BSIM_HIP(hipMemcpyAsync(&ends[0], ptr, sizeof(int), hipMemcpyDeviceToHost, st));
        BSIM_HIP(hipMemcpyAsync(&ends[1], ptr + n, sizeof(int), hipMemcpyDeviceToHost, st));
        BSIM_HIP(hipStreamSynchronize(st));
        if (ends[0] != 0 || ends[1] < 0) throw std::invalid_argument("row pointers must start at 0 and end at nnz >= 0");
        const int nnz = ends[1];
        if (nnz > 0 && (!col || !val)) throw std::invalid_argument("null argument");
        int* bad = reinterpret_cast<int*>(base);
        BSIM_HIP(hipMemsetAsync(bad, 0, sizeof(int), st));
        kCheckCsr<<<blocksFor(n), kBlock, 0, st>>>(n, nnz, ptr, col, bad);
        BSIM_LAUNCH_CHECK();
        int hbad = 0;
        BSIM_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
        BSIM_HIP(hipStreamSynchronize(st));
        if (hbad & 1) throw std::invalid_argument("row pointers must be non-decreasing within [0, nnz]");
        if (hbad & 2) throw std::invalid_argument("column index out of range");

        Pattern P;
        P.n = n;
        P.nnzStore = nnz;
        P.K = 0;
        P.ptr = ptr;
        P.col = col;
        P.sliceOff = nullptr;
        SolveRes r;
        std::size_t hlen = 0;
        try {
            DeviceSystem sys(st, std::move(P), SystemWorkspace{diag, val, base + kCheckDoubles, kWorkspaceDoubles});
            sys.update();
            sys.prepare(c);
            r = sys.solve(b, x, 0);
            // The next call may come on another stream: the array must be idle.
            BSIM_HIP(hipStreamSynchronize(st));
            hlen = sys.hist.size();
            if (hist && cap > 0) std::memcpy(hist, sys.hist.data(), sizeof(double) * std::min<std::size_t>(cap, hlen));
        } catch (...) {
            (void)hipStreamSynchronize(st);  // nothing may still use the array when the lock is released
            throw;
        }
        *res = bsim_solve_result{r.status, r.iterations, r.initial, r.final_, r.relative};
        if (len) *len = static_cast<int>(hlen);
    });
}
