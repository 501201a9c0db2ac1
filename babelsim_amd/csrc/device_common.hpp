// Shared device helpers: HIP error checks, RAII device buffers, small vector
// algebra and deterministic two-stage reductions for 64-wide wavefronts.
#pragma once

#include "pool.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace bsim {

#define BSIM_HIP(expr)                                                                              \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " +  \
                                     __FILE__ + ":" + std::to_string(__LINE__) + ": " #expr);       \
    } while (0)

#define BSIM_LAUNCH_CHECK() BSIM_HIP(hipGetLastError())

constexpr int kBlock = 256;          // 4 wavefronts of 64
constexpr int kMaxReduceBlocks = 32768;  // one row per thread up to 8.4M rows
constexpr int kTail = 1024;              // threads of a reduction tail

inline int blocksFor(long long n, int block = kBlock) {
    long long b = (n + block - 1) / block;
    return static_cast<int>(b < 1 ? 1 : b);
}

// Device bytes held by DBufs that the calling thread allocated minus those it
// released; the serving runtime reads its change across one request.
inline thread_local long long t_deviceBytes = 0;

// Device allocation of T[n]: owning (alloc, from the block cache of pool.hpp)
// or a view of memory that someone else owns (borrow).  A borrowed buffer
// never allocates or frees: growing it past its size throws instead.
template <class T> struct DBuf {
    T* p = nullptr;
    std::size_t n = 0;
    bool owned = true;
    DBuf() = default;
    explicit DBuf(std::size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n), owned(o.owned) { o.p = nullptr; o.n = 0; o.owned = true; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p; n = o.n; owned = o.owned;
            o.p = nullptr; o.n = 0; o.owned = true;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(std::size_t count) {
        if (!owned) throw std::logic_error("a borrowed device buffer cannot be reallocated");
        release();
        if (count) {
            p = static_cast<T*>(pool::deviceAlloc(sizeof(T) * count));
            t_deviceBytes += static_cast<long long>(sizeof(T) * count);
        }
        n = count;
    }
    void borrow(T* ptr, std::size_t count) {
        release();
        p = ptr;
        n = count;
        owned = false;
    }
    void ensure(std::size_t count) { if (count > n) alloc(count); }
    void release() {
        if (p && owned) {
            pool::deviceFree(p);
            t_deviceBytes -= static_cast<long long>(sizeof(T) * n);
        }
        p = nullptr;
        n = 0;
        owned = true;
    }
    void upload(const T* src, std::size_t count, hipStream_t s) {
        ensure(count);
        if (count) BSIM_HIP(hipMemcpyAsync(p, src, sizeof(T) * count, hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
    void download(T* dst, std::size_t count, hipStream_t s) const {
        if (count) BSIM_HIP(hipMemcpyAsync(dst, p, sizeof(T) * count, hipMemcpyDeviceToHost, s));
    }
    std::vector<T> toHost(hipStream_t s, std::size_t count = SIZE_MAX) const {
        std::vector<T> v(count == SIZE_MAX ? n : count);
        download(v.data(), v.size(), s);
        BSIM_HIP(hipStreamSynchronize(s));
        return v;
    }
    void zero(hipStream_t s) {
        if (n) BSIM_HIP(hipMemsetAsync(p, 0, sizeof(T) * n, s));
    }
};

// Logical workgroup index.  The hardware deals workgroups round-robin over
// the eight XCDs, each with its own L2; renumbering them so that XCD x runs
// one contiguous range of blocks keeps neighbouring rows, which gather the
// same cache lines, in one L2.  A bijection on [0, gridDim.x).
constexpr int kXcds = 8;
__device__ inline int xcdBlock() {
    const int b = blockIdx.x, q = gridDim.x / kXcds, r = gridDim.x % kXcds;
    const int x = b % kXcds;
    return x * q + (x < r ? x : r) + b / kXcds;
}
__device__ inline int globalThread() { return xcdBlock() * blockDim.x + threadIdx.x; }

// Row layout shared by the cell-gather kernels and the fine-level matrices:
// rows are grouped in slices of 64 (one wavefront); the k-th off-diagonal
// entry of the rows of slice s is stored at sliceOff[s*(K+1)+k] + lane, so a
// wavefront reading level k touches one contiguous run.  Rows in a slice are
// ordered by decreasing length when the mesh is renumbered; otherwise the gaps
// are padding entries with face and column -1.
// Every kernel that walks rows runs blocks of whole wavefronts over
// consecutive rows (or strides by a multiple of 64), so all lanes of a
// wavefront share one slice and its offsets are read with scalar loads.
static_assert(kBlock % 64 == 0, "row kernels need blocks of whole wavefronts");
struct RowRange {
    const int* so;
    int lane, K;
    __device__ RowRange(const int* sliceOff, int K_, int c)
        : so(sliceOff + static_cast<long long>(__builtin_amdgcn_readfirstlane(c >> 6)) * (K_ + 1)), lane(c & 63), K(K_) {}
    __device__ bool valid(int k) const { return k < K && lane < so[k + 1] - so[k]; }
    __device__ int at(int k) const { return so[k] + lane; }
};

// ---- device vector helpers (component-wise, AoS double[3]) ----
struct D3 {
    double x, y, z;
};
__host__ __device__ inline D3 d3(double a, double b, double c) { return {a, b, c}; }
__host__ __device__ inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline D3 operator/(D3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline D3 cross(D3 a, D3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline double comp(D3 a, int c) { return c == 0 ? a.x : (c == 1 ? a.y : a.z); }
__device__ inline D3 ld3(const double* p, long long i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ inline void st3(double* p, long long i, D3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }

// ---- deterministic reductions ----
// Stage 1 (the producing kernel): each block reduces its values with a fixed
// in-block tree and stores the K sums at partials[k * kMaxReduceBlocks + block]
// (logical block index, see xcdBlock).
// Stage 2 (kReduceTail, one block launched after it on the same stream): sums
// the partials in block order with a fixed tree and hands them to a functor.
// The kernel boundary orders the two stages, so no device-scope fence or
// completion counter is needed and the result does not depend on scheduling.
// Wavefront stage: an xor butterfly over the 64 lanes.  Both lanes of every
// exchange add the same two operands, so all lanes end with one bitwise-equal
// total and the order of additions is fixed.
template <class Op> __device__ inline double waveReduce(double v, Op op) {
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off, 64));
    return v;
}
struct SumOp {
    __device__ double operator()(double a, double b) const { return a + b; }
};
struct MaxOp {
    __device__ double operator()(double a, double b) const { return fmax(a, b); }
};
// Block stage: the per-wavefront results go through LDS (one slot per
// wavefront) and every thread combines them in wavefront order.
template <int K, int NT, class Op> __device__ inline void blockReduce(double (&v)[K], double* smem, Op op) {
    static_assert(NT % 64 == 0, "blocks of whole wavefronts");
    constexpr int NW = NT / 64;
    for (int k = 0; k < K; ++k) v[k] = waveReduce(v[k], op);
    if (NW == 1) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < K; ++k) smem[k * NW + w] = v[k];
    __syncthreads();
#pragma unroll 1
    for (int k = 0; k < K; ++k) {
        double a = smem[k * NW];
        for (int j = 1; j < NW; ++j) a = op(a, smem[k * NW + j]);
        v[k] = a;
    }
    __syncthreads();
}
// The K sums of the block, in every thread (smem: at least K * NT / 64 doubles).
template <int K, int NT = kBlock> __device__ inline void blockSum(double (&v)[K], double* smem) {
    blockReduce<K, NT>(v, smem, SumOp{});
}
// The maximum over the block, in every thread.
template <int NT = kBlock> __device__ inline double blockMaxValue(double v, double* smem) {
    double a[1] = {v};
    blockReduce<1, NT>(a, smem, MaxOp{});
    return a[0];
}

// Partials are stored under the logical block index, so the final sum runs
// over the rows in the same order whatever the hardware placement.
template <int K> __device__ inline void blockPartial(double (&v)[K], double* partials, double* smem) {
    blockSum<K>(v, smem);
    if (threadIdx.x == 0) {
        const int blk = xcdBlock();
        for (int k = 0; k < K; ++k) partials[k * kMaxReduceBlocks + blk] = v[k];
    }
}

// Launched as one block of kTail threads: thread t sums partials t, t + kTail,
// ... in order, then a fixed tree.  F: device functor called by thread 0
// with the K totals.
template <int K, class F> __global__ void __launch_bounds__(kTail) kReduceTail(int nb, const double* partials, F f) {
    __shared__ double smem[K * kTail / 64];
    double acc[K];
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    // One sum at a time with a short unroll keeps 1024 threads within 128
    // VGPRs each for every K.
    for (int k = 0; k < K; ++k) {
        const double* p = partials + static_cast<long long>(k) * kMaxReduceBlocks;
        double a = 0.0;
#pragma unroll 4
        for (int b = threadIdx.x; b < nb; b += kTail) a += p[b];
        acc[k] = a;
    }
    blockSum<K, kTail>(acc, smem);
    if (threadIdx.x == 0) f(acc);
}

// Stores the K totals to out[0..K-1].
template <int K> struct StoreSums {
    double* out;
    __device__ void operator()(const double (&v)[K]) const {
        for (int k = 0; k < K; ++k) out[k] = v[k];
    }
};

inline int reduceBlocks(long long n) {
    const int b = blocksFor(n);
    return b < kMaxReduceBlocks ? b : kMaxReduceBlocks;
}

}  // namespace bsim
