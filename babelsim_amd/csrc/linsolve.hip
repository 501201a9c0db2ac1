// Krylov solvers, preconditioners and the solve contract on the GPU.
//
// Every reduction is deterministic (fixed grid, fixed trees), every scalar of
// the recurrences (alpha, beta, omega, ...) is computed on the device by the
// last block of the reduction that produces it, and the host reads one
// residual norm per iteration to run BabelSim's convergence test.
#include "linsolve.hpp"
#include "ilu.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <type_traits>

namespace bsim {

namespace {

enum Scalar : int { S_RHO = 0, S_RHOOLD, S_ALPHA, S_OMEGA, S_OMEGAOLD, S_BETA, S_BETAOLD, S_DPI, S_A, S_DP, S_D1, S_D2,
                    S_TR, S_TMP, S_BN, S_COUNT = 16 };
enum Flag : int { F_BETA_ZERO = 1, F_INDEF_PC = 2, F_INDEF_MAT = 4, F_D1_ZERO = 8, F_D2_ZERO = 16, F_NONFINITE = 32 };

using Clock = std::chrono::steady_clock;
double since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

__device__ inline double sgn(double v) { return (v > 0) - (v < 0); }

// Row i (held in storage slot q) of A x: the diagonal first, then the
// off-diagonal entries in row order (the operation order of the oracle's
// spmv); padding entries are skipped.  The sum runs in the storage type T:
// FP64 on the FP64 path, FP32 in the FP32 AMG levels.  XF supplies the
// vector entries (a plain array, or a prolongated one computed on the fly).
// All lanes of a wavefront hold slots of one slice (see RowRange), so the
// slice offsets are read with scalar loads.
template <class T> struct PlainX {
    static constexpr bool kPlain = true;
    const T* x;
    __device__ T operator()(int j) const { return x[j]; }
};
// 1.0 in device memory: the x operand of a padded slot in the pipelined row
// product.
template <class T> struct UnitValue {
    static __device__ const T one;
};
template <class T> __device__ const T UnitValue<T>::one = T(1);
// x + alpha xc[agg] (the kProlong update) evaluated where it is read.
template <class T> struct ProlongedX {
    static constexpr bool kPlain = false;
    const T* x;
    const int* agg;
    const T* xc;
    T alpha;
    __device__ T operator()(int j) const { return x[j] + alpha * xc[agg[j]]; }
};
// One row of A times x: the diagonal first, then the stored entries in entry
// order.  Sliced rows are fetched kC levels at a time so that their column,
// value and x loads overlap.  With kAhead the column and value loads of the
// next chunk are issued before the current chunk's x loads are consumed, so a
// row of m chunks waits for about m + 2 memory latencies instead of 3 m; the
// coarse AMG levels, whose Galerkin rows run to dozens of entries over few
// rows, are bound by that chain.  Both forms add the same products in the
// same order.
template <class T, bool kAhead, class XF> __device__ inline T rowDotAt(const CsrView<T>& A, const XF& xf, int q, int i) {
    T y = A.diag[i] * xf(i);
    if (!A.sliceOff) {
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) y += A.val[k] * xf(A.col[k]);
        return y;
    }
    const int* so = A.sliceOff + static_cast<long long>(__builtin_amdgcn_readfirstlane(q >> 6)) * (A.K + 1);
    const int lane = q & 63;
    constexpr int kC = 8;
    auto fetch = [&](int k0, int (&c)[kC], T (&v)[kC]) {
#pragma unroll
        for (int j = 0; j < kC; ++j) {
            const int e = (k0 + j < A.K && lane < so[k0 + j + 1] - so[k0 + j]) ? so[k0 + j] + lane : -1;
            c[j] = e >= 0 ? A.col[e] : -1;
            v[j] = e >= 0 ? A.val[e] : T(0);
        }
    };
    if constexpr (!kAhead) {
        for (int k0 = 0; k0 < A.K; k0 += kC) {
            if (so[k0 + 1] == so[k0]) break;  // no row of the slice has more entries
            int c[kC];
            T v[kC], xv[kC];
            fetch(k0, c, v);
#pragma unroll
            for (int j = 0; j < kC; ++j) xv[j] = c[j] >= 0 ? xf(c[j]) : T(0);
#pragma unroll
            for (int j = 0; j < kC; ++j)
                if (c[j] >= 0) y += v[j] * xv[j];
        }
    } else {
        // Branch-free: a padded slot reads entry 0 and the constant 1.0 in
        // place of x and adds (-0.0) * 1.0, and y + (-0.0) is y for every y,
        // so each load is issued unconditionally, none is sunk into a
        // branch, and the waits count only the loads a product needs.  A
        // fetch keeps the raw loads and a mask of the slots in use; two
        // register sets alternate, so nothing in flight is copied.
        static_assert(XF::kPlain, "the pipelined row product reads x directly");
        if (A.K == 0 || so[1] == so[0]) return y;
        auto fetchAll = [&](int k0, int (&c)[kC], T (&v)[kC]) {
            unsigned m = 0;
#pragma unroll
            for (int j = 0; j < kC; ++j) {
                const int kk = min(k0 + j, A.K - 1);
                const bool in = (k0 + j < A.K) & (lane < so[kk + 1] - so[kk]);
                const int e = in ? so[kk] + lane : 0;
                c[j] = A.col[e];
                v[j] = A.val[e];
                m |= static_cast<unsigned>(in) << j;
            }
            return m;
        };
        auto hasChunk = [&](int k0) {
            const int kk = min(k0, A.K - 1);
            return (k0 < A.K) & (so[kk + 1] != so[kk]);
        };
        auto loadX = [&](unsigned m, const int (&c)[kC], T (&xv)[kC]) {
#pragma unroll
            for (int j = 0; j < kC; ++j) xv[j] = *(((m >> j) & 1u) && c[j] >= 0 ? xf.x + c[j] : &UnitValue<T>::one);
        };
        // A slot inside a level's width may hold no entry (column -1) when
        // the slice's rows are not sorted by length.
        auto consume = [&](unsigned m, const int (&c)[kC], const T (&v)[kC], const T (&xv)[kC]) {
#pragma unroll
            for (int j = 0; j < kC; ++j) y = y + (((m >> j) & 1u) && c[j] >= 0 ? v[j] : T(-0.0)) * xv[j];
        };
        int ca[kC], cb[kC];
        T va[kC], vb[kC], xv[kC];
        unsigned ma = fetchAll(0, ca, va), mb;
        for (int k0 = 0;; k0 += 2 * kC) {
            const bool moreB = hasChunk(k0 + kC);
            loadX(ma, ca, xv);
            __builtin_amdgcn_sched_barrier(0);  // this chunk's x loads issue first
            mb = fetchAll(k0 + kC, cb, vb);  // an empty mask past the row's end
            __builtin_amdgcn_sched_barrier(0);
            consume(ma, ca, va, xv);
            if (!moreB) break;
            const bool moreA = hasChunk(k0 + 2 * kC);
            loadX(mb, cb, xv);
            __builtin_amdgcn_sched_barrier(0);
            ma = fetchAll(k0 + 2 * kC, ca, va);
            __builtin_amdgcn_sched_barrier(0);
            consume(mb, cb, vb, xv);
            if (!moreA) break;
        }
    }
    return y;
}
template <class T, class XF> __device__ inline T rowDotG(const CsrView<T>& A, const XF& xf, int q, int i) {
    return rowDotAt<T, false>(A, xf, q, i);
}
template <class T, bool kAhead = false> __device__ inline T rowDot(const CsrView<T>& A, const T* x, int q, int i) {
    return rowDotAt<T, kAhead>(A, PlainX<T>{x}, q, i);
}
// Sum of r over the members mi[mp[I]] .. mi[mp[I + 1] - 1] in member order;
// the member indices are fetched kC at a time so that the loads of r overlap.
template <class T> __device__ inline T memberSum(const int* mp, const int* mi, const T* r, int I) {
    const int m0 = mp[I], m1 = mp[I + 1];
    constexpr int kC = 8;
    T s = 0;
    for (int k0 = m0; k0 < m1; k0 += kC) {
        int idx[kC];
        T rv[kC];
#pragma unroll
        for (int j = 0; j < kC; ++j) idx[j] = k0 + j < m1 ? mi[k0 + j] : -1;
#pragma unroll
        for (int j = 0; j < kC; ++j) rv[j] = idx[j] >= 0 ? r[idx[j]] : T(0);
#pragma unroll
        for (int j = 0; j < kC; ++j)
            if (idx[j] >= 0) s += rv[j];
    }
    return s;
}
// Krylov kernels run on matrices whose slot q holds row q.
__device__ inline double rowDot64(const CsrView<double>& A, const double* x, int i) { return rowDot(A, x, i, i); }

__global__ void __launch_bounds__(kBlock) kFill(int n, double* x, double v) {
    const int i = globalThread();
    if (i < n) x[i] = v;
}
// Reads a[0, n) and stores nothing unless the sum is non-zero (a is zero
// filled by its caller): a read-only pass that leaves no dirty lines.
__global__ void __launch_bounds__(kBlock) kTouch(long long n, const double* a, double* sink) {
    double acc = 0.0;
    for (long long i = globalThread(); i < n; i += static_cast<long long>(gridDim.x) * blockDim.x) acc += a[i];
    if (acc != 0.0) sink[0] = acc;
}
__global__ void __launch_bounds__(kBlock) kCopy(int n, const double* a, double* b) {
    const int i = globalThread();
    if (i < n) b[i] = a[i];
}
__global__ void __launch_bounds__(kBlock) kAxpy(int n, double a, const double* x, double* y) {  // y += a x
    const int i = globalThread();
    if (i < n) y[i] += a * x[i];
}

// Sum of squares of b (optionally copied to r); partial sums per block.
__global__ void __launch_bounds__(kBlock) kCopyNorm(int n, const double* b, double* r, double* part) {
    __shared__ double smem[kBlock / 64];
    double acc = 0.0;
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x) {
        const double v = b[i];
        if (r) r[i] = v;
        acc += v * v;
    }
    double v[1] = {acc};
    blockPartial<1>(v, part, smem);
}

// r = b - A x and the partial sums of its squares.
__global__ void __launch_bounds__(kBlock) kResidualNorm(CsrView<double> A, const double* b, const double* x, double* r, double* part) {
    __shared__ double smem[kBlock / 64];
    double acc = 0.0;
    for (int i = globalThread(); i < A.n; i += gridDim.x * blockDim.x) {
        const double v = b[i] - rowDot64(A, x, i);
        r[i] = v;
        acc += v * v;
    }
    double v[1] = {acc};
    blockPartial<1>(v, part, smem);
}

__global__ void __launch_bounds__(kBlock) kDot(int n, const double* a, const double* b, double* part) {
    __shared__ double smem[kBlock / 64];
    double acc = 0.0;
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x) acc += a[i] * b[i];
    double v[1] = {acc};
    blockPartial<1>(v, part, smem);
}

struct StoreSqrt {  // sc[idx] = sqrt(sum)
    double* sc;
    int idx;
    __device__ void operator()(const double (&v)[1]) const { sc[idx] = sqrt(v[0]); }
};
struct StoreValue {  // sc[idx] = sum
    double* sc;
    int idx;
    __device__ void operator()(const double (&v)[1]) const { sc[idx] = v[0]; }
};

// ---- GMRES / FGMRES (gmres.c, fgmres/fgmres.c, borthog2.c) ----
// The Krylov basis is stored as consecutive vectors of `ld` doubles.
constexpr int kGmDots = 8;  // inner products formed per pass over the new vector
// Partial sums of w . V[j0 + k] for k < cnt (the other slots hold zeros), each
// in the order of kDot.
__global__ void __launch_bounds__(kBlock) kMultiDot(int n, const double* w, const double* V, long long ld, int j0, int cnt,
                                                    double* part) {
    __shared__ double smem[kGmDots * kBlock / 64];
    double acc[kGmDots];
    for (int k = 0; k < kGmDots; ++k) acc[k] = 0.0;
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x) {
        const double wi = w[i];
        for (int k = 0; k < cnt; ++k) acc[k] += wi * V[(j0 + k) * ld + i];
    }
    blockPartial<kGmDots>(acc, part, smem);
}
// Classical Gram-Schmidt update v += sum_j (-h_j) V_j (VecMAXPY with
// lhh = -<v, V_j>, j in order) and the partial sums of the updated v . v.
__global__ void __launch_bounds__(kBlock) kGmMaxpyNorm(int n, double* v, const double* V, long long ld, int cnt,
                                                       const double* h, double* part) {
    __shared__ double smem[kBlock / 64];
    double acc = 0.0;
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x) {
        double s = v[i];
        for (int j = 0; j < cnt; ++j) s += -h[j] * V[j * ld + i];
        v[i] = s;
        acc += s * s;
    }
    double a[1] = {acc};
    blockPartial<1>(a, part, smem);
}
__global__ void __launch_bounds__(kBlock) kScale(int n, double* x, double a) {  // VecScale
    const int i = globalThread();
    if (i < n) x[i] *= a;
}
struct GmCoef {
    double c[kGmRestart];
};
// out = sum_j c_j V_j, j in order from a zero start (VecMAXPBY with beta 0).
__global__ void __launch_bounds__(kBlock) kGmCombine(int n, double* out, const double* V, long long ld, int cnt, GmCoef c) {
    const int i = globalThread();
    if (i >= n) return;
    double s = 0.0;
    for (int j = 0; j < cnt; ++j) s += c.c[j] * V[j * ld + i];
    out[i] = s;
}

// ---- CG (cg.c KSPSolve_CG, unpreconditioned norm) ----
// Stage-1 kernels return at once when a breakdown flag is set; their tails
// check the same flag so that no stale partials are consumed.
struct CgBeta {  // beta = z . r
    double* sc;
    int* fl;
    int checkSign;
    __device__ void operator()(const double (&v)[1]) const {
        if (fl[0]) return;
        const double beta = v[0];
        sc[S_BETA] = beta;
        if (!isfinite(beta)) fl[0] |= F_NONFINITE;
        else if (beta == 0.0) fl[0] |= F_BETA_ZERO;
        else if (checkSign && beta * sc[S_BETAOLD] < 0.0) fl[0] |= F_INDEF_PC;
    }
};

// z = r * (1/d) and the partial sums of z . r
__global__ void __launch_bounds__(kBlock) kJacobiDot(int n, const double* r, const double* jinv, double* z, const int* fl, double* part) {
    __shared__ double smem[kBlock / 64];
    if (fl[0]) return;
    double acc = 0.0;
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x) {
        const double v = r[i] * jinv[i];
        z[i] = v;
        acc += v * r[i];
    }
    double v[1] = {acc};
    blockPartial<1>(v, part, smem);
}
__global__ void __launch_bounds__(kBlock) kPlainDot(int n, const double* z, const double* r, const int* fl, double* part) {
    __shared__ double smem[kBlock / 64];
    if (fl[0]) return;
    double acc = 0.0;
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x) acc += z[i] * r[i];
    double v[1] = {acc};
    blockPartial<1>(v, part, smem);
}

__global__ void __launch_bounds__(kBlock) kCgDirection(int n, const double* z, double* p, const double* sc, const int* fl, int first) {
    if (fl[0]) return;
    const int i = globalThread();
    if (i >= n) return;
    if (first) p[i] = z[i];
    else {
        const double bb = sc[S_BETA] / sc[S_BETAOLD];
        p[i] = z[i] + bb * p[i];
    }
}

// w = A p and the partial sums of p . w
__global__ void __launch_bounds__(kBlock) kCgSpmvDot(CsrView<double> A, const double* p, double* w, const int* fl, double* part) {
    __shared__ double smem[kBlock / 64];
    if (fl[0]) return;
    double acc = 0.0;
    for (int i = globalThread(); i < A.n; i += gridDim.x * blockDim.x) {
        const double y = rowDot64(A, p, i);
        w[i] = y;
        acc += p[i] * y;
    }
    double v[1] = {acc};
    blockPartial<1>(v, part, smem);
}
struct CgAlpha {  // dpi = p . w, a = beta / dpi
    double* sc;
    int* fl;
    int iter;
    __device__ void operator()(const double (&v)[1]) const {
        if (fl[0]) return;
        const double dpiold = sc[S_DPI], dpi = v[0];
        sc[S_DPI] = dpi;
        if (!isfinite(dpi)) { fl[0] |= F_NONFINITE; return; }
        sc[S_BETAOLD] = sc[S_BETA];
        if (dpi == 0.0 || (iter > 0 && sgn(dpi) * sgn(dpiold) < 0)) fl[0] |= F_INDEF_MAT;
        else sc[S_A] = sc[S_BETA] / dpi;
    }
};

// x += a p, r += -a w and the partial sums of r . r
__global__ void __launch_bounds__(kBlock) kCgUpdate(int n, double* x, const double* p, double* r, const double* w, const double* sc, const int* fl,
                          double* part) {
    __shared__ double smem[kBlock / 64];
    if (fl[0]) return;
    const double a = sc[S_A];
    double acc = 0.0;
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x) {
        x[i] += a * p[i];
        const double ri = r[i] + -a * w[i];
        r[i] = ri;
        acc += ri * ri;
    }
    double v[1] = {acc};
    blockPartial<1>(v, part, smem);
}
struct CgNorm {  // dp = ||r||
    double* sc;
    const int* fl;
    __device__ void operator()(const double (&v)[1]) const {
        if (fl[0]) return;
        sc[S_DP] = sqrt(v[0]);
    }
};

// ---- BiCGSTAB (bcgs.c:14-140, right preconditioning) ----
__global__ void __launch_bounds__(kBlock) kBcgsRho(int n, const double* r, const double* rp, double* part) {
    __shared__ double smem[kBlock / 64];
    double acc = 0.0;
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x) acc += r[i] * rp[i];
    double v[1] = {acc};
    blockPartial<1>(v, part, smem);
}
struct BcgsRho {  // rho = r . rp, beta = (rho / rhoold) (alpha / omegaold)
    double* sc;
    __device__ void operator()(const double (&v)[1]) const {
        const double rho = v[0];
        sc[S_RHO] = rho;
        sc[S_BETA] = (rho / sc[S_RHOOLD]) * (sc[S_ALPHA] / sc[S_OMEGAOLD]);
    }
};
__global__ void __launch_bounds__(kBlock) kBcgsP(int n, double* p, const double* r, const double* v, const double* sc) {
    const int i = globalThread();
    if (i >= n) return;
    const double beta = sc[S_BETA];
    p[i] = r[i] + (-sc[S_OMEGAOLD] * beta) * v[i] + beta * p[i];
}
// v = A t and the partial sums of v . rp
__global__ void __launch_bounds__(kBlock) kBcgsSpmvD1(CsrView<double> A, const double* t, double* v, const double* rp, double* part) {
    __shared__ double smem[kBlock / 64];
    double acc = 0.0;
    for (int i = globalThread(); i < A.n; i += gridDim.x * blockDim.x) {
        const double y = rowDot64(A, t, i);
        v[i] = y;
        acc += y * rp[i];
    }
    double vv[1] = {acc};
    blockPartial<1>(vv, part, smem);
}
struct BcgsAlpha {  // d1 = v . rp, alpha = rho / d1
    double* sc;
    int* fl;
    __device__ void operator()(const double (&v)[1]) const {
        const double d1 = v[0];
        sc[S_D1] = d1;
        if (!isfinite(d1)) fl[0] |= F_NONFINITE;
        else if (d1 == 0.0) fl[0] |= F_D1_ZERO;
        else sc[S_ALPHA] = sc[S_RHO] / d1;
    }
};
__global__ void __launch_bounds__(kBlock) kBcgsS(int n, double* s, const double* v, const double* r, const double* sc, const int* fl) {
    if (fl[0]) return;
    const int i = globalThread();
    if (i >= n) return;
    s[i] = -sc[S_ALPHA] * v[i] + r[i];
}
// t = A rt and the partial sums of s . t and t . t
__global__ void __launch_bounds__(kBlock) kBcgsSpmvD2(CsrView<double> A, const double* rt, double* t, const double* s, const int* fl,
                            double* part) {
    __shared__ double smem[2 * kBlock / 64];
    if (fl[0]) return;
    double a1 = 0.0, a2 = 0.0;
    for (int i = globalThread(); i < A.n; i += gridDim.x * blockDim.x) {
        const double y = rowDot64(A, rt, i);
        t[i] = y;
        a1 += s[i] * y;
        a2 += y * y;
    }
    double v[2] = {a1, a2};
    blockPartial<2>(v, part, smem);
}
struct BcgsOmega {  // omega = (s . t) / (t . t)
    double* sc;
    int* fl;
    __device__ void operator()(const double (&v)[2]) const {
        if (fl[0]) return;
        sc[S_D1] = v[0];
        sc[S_D2] = v[1];
        if (v[1] == 0.0) fl[0] |= F_D2_ZERO;
        else sc[S_OMEGA] = v[0] / v[1];
    }
};
// x = alpha p + omega s + x, r = -omega t + s and the partial sums of r . r
__global__ void __launch_bounds__(kBlock) kBcgsUpdate(int n, double* x, const double* p, const double* s, double* r, const double* t,
                            const double* sc, const int* fl, double* part) {
    __shared__ double smem[kBlock / 64];
    if (fl[0]) return;
    const double alpha = sc[S_ALPHA], omega = sc[S_OMEGA];
    double acc = 0.0;
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x) {
        x[i] = alpha * p[i] + omega * s[i] + x[i];
        const double ri = -omega * t[i] + s[i];
        r[i] = ri;
        acc += ri * ri;
    }
    double v[1] = {acc};
    blockPartial<1>(v, part, smem);
}
struct BcgsNorm {  // dp = ||r||; rho and omega become the previous values
    double* sc;
    const int* fl;
    __device__ void operator()(const double (&v)[1]) const {
        if (fl[0]) return;
        sc[S_DP] = sqrt(v[0]);
        sc[S_RHOOLD] = sc[S_RHO];
        sc[S_OMEGAOLD] = sc[S_OMEGA];
    }
};
// ---- BiCGSTAB on the three components of a vector equation in lockstep ----
// Component c runs the recurrence of the kernels above on its own vectors,
// scalars (sc3 + c * S_COUNT) and flags (fl3[c]); act[c] == 0 freezes it.  Each
// component's arithmetic and reduction order are those of a solo solve, so
// the three results equal three sequential solves bit for bit, while the
// matrix is streamed once per product.  T is the storage type of the vectors
// and the matrix (double, or float inside the mixed-precision refinement);
// arithmetic is done in double and every sum accumulates in double.
template <class T> struct Vec3p {
    T* c[3];
};
template <class T> struct CVec3p {
    const T* c[3];
};

// Row i of A x_c for the components with m[c] set: the chunk's columns and
// values are loaded once for all of them, and each sum runs in the solo order
// (in the storage type, like rowDot).
template <class T>
__device__ inline void rowDot3(const CsrView<T>& A, const CVec3p<T>& x, const bool (&m)[3], int i, T (&y)[3]) {
    for (int c = 0; c < 3; ++c) y[c] = m[c] ? A.diag[i] * x.c[c][i] : T(0);
    if (A.sliceOff) {
        const RowRange R(A.sliceOff, A.K, i);
        constexpr int kC = 8;
        for (int k0 = 0; k0 < A.K; k0 += kC) {
            if (R.so[k0 + 1] == R.so[k0]) break;
            int e[kC], col[kC];
            T v[kC];
#pragma unroll
            for (int j = 0; j < kC; ++j) e[j] = R.valid(k0 + j) ? R.at(k0 + j) : -1;
#pragma unroll
            for (int j = 0; j < kC; ++j) {
                col[j] = e[j] >= 0 ? A.col[e[j]] : -1;
                v[j] = e[j] >= 0 ? A.val[e[j]] : T(0);
            }
            for (int c = 0; c < 3; ++c) {
                if (!m[c]) continue;
                T xv[kC];
#pragma unroll
                for (int j = 0; j < kC; ++j) xv[j] = col[j] >= 0 ? x.c[c][col[j]] : T(0);
#pragma unroll
                for (int j = 0; j < kC; ++j)
                    if (col[j] >= 0) y[c] += v[j] * xv[j];
            }
        }
    } else {
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
            for (int c = 0; c < 3; ++c)
                if (m[c]) y[c] += A.val[k] * x.c[c][A.col[k]];
    }
}

template <class T> __global__ void __launch_bounds__(kBlock) k3CopyNorm(int n, CVec3p<T> b, Vec3p<T> r, const int* act, double* part) {
    __shared__ double smem[3 * kBlock / 64];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x)
        for (int c = 0; c < 3; ++c) {
            if (!act[c]) continue;
            const T v = b.c[c][i];
            if (r.c[c]) r.c[c][i] = v;
            acc[c] += static_cast<double>(v) * static_cast<double>(v);
        }
    blockPartial<3>(acc, part, smem);
}
template <class T>
__global__ void __launch_bounds__(kBlock) k3ResidualNorm(CsrView<T> A, CVec3p<T> b, CVec3p<T> x, Vec3p<T> r, const int* act, double* part) {
    __shared__ double smem[3 * kBlock / 64];
    double acc[3] = {0.0, 0.0, 0.0};
    const bool m[3] = {act[0] != 0, act[1] != 0, act[2] != 0};
    for (int i = globalThread(); i < A.n; i += gridDim.x * blockDim.x) {
        T ax[3];
        rowDot3(A, x, m, i, ax);
        for (int c = 0; c < 3; ++c) {
            if (!m[c]) continue;
            const T v = b.c[c][i] - ax[c];
            r.c[c][i] = v;
            acc[c] += static_cast<double>(v) * static_cast<double>(v);
        }
    }
    blockPartial<3>(acc, part, smem);
}
template <class T> __global__ void __launch_bounds__(kBlock) k3Dot(int n, CVec3p<T> a, CVec3p<T> b, const int* act, double* part) {
    __shared__ double smem[3 * kBlock / 64];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x)
        for (int c = 0; c < 3; ++c)
            if (act[c]) acc[c] += static_cast<double>(a.c[c][i]) * static_cast<double>(b.c[c][i]);
    blockPartial<3>(acc, part, smem);
}
template <class T> __global__ void __launch_bounds__(kBlock) k3Copy(int n, CVec3p<T> a, Vec3p<T> b, const int* act) {
    const int i = globalThread();
    if (i >= n) return;
    for (int c = 0; c < 3; ++c)
        if (act[c]) b.c[c][i] = a.c[c][i];
}
template <class T> __global__ void __launch_bounds__(kBlock) k3Fill(int n, Vec3p<T> a, double v, const int* act) {
    const int i = globalThread();
    if (i >= n) return;
    for (int c = 0; c < 3; ++c)
        if (act[c]) a.c[c][i] = static_cast<T>(v);
}
template <class T> __global__ void __launch_bounds__(kBlock) k3Axpy(int n, double a, CVec3p<T> x, Vec3p<T> y, const int* act) {  // y += a x
    const int i = globalThread();
    if (i >= n) return;
    for (int c = 0; c < 3; ++c)
        if (act[c]) y.c[c][i] = static_cast<T>(static_cast<double>(y.c[c][i]) + a * static_cast<double>(x.c[c][i]));
}
// z = r * (1/d), or a copy without a preconditioner
template <class T> __global__ void __launch_bounds__(kBlock) k3Jacobi(int n, CVec3p<T> r, const T* jinv, Vec3p<T> z, const int* act) {
    const int i = globalThread();
    if (i >= n) return;
    for (int c = 0; c < 3; ++c)
        if (act[c])
            z.c[c][i] = jinv ? static_cast<T>(static_cast<double>(r.c[c][i]) * static_cast<double>(jinv[i])) : r.c[c][i];
}
// The BiCGSTAB direction p = r + beta (p - omega v), then the Jacobi apply
// z = p / d (or a copy), in one pass: z is formed from the stored p.
template <class T>
__global__ void __launch_bounds__(kBlock) k3BcgsPJacobi(int n, Vec3p<T> p, CVec3p<T> r, CVec3p<T> v, const double* sc3,
                                                        const T* jinv, Vec3p<T> z, const int* act) {
    const int i = globalThread();
    if (i >= n) return;
    for (int c = 0; c < 3; ++c) {
        if (!act[c]) continue;
        const double* sc = sc3 + c * S_COUNT;
        const double beta = sc[S_BETA];
        const T pv = static_cast<T>(static_cast<double>(r.c[c][i]) + (-sc[S_OMEGAOLD] * beta) * static_cast<double>(v.c[c][i]) +
                                    beta * static_cast<double>(p.c[c][i]));
        p.c[c][i] = pv;
        z.c[c][i] = jinv ? static_cast<T>(static_cast<double>(pv) * static_cast<double>(jinv[i])) : pv;
    }
}
template <class T>
__global__ void __launch_bounds__(kBlock) k3SpmvD1(CsrView<T> A, CVec3p<T> t, Vec3p<T> v, CVec3p<T> rp, const int* act, double* part) {
    __shared__ double smem[3 * kBlock / 64];
    double acc[3] = {0.0, 0.0, 0.0};
    const bool m[3] = {act[0] != 0, act[1] != 0, act[2] != 0};
    for (int i = globalThread(); i < A.n; i += gridDim.x * blockDim.x) {
        T y[3];
        rowDot3(A, t, m, i, y);
        for (int c = 0; c < 3; ++c) {
            if (!m[c]) continue;
            const T yv = y[c];
            v.c[c][i] = yv;
            acc[c] += static_cast<double>(yv) * static_cast<double>(rp.c[c][i]);
        }
    }
    blockPartial<3>(acc, part, smem);
}
// s = r - alpha v (components not yet stopped), then the Jacobi apply z = s / d
// (or a copy) for every active component, from s as it stands after the
// update, in one pass.
template <class T>
__global__ void __launch_bounds__(kBlock) k3BcgsSJacobi(int n, Vec3p<T> s, CVec3p<T> v, CVec3p<T> r, const double* sc3,
                                                        const int* fl3, const T* jinv, Vec3p<T> z, const int* act) {
    const int i = globalThread();
    if (i >= n) return;
    for (int c = 0; c < 3; ++c) {
        if (!act[c]) continue;
        T sv = s.c[c][i];
        if (!fl3[c]) {
            sv = static_cast<T>(-sc3[c * S_COUNT + S_ALPHA] * static_cast<double>(v.c[c][i]) + static_cast<double>(r.c[c][i]));
            s.c[c][i] = sv;
        }
        z.c[c][i] = jinv ? static_cast<T>(static_cast<double>(sv) * static_cast<double>(jinv[i])) : sv;
    }
}
// Partials: s . t of the three components at k = 0..2, t . t at k = 3..5.
template <class T>
__global__ void __launch_bounds__(kBlock) k3SpmvD2(CsrView<T> A, CVec3p<T> rt, Vec3p<T> t, CVec3p<T> s, const int* fl3, const int* act,
                         double* part) {
    __shared__ double smem[6 * kBlock / 64];
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const bool m[3] = {act[0] && !fl3[0], act[1] && !fl3[1], act[2] && !fl3[2]};
    for (int i = globalThread(); i < A.n; i += gridDim.x * blockDim.x) {
        T y[3];
        rowDot3(A, rt, m, i, y);
        for (int c = 0; c < 3; ++c) {
            if (!m[c]) continue;
            const T yv = y[c];
            const double yd = static_cast<double>(yv);
            t.c[c][i] = yv;
            acc[c] += static_cast<double>(s.c[c][i]) * yd;
            acc[3 + c] += yd * yd;
        }
    }
    blockPartial<6>(acc, part, smem);
}
template <class T>
__global__ void __launch_bounds__(kBlock) k3BcgsUpdate(int n, Vec3p<T> x, CVec3p<T> p, CVec3p<T> s, Vec3p<T> r, CVec3p<T> t, const double* sc3,
                             const int* fl3, const int* act, double* part) {
    __shared__ double smem[3 * kBlock / 64];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x)
        for (int c = 0; c < 3; ++c) {
            if (!act[c] || fl3[c]) continue;
            const double alpha = sc3[c * S_COUNT + S_ALPHA], omega = sc3[c * S_COUNT + S_OMEGA];
            x.c[c][i] = static_cast<T>(alpha * static_cast<double>(p.c[c][i]) + omega * static_cast<double>(s.c[c][i]) +
                                       static_cast<double>(x.c[c][i]));
            const T ri = static_cast<T>(-omega * static_cast<double>(t.c[c][i]) + static_cast<double>(s.c[c][i]));
            r.c[c][i] = ri;
            acc[c] += static_cast<double>(ri) * static_cast<double>(ri);
        }
    blockPartial<3>(acc, part, smem);
}
struct Tail3Sqrt {  // sc[idx] = sqrt(sum) per active component
    double* sc3;
    int idx;
    const int* act;
    __device__ void operator()(const double (&v)[3]) const {
        for (int c = 0; c < 3; ++c)
            if (act[c]) sc3[c * S_COUNT + idx] = sqrt(v[c]);
    }
};
struct Tail3Rho {
    double* sc3;
    const int* act;
    __device__ void operator()(const double (&v)[3]) const {
        for (int c = 0; c < 3; ++c)
            if (act[c]) { const double one[1] = {v[c]}; BcgsRho{sc3 + c * S_COUNT}(one); }
    }
};
struct Tail3Alpha {
    double* sc3;
    int* fl3;
    const int* act;
    __device__ void operator()(const double (&v)[3]) const {
        for (int c = 0; c < 3; ++c)
            if (act[c]) { const double one[1] = {v[c]}; BcgsAlpha{sc3 + c * S_COUNT, fl3 + c}(one); }
    }
};
struct Tail3Omega {
    double* sc3;
    int* fl3;
    const int* act;
    __device__ void operator()(const double (&v)[6]) const {
        for (int c = 0; c < 3; ++c) {
            if (!act[c]) continue;
            const double pair[2] = {v[c], v[3 + c]};
            BcgsOmega{sc3 + c * S_COUNT, fl3 + c}(pair);
        }
    }
};
struct Tail3Norm {
    double* sc3;
    const int* fl3;
    const int* act;
    __device__ void operator()(const double (&v)[3]) const {
        for (int c = 0; c < 3; ++c)
            if (act[c]) { const double one[1] = {v[c]}; BcgsNorm{sc3 + c * S_COUNT, fl3 + c}(one); }
    }
};

__global__ void __launch_bounds__(kBlock) kSpmvPlain(CsrView<double> A, const double* x, double* y) {
    const int i = globalThread();
    if (i < A.n) y[i] = rowDot64(A, x, i);
}
// ---- block Jacobi: ILU(0) inside blocks of 64 consecutive rows ----
// PETSc's bjacobi with one block per 64 rows and ILU(0) in each block
// (petsc_backend.cpp:46 names it; PCBJACOBI with -pc_bjacobi_blocks
// ceil(n/64), sub-PC ILU): one wavefront per block, lane r on row 64*b + r.
// Only entries whose column lies in the row's own block take part; the
// factors keep the matrix's pattern (no fill).  Rows past n in the last block
// are identity rows.
constexpr int kIluBlock = 64;
__device__ inline unsigned long long readlane64(unsigned long long v, int lane) {
    const unsigned lo = __builtin_amdgcn_readlane(static_cast<unsigned>(v), lane);
    const unsigned hi = __builtin_amdgcn_readlane(static_cast<unsigned>(v >> 32), lane);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}
__device__ inline double readlaneD(double v, int lane) {
    return __longlong_as_double(static_cast<long long>(readlane64(static_cast<unsigned long long>(__double_as_longlong(v)), lane)));
}
// Calls f(e) for the stored off-diagonal entries of row i (lane `lane` of
// block `blk`): the block's slice in the sliced layout, the CSR row otherwise.
template <class F> __device__ inline void forRowEntries(const CsrView<double>& A, int blk, int lane, int i, F f) {
    if (i >= A.n) return;
    if (A.sliceOff) {
        const int* so = A.sliceOff + static_cast<long long>(blk) * (A.K + 1);
        for (int k = 0; k < A.K; ++k) {
            const int width = so[k + 1] - so[k];
            if (width == 0) break;
            if (lane < width && A.col[so[k] + lane] >= 0) f(so[k] + lane);
        }
    } else {
        for (int e = A.ptr[i]; e < A.ptr[i + 1]; ++e) f(e);
    }
}
// Factorisation (IKJ order): for k = 0..62, every row r > k with (r, k) in the
// pattern takes l = a_rk / a_kk and a_rj -= l * a_kj for each j > k present in
// rows r and k.  fd: the U diagonal; fv: the factor at each in-block entry
// (L strictly below, U strictly above; a repeated column keeps it in its first
// entry and 0 in the others).
__global__ void __launch_bounds__(kIluBlock) kBlockIluFactor(CsrView<double> A, double* fd, double* fv) {
    __shared__ double T[kIluBlock][kIluBlock + 1];  // padded rows: lanes reading one column hit distinct banks
    const int blk = blockIdx.x, r = threadIdx.x, base = blk * kIluBlock, i = base + r;
    for (int j = 0; j < kIluBlock; ++j) T[r][j] = 0.0;
    unsigned long long mask = 1ull << r;
    T[r][r] = i < A.n ? A.diag[i] : 1.0;
    forRowEntries(A, blk, r, i, [&](int e) {
        const int c = A.col[e] - base;
        if (c >= 0 && c < kIluBlock) { T[r][c] += A.val[e]; mask |= 1ull << c; }
    });
    __syncthreads();
    for (int k = 0; k < kIluBlock - 1; ++k) {
        const unsigned long long mk = readlane64(mask, k);
        if (r > k && ((mask >> k) & 1ull)) {
            const double l = T[r][k] / T[k][k];
            T[r][k] = l;
            unsigned long long upd = mask & mk & ~((2ull << k) - 1ull);
            while (upd) {
                const int j = __builtin_ctzll(upd);
                upd &= upd - 1ull;
                T[r][j] -= l * T[k][j];
            }
        }
        __syncthreads();
    }
    if (i >= A.n) return;
    fd[i] = T[r][r];
    unsigned long long written = 1ull << r;
    forRowEntries(A, blk, r, i, [&](int e) {
        const int c = A.col[e] - base;
        if (c >= 0 && c < kIluBlock) {
            fv[e] = ((written >> c) & 1ull) ? 0.0 : T[r][c];
            written |= 1ull << c;
        }
    });
}
// z = (LU)^-1 r block by block.  Lane r holds its factor row in registers
// (64 doubles, statically indexed); forward and backward substitution
// broadcast each finished unknown with readlane.  Zero factors are skipped,
// so each row's sum runs over its pattern in column order.
__global__ void __launch_bounds__(kIluBlock) kBlockIluApply(CsrView<double> A, const double* fd, const double* fv,
                                                           const double* rr, double* zz) {
    const int blk = blockIdx.x, r = threadIdx.x, base = blk * kIluBlock, i = base + r;
    double T[kIluBlock];
#pragma unroll
    for (int j = 0; j < kIluBlock; ++j) T[j] = 0.0;
    unsigned long long cols = 0;  // in-block columns of this row's entries
    forRowEntries(A, blk, r, i, [&](int e) {
        const int c = A.col[e] - base;
        if (c >= 0 && c < kIluBlock) {
            const double v = fv[e];
            cols |= 1ull << c;
#pragma unroll
            for (int j = 0; j < kIluBlock; ++j) T[j] = c == j ? T[j] + v : T[j];
        }
    });
    // Columns some row of the block refers to (an xor-butterfly OR): a
    // substitution step on any other column changes no row and is skipped.
    unsigned lo = static_cast<unsigned>(cols), hi = static_cast<unsigned>(cols >> 32);
    for (int off = 32; off > 0; off >>= 1) {
        lo |= static_cast<unsigned>(__shfl_xor(static_cast<int>(lo), off, 64));
        hi |= static_cast<unsigned>(__shfl_xor(static_cast<int>(hi), off, 64));
    }
    const unsigned long long used = (static_cast<unsigned long long>(__builtin_amdgcn_readfirstlane(hi)) << 32) |
                                    __builtin_amdgcn_readfirstlane(lo);
    const double d = i < A.n ? fd[i] : 1.0;
    double y = i < A.n ? rr[i] : 0.0;
#pragma unroll
    for (int s = 0; s < kIluBlock - 1; ++s) {
        if (!((used >> s) & 1ull)) continue;
        const double ys = readlaneD(y, s);
        const double t = r > s ? T[s] : 0.0;
        if (t != 0.0) y = y - t * ys;
    }
#pragma unroll
    for (int s = kIluBlock - 1; s >= 0; --s) {
        if (r == s) y = y / d;
        if (!((used >> s) & 1ull)) continue;
        const double xs = readlaneD(y, s);
        const double t = r < s ? T[s] : 0.0;
        if (t != 0.0) y = y - t * xs;
    }
    if (i < A.n) zz[i] = y;
}
__global__ void __launch_bounds__(kBlock) kJacobiApply(int n, const double* r, const double* jinv, double* z) {
    const int i = globalThread();
    if (i < n) z[i] = r[i] * jinv[i];
}
__global__ void __launch_bounds__(kBlock) kInvDiag(int n, const double* d, double* inv) {
    const int i = globalThread();
    if (i < n) inv[i] = 1.0 / d[i];
}
// Counts entries where (a0, a1) and (b0, b1) differ, as partial sums per block.
__global__ void __launch_bounds__(kBlock) kCompare(int n0, const double* a0, const double* b0, int n1, const double* a1, const double* b1,
                         double* part) {
    __shared__ double smem[kBlock / 64];
    double cnt = 0.0;
    const int n = n0 + n1;
    for (int i = globalThread(); i < n; i += gridDim.x * blockDim.x) {
        const bool same = i < n0 ? a0[i] == b0[i] : a1[i - n0] == b1[i - n0];
        if (!same) cnt += 1.0;
    }
    double v[1] = {cnt};
    blockPartial<1>(v, part, smem);
}

// ---- AMG ----
// The AMG kernels below compute in the storage type T of their level.
template <class T> __global__ void __launch_bounds__(kBlock) kJacobiFirst(int n, const T* d, const T* b, T* x, double omega) {
    const int i = globalThread();
    if (i < n) x[i] = static_cast<T>(omega) * b[i] / d[i];
}
// kAhead selects the pipelined row product (coarse levels, see rowDotAt).
template <class T, bool kAhead = false>
__global__ void __launch_bounds__(kBlock) kJacobiSweep(CsrView<T> A, const T* b, const T* x, T* xo, double omega) {
    const int q = globalThread();
    if (q >= A.n) return;
    const int i = A.perm ? A.perm[q] : q;
    const T ax = rowDot<T, kAhead>(A, x, q, i);
    xo[i] = x[i] + static_cast<T>(omega) * (b[i] - ax) / A.diag[i];
}
// The last fine-level sweep of a preconditioner application inside CG, with
// the partial sums of z . r (z the result, r = b) that kPlainDot would form.
__global__ void __launch_bounds__(kBlock) kJacobiSweepDot(CsrView<double> A, const double* b, const double* x, double* xo, double omega,
                                double* part) {
    __shared__ double smem[kBlock / 64];
    double acc = 0.0;
    for (int q = globalThread(); q < A.n; q += gridDim.x * blockDim.x) {
        const int i = A.perm ? A.perm[q] : q;
        const double ax = rowDot(A, x, q, i);
        const double z = x[i] + omega * (b[i] - ax) / A.diag[i];
        xo[i] = z;
        acc += z * b[i];
    }
    double v[1] = {acc};
    blockPartial<1>(v, part, smem);
}
template <class T, bool kAhead = false>
__global__ void __launch_bounds__(kBlock) kResidual(CsrView<T> A, const T* b, const T* x, T* r) {
    const int q = globalThread();
    if (q >= A.n) return;
    const int i = A.perm ? A.perm[q] : q;
    r[i] = b[i] - rowDot<T, kAhead>(A, x, q, i);
}
template <class T> __global__ void __launch_bounds__(kBlock) kRestrict(int nc, const int* mp, const int* mi, const T* r, T* bc) {
    const int I = globalThread();
    if (I >= nc) return;
    bc[I] = memberSum(mp, mi, r, I);
}
template <class T> __global__ void __launch_bounds__(kBlock) kProlong(int n, const int* agg, const T* xc, T* x, double alpha) {
    const int i = globalThread();
    if (i < n) x[i] = x[i] + static_cast<T>(alpha) * xc[agg[i]];
}

// Fused V-cycle steps.  Each computes, element by element, exactly what the
// unfused kernel sequence computes, so a cycle gives the same bits either
// way; they only save launches on the levels where launches, not bytes,
// set the time.
//
// kRestrict, then the coarse level's first damped-Jacobi step (kJacobiFirst)
// into xc.
template <class T>
__global__ void __launch_bounds__(kBlock) kRestrictFirst(int nc, const int* mp, const int* mi, const T* r, T* bc, const T* dc, T* xc,
                               double omega) {
    const int I = globalThread();
    if (I >= nc) return;
    const T s = memberSum(mp, mi, r, I);
    bc[I] = s;
    xc[I] = static_cast<T>(omega) * s / dc[I];
}
// kProlong followed by one kJacobiSweep: the prolongated vector is formed
// where the sweep reads it and is not stored.
template <class T>
__global__ void __launch_bounds__(kBlock) kProlongSweep(CsrView<T> A, const int* agg, const T* xc, double alpha, const T* b, const T* x, T* xo,
                              double omega) {
    const int q = globalThread();
    if (q >= A.n) return;
    const int i = A.perm ? A.perm[q] : q;
    const ProlongedX<T> xf{x, agg, xc, static_cast<T>(alpha)};
    const T xi = xf(i);
    const T ax = rowDotG<T>(A, xf, q, i);
    xo[i] = xi + static_cast<T>(omega) * (b[i] - ax) / A.diag[i];
}

// The levels at the bottom of the hierarchy, run by one workgroup: a barrier
// between the steps replaces a kernel launch.  Level k of the tail has the
// matrix `A[k]`, the aggregation of its rows into level k+1 (none on the
// last level, which is solved by the dense inverse) and its x (result), b,
// r and t (second sweep vector) arrays.
constexpr int kTailThreads = 1024;
constexpr int kMaxTailLevels = 12;
// A level joins the single-workgroup tail when its rows and stored entries
// fit one pass of the workgroup's threads (one row each, about a cache
// line of entries): beyond that one compute unit's bandwidth is slower than
// the launches it saves.
constexpr int kTailRows = kTailThreads;
constexpr int kTailEntries = 8 * kTailThreads;
// Prolongation and the first post-sweep run as one kernel on coarse levels
// with thin rows (the fused sweep gathers agg and xc for every entry).
constexpr int kFuseMaxRowLength = 16;
// Launched coarse levels take the pipelined row product when their rows span
// more than one chunk and the level is small enough that memory latency, not
// bandwidth, sets a sweep's time; on larger levels the plain form's lower
// register count (more waves per SIMD) wins.  Inside kCycleTail the plain
// form measured faster on every level.
constexpr int kAheadMinRowLength = 9;
constexpr int kAheadMaxRows = 65536;
template <class T> struct TailLevels {
    int count;
    CsrView<T> A[kMaxTailLevels];
    const int* agg[kMaxTailLevels];
    const int* memPtr[kMaxTailLevels];
    const int* memIdx[kMaxTailLevels];
    T* x[kMaxTailLevels];
    T* b[kMaxTailLevels];
    T* r[kMaxTailLevels];
    T* t[kMaxTailLevels];
    const double* inv;  // dense inverse of the last level, column-major
    double omega, alpha;
    int pre, post;
};
template <class T> __global__ void __launch_bounds__(kTailThreads) kCycleTail(TailLevels<T> L) {
    const int tid = threadIdx.x;
    const int later = L.pre - 1 + L.post;
    // Each level starts in x when the (pre - 1) + post swaps are even, so
    // that its result lands in x; after the pre-smoothing swaps the current
    // iterate is in x exactly when curIsX holds.
    const bool startX = later % 2 == 0;
    const bool curIsX = (L.pre > 1 ? L.pre - 1 : 0) % 2 == 0 ? startX : !startX;
    const int last = L.count - 1;
    for (int k = 0; k < last; ++k) {  // descend: pre-smoothing, residual, restriction
        const CsrView<T>& A = L.A[k];
        const T* b = L.b[k];
        T* c = later % 2 == 0 ? L.x[k] : L.t[k];
        T* o = later % 2 == 0 ? L.t[k] : L.x[k];
        for (int i = tid; i < A.n; i += kTailThreads) c[i] = static_cast<T>(L.omega) * b[i] / A.diag[i];
        __syncthreads();
        for (int s = 1; s < L.pre; ++s) {
            for (int q = tid; q < A.n; q += kTailThreads) {
                const int i = A.perm ? A.perm[q] : q;
                const T ax = rowDot(A, c, q, i);
                o[i] = c[i] + static_cast<T>(L.omega) * (b[i] - ax) / A.diag[i];
            }
            __syncthreads();
            T* w = c; c = o; o = w;
        }
        for (int q = tid; q < A.n; q += kTailThreads) {
            const int i = A.perm ? A.perm[q] : q;
            L.r[k][i] = b[i] - rowDot(A, c, q, i);
        }
        __syncthreads();
        const int nc = L.A[k + 1].n;
        for (int I = tid; I < nc; I += kTailThreads) {
            L.b[k + 1][I] = memberSum(L.memPtr[k], L.memIdx[k], L.r[k], I);
        }
        __syncthreads();
    }
    {  // the last level: x = inv b (the kDenseApply sum)
        const int n = L.A[last].n;
        const T* b = L.b[last];
        for (int i = tid; i < n; i += kTailThreads) {
            double s = 0.0;
            int j = 0;
            for (; j + 8 <= n; j += 8) {
                double v[8], w[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    v[u] = L.inv[static_cast<long long>(j + u) * n + i];
                    w[u] = static_cast<double>(b[j + u]);
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) s += v[u] * w[u];
            }
            for (; j < n; ++j) s += L.inv[static_cast<long long>(j) * n + i] * static_cast<double>(b[j]);
            L.x[last][i] = static_cast<T>(s);
        }
        __syncthreads();
    }
    for (int k = last - 1; k >= 0; --k) {  // ascend: prolongation, post-smoothing
        const CsrView<T>& A = L.A[k];
        const T* b = L.b[k];
        T* c = curIsX ? L.x[k] : L.t[k];
        T* o = curIsX ? L.t[k] : L.x[k];
        const T* xc = L.x[k + 1];
        for (int i = tid; i < A.n; i += kTailThreads) c[i] = c[i] + static_cast<T>(L.alpha) * xc[L.agg[k][i]];
        __syncthreads();
        for (int s = 0; s < L.post; ++s) {
            for (int q = tid; q < A.n; q += kTailThreads) {
                const int i = A.perm ? A.perm[q] : q;
                const T ax = rowDot(A, c, q, i);
                o[i] = c[i] + static_cast<T>(L.omega) * (b[i] - ax) / A.diag[i];
            }
            __syncthreads();
            T* w = c; c = o; o = w;
        }
    }
}
__global__ void __launch_bounds__(kBlock) kGalerkin(int n, int nc, int ne, const int* gp, const int* gs, const double* fd, const double* fv,
                          double* cd, double* cv) {
    const int e = globalThread();
    if (e >= ne) return;
    double s = 0.0;
    for (int t = gp[e]; t < gp[e + 1]; ++t) {
        const int src = gs[t];
        s += src < n ? fd[src] : fv[src - n];
    }
    if (e < nc) cd[e] = s;
    else cv[e - nc] = s;
}
__global__ void __launch_bounds__(kBlock) kToFloat(int n, const double* a, float* b) {
    const int i = globalThread();
    if (i < n) b[i] = static_cast<float>(a[i]);
}
__global__ void __launch_bounds__(kBlock) kToDouble(int n, const float* a, double* b) {
    const int i = globalThread();
    if (i < n) b[i] = static_cast<double>(a[i]);
}
// Dense fill + LU with partial pivoting of the coarsest matrix (one block).
// The coarsest level can be stored as CSR or sliced (a one-level hierarchy is
// the fine matrix itself); rows are filled in entry order either way.
__global__ void __launch_bounds__(1024) kDenseLU(CsrView<double> A, double* a, int* piv, int* flag) {
    const int n = A.n;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) a[e] = 0.0;
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int i = A.perm ? A.perm[q] : q;
        double* row = a + static_cast<long long>(i) * n;
        row[i] = A.diag[i];
        if (A.sliceOff) {
            const RowRange R(A.sliceOff, A.K, q);
            for (int k = 0; R.valid(k); ++k) {
                const int e = R.at(k), c = A.col[e];
                if (c >= 0) row[c] += A.val[e];
            }
        } else {
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) row[A.col[k]] += A.val[k];
        }
    }
    __syncthreads();
    __shared__ int sp;
    for (int k = 0; k < n; ++k) {
        if (threadIdx.x == 0) {
            int p = k;
            for (int i = k + 1; i < n; ++i)
                if (fabs(a[static_cast<long long>(i) * n + k]) > fabs(a[static_cast<long long>(p) * n + k])) p = i;
            piv[k] = p;
            sp = p;
            if (!(fabs(a[static_cast<long long>(p) * n + k]) > 0.0)) *flag = 1;
        }
        __syncthreads();
        const int p = sp;
        if (p != k)
            for (int j = threadIdx.x; j < n; j += blockDim.x) {
                const double t = a[static_cast<long long>(k) * n + j];
                a[static_cast<long long>(k) * n + j] = a[static_cast<long long>(p) * n + j];
                a[static_cast<long long>(p) * n + j] = t;
            }
        __syncthreads();
        const double pv = a[static_cast<long long>(k) * n + k];
        for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) {
            const double l = a[static_cast<long long>(i) * n + k] / pv;
            a[static_cast<long long>(i) * n + k] = l;
            for (int j = k + 1; j < n; ++j) a[static_cast<long long>(i) * n + j] -= l * a[static_cast<long long>(k) * n + j];
        }
        __syncthreads();
    }
}
// Explicit inverse of the coarsest matrix from its LU factors: column j is
// the LU solve of the unit vector e_j (thread j works in column j of inv).
// The inverse is stored column-major (thread j owns column j), so that the
// dense apply below reads it with consecutive lanes on consecutive rows.
__global__ void __launch_bounds__(kBlock) kDenseInverse(int n, const double* a, const int* piv, double* inv) {
    const int j = globalThread();
    if (j >= n) return;
    auto w = [&](int i) -> double& { return inv[static_cast<long long>(j) * n + i]; };
    for (int i = 0; i < n; ++i) w(i) = i == j ? 1.0 : 0.0;
    for (int k = 0; k < n; ++k) { const double t = w(k); w(k) = w(piv[k]); w(piv[k]) = t; }
    for (int i = 0; i < n; ++i)
        for (int q = 0; q < i; ++q) w(i) -= a[static_cast<long long>(i) * n + q] * w(q);
    for (int i = n - 1; i >= 0; --i) {
        for (int q = i + 1; q < n; ++q) w(i) -= a[static_cast<long long>(i) * n + q] * w(q);
        w(i) /= a[static_cast<long long>(i) * n + i];
    }
}
// Coarsest solve as a dense product x = inv * b (FP64 accumulation).
template <class T> __global__ void __launch_bounds__(kBlock) kDenseApply(int n, const double* inv, const T* b, T* xo) {
    const int i = globalThread();
    if (i >= n) return;
    double s = 0.0;
    int j = 0;
    for (; j + 8 <= n; j += 8) {  // eight independent loads in flight, summed in order
        double v[8], w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v[u] = inv[static_cast<long long>(j + u) * n + i];
            w[u] = static_cast<double>(b[j + u]);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u] * w[u];
    }
    for (; j < n; ++j) s += inv[static_cast<long long>(j) * n + i] * static_cast<double>(b[j]);
    xo[i] = static_cast<T>(s);
}

}  // namespace

bool SolverCfg::operator==(const SolverCfg& o) const {
    return solver == o.solver && pc == o.pc && atol == o.atol && rtol == o.rtol && maxIter == o.maxIter &&
           warmStart == o.warmStart && rebuildInterval == o.rebuildInterval && reuseGrowth == o.reuseGrowth &&
           amgMaxLevels == o.amgMaxLevels && amgCoarseSize == o.amgCoarseSize && amgPre == o.amgPre &&
           amgPost == o.amgPost && amgOmega == o.amgOmega && amgStrength == o.amgStrength &&
           amgCorrection == o.amgCorrection && precision == o.precision && innerRtol == o.innerRtol;
}

// ------------------------------------------------------------ probe
KernelProbe::~KernelProbe() {
    std::lock_guard<std::recursive_mutex> rl(pool::runtimeLock());
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
}
void KernelProbe::start(int capacity) {
    if (capacity <= 0) throw std::invalid_argument("probe capacity must be positive");
    while (static_cast<int>(ev.size()) < 2 * capacity) {
        // Timing-only events: no system-scope fence (cache writeback and
        // invalidate) when they complete, so they do not perturb the kernels.
        hipEvent_t e;
        std::lock_guard<std::recursive_mutex> rl(pool::runtimeLock());
        BSIM_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        ev.push_back(e);
    }
    used = 0;
    on = true;
}
void KernelProbe::begin(hipStream_t s) {
    pending = 0;
    if (!on || 2 * (used + 1) > static_cast<int>(ev.size())) return;
    BSIM_HIP(hipEventRecord(ev[2 * used], s));
    pending = 1;
}
void KernelProbe::end(hipStream_t s) {
    if (!pending) return;
    BSIM_HIP(hipEventRecord(ev[2 * used + 1], s));
    ++used;
    pending = 0;
}
double KernelProbe::stop(int& launches) {
    on = false;
    BSIM_HIP(hipDeviceSynchronize());
    double total = 0.0;
    for (int k = 0; k < used; ++k) {
        BSIM_HIP(hipEventSynchronize(ev[2 * k + 1]));
        float ms = 0.0f;
        BSIM_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
        total += 1e-3 * ms;
    }
    launches = used;
    used = 0;
    return total;
}

// ------------------------------------------------------------------ AMG
void AmgDevice::setup(const CsrView<double>& A, const Pattern& P, const SolverCfg& c) {
    const auto t0 = Clock::now();
    dropGraphs();  // the cycle's settings and arrays may change below
    cfg = c;
    mixed = c.precision == 1;
    hipStream_t s = stream;
    if (!aggregated) {
        // Aggregation runs once, on the row-major CSR of the first matrix.
        HostCsr A0;
        A0.n = A.n;
        A0.ptr = P.hptr;
        A0.col = P.hcol;
        A0.diag.resize(A.n);
        std::vector<double> stored(std::max(P.nnzStore, 1));
        BSIM_HIP(hipMemcpyAsync(A0.diag.data(), A.diag, sizeof(double) * A.n, hipMemcpyDeviceToHost, s));
        if (P.nnzStore) BSIM_HIP(hipMemcpyAsync(stored.data(), A.val, sizeof(double) * P.nnzStore, hipMemcpyDeviceToHost, s));
        BSIM_HIP(hipStreamSynchronize(s));
        A0.val.resize(P.hcol.size());
        for (std::size_t q = 0; q < P.hcol.size(); ++q) A0.val[q] = stored[P.csrToStore.empty() ? q : P.csrToStore[q]];
        std::vector<HostLevel> H = buildHierarchy(A0, c.amgStrength, c.amgMaxLevels, c.amgCoarseSize);
        // Level-0 Galerkin sources address the stored (possibly sliced) values.
        if (!P.csrToStore.empty() && !H[0].gSrc.empty()) remapGatherSrc(H[0].gSrc, A.n, P.csrToStore);
        // Smoothed coarse levels are stored sliced (the coarsest stays CSR for
        // the dense factorisation); the gathers that write and read their
        // values are rewritten to the sliced positions.
        std::vector<SlicedPattern> sliced(H.size());
        for (std::size_t l = 1; l + 1 < H.size(); ++l) {
            sliced[l] = sliceRows(H[l].A);
            remapGatherDest(H[l - 1].gPtr, H[l - 1].gSrc, H[l].A.n, sliced[l].csrToStore);
            remapGatherSrc(H[l].gSrc, H[l].A.n, sliced[l].csrToStore);
        }
        L.clear();
        for (std::size_t l = 0; l < H.size(); ++l) {
            auto lv = std::make_unique<Level>();
            lv->n = H[l].A.n;
            lv->nnz = static_cast<int>(H[l].A.col.size());
            if (l == 0) { lv->ptr = A.ptr; lv->col = A.col; lv->sliceOff = A.sliceOff; lv->K = A.K; lv->nnz = P.nnzStore; }
            else if (l + 1 < H.size()) {
                const SlicedPattern& S = sliced[l];
                lv->nnz = S.nnz;
                lv->K = S.K;
                lv->colOwn.upload(S.col, s);
                lv->sliceOwn.upload(S.sliceOff, s);
                lv->permOwn.upload(S.perm, s);
                lv->col = lv->colOwn.p;
                lv->sliceOff = lv->sliceOwn.p;
                lv->perm = lv->permOwn.p;
            } else {
                lv->ptrOwn.upload(H[l].A.ptr, s);
                lv->colOwn.upload(H[l].A.col, s);
                lv->ptr = lv->ptrOwn.p;
                lv->col = lv->colOwn.p;
            }
            lv->nc = H[l].nc;
            if (!H[l].agg.empty()) {
                lv->agg.upload(H[l].agg, s);
                lv->memPtr.upload(H[l].memPtr, s);
                lv->memIdx.upload(H[l].memIdx, s);
                lv->gPtr.upload(H[l].gPtr, s);
                lv->gSrc.upload(H[l].gSrc, s);
            }
            lv->diag.alloc(lv->n);
            lv->val.alloc(std::max(lv->nnz, 1));
            lv->x.alloc(lv->n); lv->b.alloc(lv->n); lv->r.alloc(lv->n); lv->t.alloc(lv->n);
            L.push_back(std::move(lv));
        }
        hostAgg0 = H[0].agg;
        // The tail: the coarsest level and the levels above it that are within
        // kTailRows rows and kTailEntries stored entries (never level 0,
        // whose cycle also forms z . r).
        const int nl = static_cast<int>(L.size());
        tailStart = nl - 1;
        while (tailStart > 1 && L[tailStart - 1]->n <= kTailRows && L[tailStart - 1]->nnz <= kTailEntries) --tailStart;
        if (nl - tailStart > kMaxTailLevels) tailStart = nl - kMaxTailLevels;
        if (tailStart < 1) tailStart = nl - 1;
        const int nco = L.back()->n;
        lu.alloc(static_cast<std::size_t>(nco) * nco);
        piv.alloc(nco);
        luFlag.alloc(1);
        aggregated = true;
    }
    if (mixed) {
        for (auto& lv : L) {
            lv->diag32.ensure(lv->n); lv->val32.ensure(std::max(lv->nnz, 1));
            lv->x32.ensure(lv->n); lv->b32.ensure(lv->n); lv->r32.ensure(lv->n); lv->t32.ensure(lv->n);
        }
    }
    // Numeric setup: the preconditioner keeps its own copy of the fine matrix.
    BSIM_HIP(hipMemcpyAsync(L[0]->diag.p, A.diag, sizeof(double) * A.n, hipMemcpyDeviceToDevice, s));
    if (L[0]->nnz) BSIM_HIP(hipMemcpyAsync(L[0]->val.p, A.val, sizeof(double) * L[0]->nnz, hipMemcpyDeviceToDevice, s));
    for (std::size_t l = 0; l + 1 < L.size(); ++l) {
        Level& f = *L[l];
        Level& cc = *L[l + 1];
        const int ne = cc.n + cc.nnz;
        kGalerkin<<<blocksFor(ne), kBlock, 0, s>>>(f.n, cc.n, ne, f.gPtr.p, f.gSrc.p, f.diag.p, f.val.p, cc.diag.p, cc.val.p);
    }
    if (mixed)
        for (auto& lv : L) {
            kToFloat<<<blocksFor(lv->n), kBlock, 0, s>>>(lv->n, lv->diag.p, lv->diag32.p);
            if (lv->nnz) kToFloat<<<blocksFor(lv->nnz), kBlock, 0, s>>>(lv->nnz, lv->val.p, lv->val32.p);
        }
    Level& co = *L.back();
    luFlag.zero(s);
    if (!co.ptr && !co.sliceOff) throw std::logic_error("AMG coarsest level has no row layout");
    const CsrView<double> Aco{co.n, co.ptr, co.col, co.diag.p, co.val.p, co.sliceOff, co.K, co.perm};
    kDenseLU<<<1, 1024, 0, s>>>(Aco, lu.p, piv.p, luFlag.p);
    BSIM_LAUNCH_CHECK();
    if (luFlag.toHost(s)[0]) throw std::runtime_error("AMG coarsest matrix is singular");
    inv.ensure(static_cast<std::size_t>(co.n) * co.n);
    kDenseInverse<<<blocksFor(co.n), kBlock, 0, s>>>(co.n, lu.p, piv.p, inv.p);
    BSIM_LAUNCH_CHECK();
    setupSeconds += since(t0);
}

std::vector<int> AmgDevice::sizes() const {
    std::vector<int> v;
    for (auto& lv : L) v.push_back(lv->n);
    return v;
}

std::vector<double> AmgDevice::coarseDiag(int level) const { return L.at(level)->diag.toHost(stream); }

template <class T>
void AmgDevice::levelArrays(const Level& lv, const T*& dg, const T*& vl, T*& x, T*& b, T*& r, T*& t) const {
    if constexpr (std::is_same<T, float>::value) {
        dg = lv.diag32.p; vl = lv.val32.p; x = lv.x32.p; b = lv.b32.p; r = lv.r32.p; t = lv.t32.p;
    } else {
        dg = lv.diag.p; vl = lv.val.p; x = lv.x.p; b = lv.b.p; r = lv.r.p; t = lv.t.p;
    }
}

// The levels from tailStart to the coarsest in one launch of kCycleTail: the
// right-hand side is in level tailStart's b array and the result goes to its
// x array, as cycle(tailStart, b, x) would leave them.
template <class T> void AmgDevice::tail() {
    TailLevels<T> a{};
    a.count = static_cast<int>(L.size()) - tailStart;
    for (int k = 0; k < a.count; ++k) {
        const Level& lv = *L[tailStart + k];
        const T *dg, *vl;
        levelArrays<T>(lv, dg, vl, a.x[k], a.b[k], a.r[k], a.t[k]);
        a.A[k] = CsrView<T>{lv.n, lv.ptr, lv.col, dg, vl, lv.sliceOff, lv.K, lv.perm};
        a.agg[k] = lv.agg.p;
        a.memPtr[k] = lv.memPtr.p;
        a.memIdx[k] = lv.memIdx.p;
    }
    a.inv = inv.p;
    a.omega = cfg.amgOmega;
    a.alpha = cfg.amgCorrection;
    a.pre = cfg.amgPre;
    a.post = cfg.amgPost;
    kCycleTail<T><<<1, kTailThreads, 0, stream>>>(a);
}

// One V-cycle on level l: b is the right-hand side, and the result is left
// in xOut.  Sweeps alternate between xOut and the level's spare vector; the
// first write goes to whichever of the two makes the last one land in xOut.
// firstDone: the first damped-Jacobi step is already in that vector (the
// parent's restriction wrote it).  On coarse levels with thin rows the
// prolongation and the first post-sweep run as one kernel.
template <class T> void AmgDevice::cycle(int l, const T* b, T* xOut, bool firstDone) {
    Level& lv = *L[l];
    hipStream_t s = stream;
    const T *dg, *vl;
    T *xl, *bl, *r, *spare;
    levelArrays<T>(lv, dg, vl, xl, bl, r, spare);
    const CsrView<T> A{lv.n, lv.ptr, lv.col, dg, vl, lv.sliceOff, lv.K, lv.perm};
    const int nb = blocksFor(lv.n);
    if (l == static_cast<int>(L.size()) - 1) {
        kDenseApply<T><<<blocksFor(lv.n), kBlock, 0, s>>>(lv.n, inv.p, b, xOut);
        return;
    }
    const double w = cfg.amgOmega;
    const int pre = cfg.amgPre, post = cfg.amgPost;
    const int later = pre - 1 + post;
    T* cur = later % 2 == 0 ? xOut : spare;
    T* nxt = later % 2 == 0 ? spare : xOut;
    KernelProbe* pr = l == 0 ? probe : nullptr;
    const bool ahead = l > 0 && lv.sliceOff && lv.K >= kAheadMinRowLength && lv.n <= kAheadMaxRows;
    auto sweep = [&]() {  // cur -> nxt (the probed kernel)
        if (pr) pr->begin(s);
        if (ahead) kJacobiSweep<T, true><<<nb, kBlock, 0, s>>>(A, b, cur, nxt, w);
        else kJacobiSweep<T><<<nb, kBlock, 0, s>>>(A, b, cur, nxt, w);
        if (pr) pr->end(s);
        std::swap(cur, nxt);
    };
    if (!firstDone) kJacobiFirst<T><<<nb, kBlock, 0, s>>>(lv.n, dg, b, cur, w);
    for (int k = 1; k < pre; ++k) sweep();
    Level& cl = *L[l + 1];
    const T *cdg, *cvl;
    T *cx, *cb, *cr, *ct;
    levelArrays<T>(cl, cdg, cvl, cx, cb, cr, ct);
    const bool childTail = l + 1 >= tailStart;
    T* childCur = childTail ? nullptr : (later % 2 == 0 ? cx : ct);
    const bool fuse = l >= 1 && lv.sliceOff && lv.K <= kFuseMaxRowLength;
    if (ahead) kResidual<T, true><<<nb, kBlock, 0, s>>>(A, b, cur, r);
    else kResidual<T><<<nb, kBlock, 0, s>>>(A, b, cur, r);
    if (childCur)
        kRestrictFirst<T><<<blocksFor(cl.n), kBlock, 0, s>>>(cl.n, lv.memPtr.p, lv.memIdx.p, r, cb, cdg, childCur, w);
    else
        kRestrict<T><<<blocksFor(cl.n), kBlock, 0, s>>>(cl.n, lv.memPtr.p, lv.memIdx.p, r, cb);
    if (childTail) tail<T>();
    else cycle<T>(l + 1, cb, cx, true);
    int k0 = 0;
    if (fuse && post > 0) {
        kProlongSweep<T><<<nb, kBlock, 0, s>>>(A, lv.agg.p, cx, cfg.amgCorrection, b, cur, nxt, w);
        std::swap(cur, nxt);
        k0 = 1;
    } else {
        kProlong<T><<<nb, kBlock, 0, s>>>(lv.n, lv.agg.p, cx, cur, cfg.amgCorrection);
    }
    for (int k = k0; k < post; ++k) {
        if constexpr (std::is_same<T, double>::value) {
            if (l == 0 && k == post - 1 && dotPartials) {  // last sweep also forms z . r
                kJacobiSweepDot<<<reduceBlocks(lv.n), kBlock, 0, s>>>(A, b, cur, nxt, w, dotPartials);
                std::swap(cur, nxt);
                continue;
            }
        }
        sweep();
    }
    if (cur != xOut) throw std::logic_error("V-cycle result left outside its output vector");
}

void AmgDevice::run(const double* r, double* z, double* zrPartials) {
    Level& l0 = *L[0];
    hipStream_t s = stream;
    const int nb = blocksFor(l0.n);
    if (mixed) {
        kToFloat<<<nb, kBlock, 0, s>>>(l0.n, r, l0.b32.p);
        cycle<float>(0, l0.b32.p, l0.x32.p, false);
        kToDouble<<<nb, kBlock, 0, s>>>(l0.n, l0.x32.p, z);
    } else {
        dotPartials = zrPartials;
        cycle<double>(0, r, z, false);
        dotPartials = nullptr;
    }
    BSIM_LAUNCH_CHECK();
}

namespace {
std::atomic<bool> g_graphReplay{true};
}
void setGraphReplay(bool on) { g_graphReplay = on; }
bool graphReplay() { return g_graphReplay; }

bool AmgDevice::apply(const double* r, double* z, double* zrPartials) {
    const bool fused = !mixed && zrPartials && L.size() > 1 && cfg.amgPost > 0;
    double* partials = fused ? zrPartials : nullptr;
    // Events bracket single kernels only in launches from the stream.
    if ((probe && probe->on) || !graphReplay()) {
        run(r, z, partials);
        return fused;
    }
    const GraphKey key{r, z, partials};
    auto it = graphs.find(key);
    if (it == graphs.end()) {
        // The cycle is launches only (no copies, fills or host waits), all
        // with arguments fixed by the hierarchy and the key.  No other
        // thread returns memory to the runtime (a device-wide wait) while
        // the capture is open.
        std::lock_guard<std::recursive_mutex> rl(pool::runtimeLock());
        BSIM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        try {
            run(r, z, partials);
        } catch (...) {
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        BSIM_HIP(hipStreamEndCapture(stream, &g));
        hipGraphExec_t ge = nullptr;
        const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e != hipSuccess) (void)hipGraphDestroy(g);
        BSIM_HIP(e);
        it = graphs.emplace(key, Captured{g, ge}).first;  // the graph lives as long as its executable
    }
    BSIM_HIP(hipGraphLaunch(it->second.exec, stream));
    return fused;
}

void AmgDevice::dropGraphs() {
    std::lock_guard<std::recursive_mutex> rl(pool::runtimeLock());
    for (auto& kv : graphs) {
        (void)hipGraphExecDestroy(kv.second.exec);
        (void)hipGraphDestroy(kv.second.graph);
    }
    graphs.clear();
}

AmgDevice::~AmgDevice() { dropGraphs(); }

// ------------------------------------------------------------- systems
DeviceSystem::DeviceSystem(hipStream_t st, Pattern pattern)
    : n(pattern.n), nnz(pattern.nnzStore), stream(st), pat(std::move(pattern)) {
    const auto t0 = Clock::now();
    for (int k = 0; k < 2; ++k) { diagBuf[k].alloc(n); valBuf[k].alloc(std::max(nnz, 1)); }
    sc.alloc(S_COUNT);
    flags.alloc(1);
    partials.alloc(6 * kMaxReduceBlocks);
    hostScalars = static_cast<double*>(pool::hostAlloc(sizeof(double) * (S_COUNT + 2)));
    for (DBuf<double>* b : {&r, &z, &p, &w, &rp, &v, &t, &s, &tmp, &guess, &saved}) b->alloc(n);
    sc3.alloc(3 * S_COUNT);
    fl3.alloc(3);
    act3.alloc(3);
    hsc3 = static_cast<double*>(pool::hostAlloc(sizeof(double) * (3 * S_COUNT + 2)));
    hact3 = static_cast<int*>(pool::hostAlloc(sizeof(int) * 4 * kActSlots));
    counters.patternSeconds = since(t0);
}

namespace {
constexpr std::size_t kWsAlign = 32;  // doubles: 256 bytes
std::size_t wsRound(std::size_t v) { return (v + kWsAlign - 1) / kWsAlign * kWsAlign; }
constexpr int kWsVectors = 12;        // r z p w rp v t s tmp guess saved jinv
}  // namespace

std::size_t DeviceSystem::workspaceDoubles(int n) {
    return 2 * kWsAlign + 6 * static_cast<std::size_t>(kMaxReduceBlocks) +
           kWsVectors * wsRound(static_cast<std::size_t>(n));
}

DeviceSystem::DeviceSystem(hipStream_t st, Pattern pattern, const SystemWorkspace& ws)
    : n(pattern.n), nnz(pattern.nnzStore), stream(st), pat(std::move(pattern)) {
    static_assert(S_COUNT + 2 <= 32, "host scalar mirror");
    if (!ws.diag || (!ws.val && nnz > 0) || !ws.scratch) throw std::invalid_argument("null workspace");
    if (ws.doubles < workspaceDoubles(n)) throw std::invalid_argument("workspace is too small for the system");
    borrowed = true;
    // Both matrix slots view the caller's values: the first update() makes
    // them effective without a comparison, and nothing here writes them.
    for (int k = 0; k < 2; ++k) {
        diagBuf[k].borrow(const_cast<double*>(ws.diag), n);
        valBuf[k].borrow(const_cast<double*>(ws.val), nnz);
    }
    double* q = ws.scratch;
    sc.borrow(q, S_COUNT);
    q += kWsAlign;
    flags.borrow(reinterpret_cast<int*>(q), 1);
    q += kWsAlign;
    partials.borrow(q, 6 * static_cast<std::size_t>(kMaxReduceBlocks));
    q += 6 * static_cast<std::size_t>(kMaxReduceBlocks);
    const std::size_t step = wsRound(static_cast<std::size_t>(n));
    for (DBuf<double>* b : {&r, &z, &p, &w, &rp, &v, &t, &s, &tmp, &guess, &saved, &jinv}) {
        b->borrow(q, static_cast<std::size_t>(n));
        q += step;
    }
    hostScalars = hostLocal;
}

DeviceSystem::~DeviceSystem() {
    if (borrowed) return;
    if (hostScalars) pool::hostFree(hostScalars);
    if (hsc3) pool::hostFree(hsc3);
    if (hact3) pool::hostFree(hact3);
    if (gmHost) pool::hostFree(gmHost);
}

bool DeviceSystem::update() {
    const auto t0 = Clock::now();
    bool changed = true;
    if (ready) {
        const int rb = reduceBlocks(static_cast<long long>(n) + nnz);
        kCompare<<<rb, kBlock, 0, stream>>>(n, diagBuf[0].p, diagBuf[1].p, nnz, valBuf[0].p, valBuf[1].p, partials.p);
        kReduceTail<1><<<1, kTail, 0, stream>>>(rb, partials.p, StoreValue{sc.p, S_TMP});
        BSIM_LAUNCH_CHECK();
        changed = readScalar(S_TMP) != 0.0;
    }
    if (changed) {
        cand = 1 - cand;  // the candidate becomes the effective matrix
        ready = true;
        setupPending = true;
        ++counters.matrixUpdates;
    } else {
        ++counters.rhsOnly;
    }
    counters.matrixUpdateSeconds += since(t0);
    return changed;
}

void DeviceSystem::prepare(const SolverCfg& c) {
    if (borrowed && (c.pc >= 2 || c.precision != 0 || c.solver < 0 || c.solver > 1))
        throw std::invalid_argument("a system in a fixed workspace runs FP64 CG or BiCGSTAB with no "
                                    "preconditioner or Jacobi only");
    if (!configured || !(c == cfg)) {
        cfg = c;
        configured = true;
        setupPending = true;
        forceRebuild = true;
    }
    if (!setupPending && !forceRebuild) return;
    const bool reuse = cfg.pc != 0 && counters.pcSetups != 0 && !forceRebuild &&
                       counters.matrixUpdates - pcMatrixUpdate < static_cast<std::uint64_t>(cfg.rebuildInterval) &&
                       pcRhsSolves < static_cast<std::uint64_t>(cfg.rebuildInterval);
    setup(reuse);
}

void DeviceSystem::setup(bool reuse) {
    const auto t0 = Clock::now();
    if (!reuse) {
        if (cfg.pc == 1) {
            jinv.ensure(n);
            kInvDiag<<<blocksFor(n), kBlock, 0, stream>>>(n, diag(), jinv.p);
            BSIM_LAUNCH_CHECK();
        } else if (cfg.pc == 3) {
            iluDiag.ensure(n);
            iluVal.ensure(std::max(nnz, 1));
            kBlockIluFactor<<<(n + kIluBlock - 1) / kIluBlock, kIluBlock, 0, stream>>>(view(), iluDiag.p, iluVal.p);
            BSIM_LAUNCH_CHECK();
        } else if (cfg.pc == 4 || cfg.pc == 5) {
            // The symbolic analysis depends on the pattern and the kind only.
            const bool chol = cfg.pc == 4;
            if (!lilu || lilu->cholesky() != chol) {
                lilu = std::make_unique<LevelIlu>(stream, chol);
                lilu->analyse(pat, n);
            }
            lilu->factor(view());
        } else if (cfg.pc == 2) {
            if (!amg) amg = std::make_unique<AmgDevice>(stream);
            amg->probe = &probe;
            amg->setup(view(), pat, cfg);
            // Fine-level sweep: off-diagonal values and columns, then the
            // diagonal, b, x (each read once) and the result.
            const double vb = cfg.precision == 1 ? 4.0 : 8.0;
            probe.bytesPerLaunch = (vb + 4.0) * nnz + 4.0 * vb * n;
        }
        if (cfg.pc != 0) ++counters.pcSetups;
        pcMatrixUpdate = counters.matrixUpdates;
        pcRhsSolves = 0;
        freshIterations[0] = freshIterations[1] = freshIterations[2] = 0;
    }
    setupPending = false;
    forceRebuild = false;
    counters.setupSeconds += since(t0);
}

void DeviceSystem::applyPC(const double* rr, double* zz) {
    ++counters.pcApplications;
    if (cfg.pc == 0) kCopy<<<blocksFor(n), kBlock, 0, stream>>>(n, rr, zz);
    else if (cfg.pc == 1) kJacobiApply<<<blocksFor(n), kBlock, 0, stream>>>(n, rr, jinv.p, zz);
    else if (cfg.pc == 3)
        kBlockIluApply<<<(n + kIluBlock - 1) / kIluBlock, kIluBlock, 0, stream>>>(view(), iluDiag.p, iluVal.p, rr, zz);
    else if (cfg.pc >= 4) lilu->apply(rr, zz);
    else amg->apply(rr, zz);
    BSIM_LAUNCH_CHECK();
}

double DeviceSystem::readScalar(int idx, int* fl) {
    BSIM_HIP(hipMemcpyAsync(hostScalars, sc.p + idx, sizeof(double), hipMemcpyDeviceToHost, stream));
    if (fl) BSIM_HIP(hipMemcpyAsync(reinterpret_cast<int*>(hostScalars + S_COUNT), flags.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    BSIM_HIP(hipStreamSynchronize(stream));
    if (fl) *fl = *reinterpret_cast<int*>(hostScalars + S_COUNT);
    return hostScalars[0];
}

void DeviceSystem::spmv(const double* x, double* y) {
    kSpmvPlain<<<blocksFor(n), kBlock, 0, stream>>>(view(), x, y);
    BSIM_LAUNCH_CHECK();
}

namespace {
// HIP events destroyed on every path out of a timing routine.
struct EventSet {
    std::vector<hipEvent_t> ev;
    explicit EventSet(int count) {
        ev.reserve(count);
        for (int k = 0; k < count; ++k) {
            hipEvent_t e;
            std::lock_guard<std::recursive_mutex> rl(pool::runtimeLock());
            BSIM_HIP(hipEventCreate(&e));
            ev.push_back(e);
        }
    }
    ~EventSet() {
        std::lock_guard<std::recursive_mutex> rl(pool::runtimeLock());
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    EventSet(const EventSet&) = delete;
    EventSet& operator=(const EventSet&) = delete;
};
}  // namespace

// A diagnostic: the cold mode allocates a 512 MiB buffer outside any serving
// reservation, so it is not for contexts under admission control.
double DeviceSystem::timeSpmv(int reps, bool cold) {
    if (reps <= 0) throw std::invalid_argument("reps must be positive");
    if (borrowed) throw std::logic_error("SpMV timing allocates; not available in a fixed workspace");
    kFill<<<blocksFor(n), kBlock, 0, stream>>>(n, p.p, 1.0);
    kSpmvPlain<<<blocksFor(n), kBlock, 0, stream>>>(view(), p.p, w.p);  // warm
    if (!cold) {
        EventSet E(2);
        BSIM_HIP(hipEventRecord(E.ev[0], stream));
        for (int k = 0; k < reps; ++k) kSpmvPlain<<<blocksFor(n), kBlock, 0, stream>>>(view(), p.p, w.p);
        BSIM_HIP(hipEventRecord(E.ev[1], stream));
        BSIM_HIP(hipEventSynchronize(E.ev[1]));
        float ms = 0.0f;
        BSIM_HIP(hipEventElapsedTime(&ms, E.ev[0], E.ev[1]));
        return ms * 1e-3 / reps;
    }
    constexpr int kFlush = 1 << 26;  // 2^26 doubles = 512 MiB
    DBuf<double> flush;
    flush.alloc(kFlush);
    kFill<<<blocksFor(kFlush), kBlock, 0, stream>>>(kFlush, flush.p, 0.0);
    EventSet E(2 * reps);
    for (int k = 0; k < reps; ++k) {
        kTouch<<<4096, kBlock, 0, stream>>>(kFlush, flush.p, w.p);
        BSIM_HIP(hipEventRecord(E.ev[2 * k], stream));
        kSpmvPlain<<<blocksFor(n), kBlock, 0, stream>>>(view(), p.p, w.p);
        BSIM_HIP(hipEventRecord(E.ev[2 * k + 1], stream));
    }
    BSIM_HIP(hipEventSynchronize(E.ev.back()));
    double sec = 0.0;
    for (int k = 0; k < reps; ++k) {
        float ms = 0.0f;
        BSIM_HIP(hipEventElapsedTime(&ms, E.ev[2 * k], E.ev[2 * k + 1]));
        sec += ms * 1e-3;
    }
    // The buffer is freed only after the last launch that reads it.
    BSIM_HIP(hipStreamSynchronize(stream));
    flush.release();
    return sec / reps;
}

double DeviceSystem::trueResidual(const double* b, const double* x, bool check) {
    const auto t0 = Clock::now();
    const int rb = reduceBlocks(n);
    kResidualNorm<<<rb, kBlock, 0, stream>>>(view(), b, x, tmp.p, partials.p);
    kReduceTail<1><<<1, kTail, 0, stream>>>(rb, partials.p, StoreSqrt{sc.p, S_TR});
    BSIM_LAUNCH_CHECK();
    ++counters.spmvs;
    if (check) ++counters.trueResidualChecks;
    const double v = readScalar(S_TR);
    counters.residualSeconds += since(t0);
    return v;
}

void DeviceSystem::cg(const double* b, double* x, bool guessZero, int maxIt, double target, int& its, int& reason) {
    double* part = partials.p;
    const int rb = reduceBlocks(n), nb = blocksFor(n);
    flags.zero(stream);
    {
        const double init[S_COUNT] = {0, 1.0, 1.0, 0, 1.0, 0, 1.0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        BSIM_HIP(hipMemcpyAsync(sc.p, init, sizeof(init), hipMemcpyHostToDevice, stream));
    }
    if (!guessZero) { kResidualNorm<<<rb, kBlock, 0, stream>>>(view(), b, x, r.p, part); ++counters.spmvs; }
    else kCopyNorm<<<rb, kBlock, 0, stream>>>(n, b, r.p, part);
    kReduceTail<1><<<1, kTail, 0, stream>>>(rb, part, StoreSqrt{sc.p, S_DP});
    BSIM_LAUNCH_CHECK();
    double dp = readScalar(S_DP);
    its = 0;
    reason = 0;
    hist.push_back(dp);
    auto test = [&](double v) {
        if (!std::isfinite(v)) { reason = -9; return true; }
        if (v <= target) { reason = 2; return true; }
        return false;
    };
    if (test(dp)) return;
    auto pcDot = [&](int checkSign) {
        if (cfg.pc == 1) {
            kJacobiDot<<<rb, kBlock, 0, stream>>>(n, r.p, jinv.p, z.p, flags.p, part);
            ++counters.pcApplications;
        } else if (cfg.pc == 2) {
            ++counters.pcApplications;
            if (!amg->apply(r.p, z.p, part)) kPlainDot<<<rb, kBlock, 0, stream>>>(n, z.p, r.p, flags.p, part);
        } else {
            applyPC(r.p, z.p);
            kPlainDot<<<rb, kBlock, 0, stream>>>(n, z.p, r.p, flags.p, part);
        }
        kReduceTail<1><<<1, kTail, 0, stream>>>(rb, part, CgBeta{sc.p, flags.p, checkSign});
        BSIM_LAUNCH_CHECK();
    };
    pcDot(0);
    int i = 0;
    int fl = 0;
    do {
        its = i + 1;
        kCgDirection<<<nb, kBlock, 0, stream>>>(n, z.p, p.p, sc.p, flags.p, i == 0);
        kCgSpmvDot<<<rb, kBlock, 0, stream>>>(view(), p.p, w.p, flags.p, part);
        kReduceTail<1><<<1, kTail, 0, stream>>>(rb, part, CgAlpha{sc.p, flags.p, i});
        kCgUpdate<<<rb, kBlock, 0, stream>>>(n, x, p.p, r.p, w.p, sc.p, flags.p, part);
        kReduceTail<1><<<1, kTail, 0, stream>>>(rb, part, CgNorm{sc.p, flags.p});
        BSIM_LAUNCH_CHECK();
        ++counters.spmvs;
        dp = readScalar(S_DP, &fl);
        if (fl & F_BETA_ZERO) { reason = 3; break; }
        if (fl & F_INDEF_PC) { reason = -8; break; }
        if (fl & F_NONFINITE) {
            // A non-finite z . r ends the solve where it was formed, one
            // iteration earlier (KSPCheckDot after the preconditioner).
            reason = -9;
            if (!std::isfinite(readScalar(S_BETA))) its = i;
            break;
        }
        if (fl & F_INDEF_MAT) { reason = -10; break; }
        hist.push_back(dp);
        if (test(dp)) break;
        pcDot(1);
        ++i;
    } while (i < maxIt);
    if (i >= maxIt) reason = -3;
}

void DeviceSystem::bcgs(const double* b, double* x, bool guessZero, int maxIt, double target, int& its, int& reason) {
    double* part = partials.p;
    const int rb = reduceBlocks(n), nb = blocksFor(n);
    flags.zero(stream);
    {
        const double init[S_COUNT] = {0, 1.0, 1.0, 0, 1.0, 0, 1.0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        BSIM_HIP(hipMemcpyAsync(sc.p, init, sizeof(init), hipMemcpyHostToDevice, stream));
    }
    bool haveGuess = false;
    if (!guessZero) {
        kResidualNorm<<<rb, kBlock, 0, stream>>>(view(), b, x, r.p, part);
        ++counters.spmvs;
        kCopy<<<nb, kBlock, 0, stream>>>(n, x, guess.p);
        kFill<<<nb, kBlock, 0, stream>>>(n, x, 0.0);
        haveGuess = true;
    } else {
        kCopyNorm<<<rb, kBlock, 0, stream>>>(n, b, r.p, part);
    }
    kReduceTail<1><<<1, kTail, 0, stream>>>(rb, part, StoreSqrt{sc.p, S_DP});
    BSIM_LAUNCH_CHECK();
    double dp = readScalar(S_DP);
    its = 0;
    reason = 0;
    hist.push_back(dp);
    auto test = [&](double vv) {
        if (!std::isfinite(vv)) { reason = -9; return true; }
        if (vv <= target) { reason = 2; return true; }
        return false;
    };
    if (test(dp)) {
        if (haveGuess) kAxpy<<<nb, kBlock, 0, stream>>>(n, 1.0, guess.p, x);
        return;
    }
    kCopy<<<nb, kBlock, 0, stream>>>(n, r.p, rp.p);
    kFill<<<nb, kBlock, 0, stream>>>(n, p.p, 0.0);
    kFill<<<nb, kBlock, 0, stream>>>(n, v.p, 0.0);
    int i = 0, fl = 0;
    do {
        kBcgsRho<<<rb, kBlock, 0, stream>>>(n, r.p, rp.p, part);
        kReduceTail<1><<<1, kTail, 0, stream>>>(rb, part, BcgsRho{sc.p});
        kBcgsP<<<nb, kBlock, 0, stream>>>(n, p.p, r.p, v.p, sc.p);
        applyPC(p.p, tmp.p);
        kBcgsSpmvD1<<<rb, kBlock, 0, stream>>>(view(), tmp.p, v.p, rp.p, part);
        kReduceTail<1><<<1, kTail, 0, stream>>>(rb, part, BcgsAlpha{sc.p, flags.p});
        kBcgsS<<<nb, kBlock, 0, stream>>>(n, s.p, v.p, r.p, sc.p, flags.p);
        BSIM_LAUNCH_CHECK();
        applyPC(s.p, r.p);
        kBcgsSpmvD2<<<rb, kBlock, 0, stream>>>(view(), r.p, t.p, s.p, flags.p, part);
        kReduceTail<2><<<1, kTail, 0, stream>>>(rb, part, BcgsOmega{sc.p, flags.p});
        kBcgsUpdate<<<rb, kBlock, 0, stream>>>(n, x, p.p, s.p, r.p, t.p, sc.p, flags.p, part);
        kReduceTail<1><<<1, kTail, 0, stream>>>(rb, part, BcgsNorm{sc.p, flags.p});
        BSIM_LAUNCH_CHECK();
        counters.spmvs += 2;
        const double newDp = readScalar(S_DP, &fl);
        if (fl & F_NONFINITE) { reason = -9; break; }
        if (fl & F_D1_ZERO) { reason = -5; break; }
        if (fl & F_D2_ZERO) {
            kDot<<<rb, kBlock, 0, stream>>>(n, s.p, s.p, part);
            kReduceTail<1><<<1, kTail, 0, stream>>>(rb, part, StoreValue{sc.p, S_TMP});
            BSIM_LAUNCH_CHECK();
            if (readScalar(S_TMP) != 0.0) { reason = -5; break; }
            const double alpha = readScalar(S_ALPHA);
            kAxpy<<<nb, kBlock, 0, stream>>>(n, alpha, p.p, x);
            ++its;
            reason = 2;
            hist.push_back(dp);
            break;
        }
        dp = newDp;
        ++its;
        hist.push_back(dp);
        if (test(dp)) break;
        if (readScalar(S_RHO) == 0.0) { reason = -5; break; }
        ++i;
    } while (i < maxIt);
    if (i >= maxIt) reason = -3;
    applyPC(x, tmp.p);  // unwind right preconditioning
    kCopy<<<nb, kBlock, 0, stream>>>(n, tmp.p, x);
    if (haveGuess) kAxpy<<<nb, kBlock, 0, stream>>>(n, 1.0, guess.p, x);
    BSIM_LAUNCH_CHECK();
}

// KSPSolve_GMRES (gmres.c:88-235) and KSPSolve_FGMRES (fgmres/fgmres.c:56-244)
// as the reference configures them: the unpreconditioned norm
// (petsc_linear_system.cpp:181) puts GMRES on the right side, so both methods
// precondition from the right and test the residual norm of the Arnoldi
// recursion.  Restart kGmRestart, classical Gram-Schmidt without refinement
// (borthog2.c), happy-breakdown tolerance 1e-30 and, for GMRES, the breakdown
// test at a restart with tolerance 0.1.  The Hessenberg matrix, its plane
// rotations (KSPGMRESUpdateHessenberg) and the back substitution
// (KSPGMRESBuildSoln) are host scalars: each iteration reads back its new
// column of inner products and the norm of the new basis vector.
void DeviceSystem::gmres(const double* b, double* x, bool guessZero, int maxIt, double target, int& its, int& reason,
                         bool flexible) {
    if (borrowed) throw std::logic_error("GMRES allocates its Krylov basis; not available in a fixed workspace");
    constexpr int m = kGmRestart;
    constexpr double hapTol = 1e-30, breakdownTol = 0.1;
    const long long ld = (static_cast<long long>(n) + 31) / 32 * 32;
    gmV.ensure(static_cast<std::size_t>((m + 1) * ld));
    if (flexible) gmZ.ensure(static_cast<std::size_t>(m * ld));
    gmPart.ensure(static_cast<std::size_t>(kGmDots) * kMaxReduceBlocks);
    constexpr int ttSlot = m + kGmDots;  // past every chunk of inner products
    gmH.ensure(ttSlot + 1);
    if (!gmHost) gmHost = static_cast<double*>(pool::hostAlloc(sizeof(double) * (ttSlot + 1)));
    double* const V = gmV.p;
    auto Vj = [&](int j) { return V + j * ld; };
    auto Zj = [&](int j) { return gmZ.p + j * ld; };
    double* part = partials.p;
    const int rb = reduceBlocks(n), nb = blocksFor(n);
    // HH(i, j) at hh[i * (m + 1) + j]; rotations cc, ss; right-hand side grs.
    std::vector<double> hh(static_cast<std::size_t>(m + 1) * (m + 1), 0.0), cc(m + 1, 0.0), ss(m + 1, 0.0),
        grs(m + 1, 0.0), nrs(m + 1, 0.0);
    auto H = [&](int i, int j) -> double& { return hh[static_cast<std::size_t>(i) * (m + 1) + j]; };
    its = 0;
    reason = 0;
    auto test = [&](double v) {  // petsc_linear_system.cpp:164-175
        if (!std::isfinite(v)) { reason = -9; return true; }
        if (v <= target) { reason = 2; return true; }
        return false;
    };
    double rnorm = -1.0, rnorm0 = 0.0;  // ksp->rnorm (-1: first cycle) and gmres->rnorm0
    bool zero = guessZero;
    while (true) {
        // Initial residual (KSPInitialResidual / KSPFGMRESResidual) into V_0.
        if (zero) kCopyNorm<<<rb, kBlock, 0, stream>>>(n, b, Vj(0), part);
        else { kResidualNorm<<<rb, kBlock, 0, stream>>>(view(), b, x, Vj(0), part); ++counters.spmvs; }
        kReduceTail<1><<<1, kTail, 0, stream>>>(rb, part, StoreSqrt{sc.p, S_DP});
        BSIM_LAUNCH_CHECK();
        double res = readScalar(S_DP);
        zero = false;
        if (!std::isfinite(res)) { reason = -9; return; }  // KSPCheckNorm
        if (!flexible) {
            if (res != 0.0) kScale<<<nb, kBlock, 0, stream>>>(n, Vj(0), 1.0 / res);  // VecNormalize
            if (rnorm > 0.0 && std::fabs(res - rnorm) > breakdownTol * rnorm0) { reason = -5; return; }
        }
        grs[0] = rnorm0 = res;
        rnorm = res;
        hist.push_back(res);
        if (!flexible && res == 0.0) { reason = 3; return; }
        test(res);
        int it = 0;
        bool hapend = false;
        if (!reason && flexible) kScale<<<nb, kBlock, 0, stream>>>(n, Vj(0), 1.0 / res);
        bool built = true;  // false when a norm check returns from the cycle
        while (!reason && it < m && its < maxIt) {
            double* w = Vj(it + 1);
            if (flexible) {
                applyPC(Vj(it), Zj(it));
                kSpmvPlain<<<nb, kBlock, 0, stream>>>(view(), Zj(it), w);
            } else {
                applyPC(Vj(it), tmp.p);
                kSpmvPlain<<<nb, kBlock, 0, stream>>>(view(), tmp.p, w);
            }
            ++counters.spmvs;
            for (int j0 = 0; j0 <= it; j0 += kGmDots) {
                const int cnt = std::min(kGmDots, it + 1 - j0);
                kMultiDot<<<rb, kBlock, 0, stream>>>(n, w, V, ld, j0, cnt, gmPart.p);
                kReduceTail<kGmDots><<<1, kTail, 0, stream>>>(rb, gmPart.p, StoreSums<kGmDots>{gmH.p + j0});
            }
            kGmMaxpyNorm<<<rb, kBlock, 0, stream>>>(n, w, V, ld, it + 1, gmH.p, part);
            kReduceTail<1><<<1, kTail, 0, stream>>>(rb, part, StoreSqrt{gmH.p, ttSlot});
            BSIM_LAUNCH_CHECK();
            BSIM_HIP(hipMemcpyAsync(gmHost, gmH.p, sizeof(double) * (ttSlot + 1), hipMemcpyDeviceToHost, stream));
            BSIM_HIP(hipStreamSynchronize(stream));
            bool dotFailed = false;  // KSPCheckDot: the orthogonalisation stops
            for (int j = 0; j <= it; ++j)
                if (!std::isfinite(gmHost[j])) dotFailed = true;
            if (dotFailed) { reason = -9; break; }
            for (int j = 0; j <= it; ++j) H(j, it) = 0.0 - (-gmHost[j]);
            const double tt = gmHost[ttSlot];
            if (!std::isfinite(tt)) { reason = -9; built = false; break; }  // KSPCheckNorm returns from the cycle
            H(it + 1, it) = tt;
            const double hapbnd = std::min(std::fabs(tt / grs[it]), hapTol);
            if (flexible) {
                if (tt > hapbnd) kScale<<<nb, kBlock, 0, stream>>>(n, w, 1.0 / tt);
                else hapend = true;
            } else {
                if (tt != 0.0) kScale<<<nb, kBlock, 0, stream>>>(n, w, 1.0 / tt);
                if (tt < hapbnd) hapend = true;
            }
            // KSPGMRESUpdateHessenberg: earlier rotations, then a new one.
            bool nullDir = false;
            for (int j = 1; j <= it; ++j) {
                const double t0 = H(j - 1, it);
                H(j - 1, it) = cc[j - 1] * t0 + ss[j - 1] * H(j, it);
                H(j, it) = cc[j - 1] * H(j, it) - ss[j - 1] * t0;
            }
            if (!hapend) {
                const double a = H(it, it), c = H(it + 1, it);
                const double r = std::sqrt(a * a + c * c);
                if (r == 0.0) {
                    nullDir = true;
                } else {
                    cc[it] = a / r;
                    ss[it] = c / r;
                    grs[it + 1] = -(ss[it] * grs[it]);
                    grs[it] = cc[it] * grs[it];
                    H(it, it) = cc[it] * a + ss[it] * c;
                    res = std::fabs(grs[it + 1]);
                }
            } else {
                res = 0.0;
            }
            if (nullDir) {
                reason = -2;  // KSP_DIVERGED_NULL
                if (flexible) break;
            }
            ++it;
            ++its;
            rnorm = res;
            if (reason) break;
            hist.push_back(res);
            test(res);
            if (hapend && !reason) { reason = -5; break; }
        }
        if (built && it > 0) {
            // KSPGMRESBuildSoln / KSPFGMRESBuildSoln on columns 0 .. it-1.
            const int k1 = it - 1;
            bool solved = true;
            if (H(k1, k1) != 0.0) nrs[k1] = grs[k1] / H(k1, k1);
            else if (flexible) nrs[k1] = 0.0;
            else { reason = -5; solved = false; }
            for (int k = k1 - 1; solved && k >= 0; --k) {
                double t0 = grs[k];
                for (int j = k + 1; j <= k1; ++j) t0 = t0 - H(k, j) * nrs[j];
                if (!flexible && H(k, k) == 0.0) { reason = -5; solved = false; break; }
                nrs[k] = t0 / H(k, k);
            }
            if (solved) {
                GmCoef c{};
                for (int j = 0; j <= k1; ++j) c.c[j] = nrs[j];
                double* temp = flexible ? tmp.p : this->w.p;
                kGmCombine<<<nb, kBlock, 0, stream>>>(n, temp, flexible ? gmZ.p : V, ld, it, c);
                if (!flexible) {
                    applyPC(temp, tmp.p);  // KSPUnwindPreconditioner
                    temp = tmp.p;
                }
                kAxpy<<<nb, kBlock, 0, stream>>>(n, 1.0, temp, x);
                BSIM_LAUNCH_CHECK();
            }
        }
        if (!built) return;
        if (reason == 0 && its >= maxIt) reason = -3;
        if (reason) return;
    }
}

SolveRes DeviceSystem::solve(const double* b, double* x, int rhsIndex) {
    if (ready && configured && !setupPending && cfg.solver == 1 && cfg.precision == 1 && cfg.pc != 2) {
        const double* const b3[3] = {b, b, b};
        double* const x3[3] = {x, x, x};
        SolveRes res[3];
        std::vector<double> h[3];
        solve3Mixed(b3, x3, res, h, 1);
        hist = h[0];
        return res[0];
    }
    return solveFrom(b, x, rhsIndex, nullptr);
}

// The solve contract of petsc_linear_system.cpp:355-481.  With `first`, the
// prelude (norms) and the first Krylov call of attempt 0 were already run on
// (b, x) by the lockstep solver; their results, history and counters are
// taken from it, and `saved` must hold the component's initial guess.
SolveRes DeviceSystem::solveFrom(const double* b, double* x, int rhsIndex, const FirstRun* first) {
    if (!ready || !configured || setupPending)
        throw std::logic_error("linear system must update and prepare before solving a right-hand side");
    const auto tRhs = Clock::now();
    const int nb = blocksFor(n);
    hist.clear();
    ++counters.solves;
    auto resetGuess = [&]() {
        if (cfg.warmStart) kCopy<<<nb, kBlock, 0, stream>>>(n, saved.p, x);
        else kFill<<<nb, kBlock, 0, stream>>>(n, x, 0.0);
    };
    double rhsNorm, initial;
    if (first) {
        hist = first->hist;
        counters.spmvs += first->delta.spmvs;
        counters.trueResidualChecks += first->delta.trueResidualChecks;
        counters.pcApplications += first->delta.pcApplications;
        rhsNorm = first->rhsNorm;
        initial = first->initial;
    } else {
        // x0 = 0, or the current field when warm starting.
        kCopy<<<nb, kBlock, 0, stream>>>(n, x, saved.p);
        resetGuess();
        BSIM_LAUNCH_CHECK();
        kCopyNorm<<<reduceBlocks(n), kBlock, 0, stream>>>(n, b, nullptr, partials.p);
        kReduceTail<1><<<1, kTail, 0, stream>>>(reduceBlocks(n), partials.p, StoreSqrt{sc.p, S_BN});
        rhsNorm = readScalar(S_BN);
        counters.rhsSeconds += since(tRhs);
        initial = cfg.warmStart ? trueResidual(b, x, false) : rhsNorm;  // petsc_linear_system.cpp:410
    }
    const double scale = std::max(initial, rhsNorm) > 0.0 ? std::max(initial, rhsNorm) : 1.0;
    const double target = std::max(cfg.atol, cfg.rtol * scale);
    const bool reused = cfg.pc != 0 && pcMatrixUpdate < counters.matrixUpdates;
    if (reused) ++counters.pcReuses;  // per right-hand side (petsc_backend.cpp:262-266)
    SolveRes res;
    if (std::isfinite(initial) && initial <= target) {
        res.status = kConverged; res.iterations = 0; res.initial = initial; res.final_ = initial;
        res.relative = initial / scale;
        ++counters.trueResidualChecks;
        counters.last(res.status, res.initial, res.final_, res.relative);
        return res;
    }
    int total = 0;
    bool pcReused = reused;
    for (int attempt = 0; attempt < 2; ++attempt) {
        int attemptIts = 0, reason = 0;
        double fin = 0.0, internalTarget = target;
        bool guessZero = !cfg.warmStart;
        for (int restart = 0; restart < 3; ++restart) {
            int its = 0;
            const bool given = first && attempt == 0 && restart == 0;
            const auto tKsp = Clock::now();
            if (given) { its = first->its; reason = first->reason; }
            else if (cfg.solver == 0) cg(b, x, guessZero, cfg.maxIter - attemptIts, internalTarget, its, reason);
            else if (cfg.solver == 1) bcgs(b, x, guessZero, cfg.maxIter - attemptIts, internalTarget, its, reason);
            else gmres(b, x, guessZero, cfg.maxIter - attemptIts, internalTarget, its, reason, cfg.solver == 3);
            if (!given) counters.solveSeconds += since(tKsp);
            attemptIts += its;
            total += its;
            fin = given ? first->fin : trueResidual(b, x);
            if ((std::isfinite(fin) && fin <= target) || reason <= 0 || !std::isfinite(fin) || attemptIts >= cfg.maxIter)
                break;
            internalTarget *= 0.5;
            guessZero = false;
        }
        int status = kNumericalFailure;
        if (std::isfinite(fin) && fin <= target) status = kConverged;
        else if (reason == -3 || attemptIts >= cfg.maxIter) status = kMaxIterations;
        res.status = status; res.iterations = total; res.initial = initial; res.final_ = fin; res.relative = fin / scale;
        if (status != kConverged && attempt == 0 && pcReused) {
            setup(false);
            ++counters.retries;
            resetGuess();
            pcReused = false;
            continue;
        }
        if (status == kConverged) {
            int& fresh = freshIterations[rhsIndex];
            if (pcMatrixUpdate == counters.matrixUpdates) fresh = std::max(fresh, attemptIts);
            else if (attemptIts > cfg.reuseGrowth * static_cast<double>(std::max(fresh, 1))) forceRebuild = true;
        }
        break;
    }
    if (total > 0) ++pcRhsSolves;
    counters.krylovIterations += total;
    counters.last(res.status, res.initial, res.final_, res.relative);
    return res;
}

// The active mask goes to the device from a ring of pinned slots: a slot is
// only rewritten after the stream has synchronised since its last copy.
void DeviceSystem::uploadAct(const bool (&m)[3]) {
    int* slot = hact3 + 4 * (actSlot++ % kActSlots);
    for (int c = 0; c < 3; ++c) slot[c] = m[c] ? 1 : 0;
    BSIM_HIP(hipMemcpyAsync(act3.p, slot, sizeof(int) * 3, hipMemcpyHostToDevice, stream));
}

void DeviceSystem::syncScalars3() {
    BSIM_HIP(hipMemcpyAsync(hsc3, sc3.p, sizeof(double) * 3 * S_COUNT, hipMemcpyDeviceToHost, stream));
    BSIM_HIP(hipMemcpyAsync(reinterpret_cast<int*>(hsc3 + 3 * S_COUNT), fl3.p, sizeof(int) * 3, hipMemcpyDeviceToHost,
                            stream));
    BSIM_HIP(hipStreamSynchronize(stream));
}

namespace {
template <class T> Vec3p<T> vec3(DBuf<T> (&a)[3]) { return Vec3p<T>{{a[0].p, a[1].p, a[2].p}}; }
template <class T> CVec3p<T> cvec3(DBuf<T> (&a)[3]) { return CVec3p<T>{{a[0].p, a[1].p, a[2].p}}; }
__global__ void __launch_bounds__(kBlock) k3Narrow(int n, CVec3p<double> a, Vec3p<float> b, const int* act) {  // b = float(a)
    const int i = globalThread();
    if (i >= n) return;
    for (int c = 0; c < 3; ++c)
        if (act[c]) b.c[c][i] = static_cast<float>(a.c[c][i]);
}
__global__ void __launch_bounds__(kBlock) k3AddWide(int n, CVec3p<float> d, Vec3p<double> x, const int* act) {  // x += double(d)
    const int i = globalThread();
    if (i >= n) return;
    for (int c = 0; c < 3; ++c)
        if (act[c]) x.c[c][i] += static_cast<double>(d.c[c][i]);
}
}  // namespace

template <class T> void DeviceSystem::Work3<T>::ensure(int n) {
    for (int c = 0; c < 3; ++c)
        for (DBuf<T>* d : {&r[c], &p[c], &v[c], &t[c], &s[c], &rp[c], &tmp[c], &guess[c]}) d->ensure(n);
}

// One BiCGSTAB call (bcgs() above, bcgs.c:14-140) on every component with
// run[c], in lockstep.  Each component keeps the solo control flow: its own
// convergence test, breakdown exits and iteration count; a component that
// stops is frozen through act3 while the others go on.  Counters go to dl[c].
template <class T>
void DeviceSystem::bcgs3(Work3<T>& w, const CsrView<T>& A, const T* jv, const T* const b[3], T* const x[3],
                         const bool (&run)[3], bool guessZero, int maxIt, const double (&target)[3], int (&its)[3],
                         int (&reason)[3], std::vector<double> (&h)[3], SystemCounters (&dl)[3]) {
    double* part = partials.p;
    const int rb = reduceBlocks(n), nb = blocksFor(n);
    const CVec3p<T> B{{b[0], b[1], b[2]}};
    const Vec3p<T> X{{x[0], x[1], x[2]}};
    const CVec3p<T> cX{{x[0], x[1], x[2]}};
    int* act = act3.p;
    auto upload = [&](const bool (&m)[3]) { uploadAct(m); };
    bool on[3] = {run[0], run[1], run[2]};
    upload(on);
    BSIM_HIP(hipMemsetAsync(fl3.p, 0, sizeof(int) * 3, stream));
    {
        const double one[S_COUNT] = {0, 1.0, 1.0, 0, 1.0, 0, 1.0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int c = 0; c < 3; ++c) std::memcpy(hsc3 + c * S_COUNT, one, sizeof(one));
        BSIM_HIP(hipMemcpyAsync(sc3.p, hsc3, sizeof(double) * 3 * S_COUNT, hipMemcpyHostToDevice, stream));
    }
    const bool haveGuess = !guessZero;
    if (!guessZero) {
        k3ResidualNorm<T><<<rb, kBlock, 0, stream>>>(A, B, cX, vec3(w.r), act, part);
        for (int c = 0; c < 3; ++c) if (on[c]) ++dl[c].spmvs;
        k3Copy<T><<<nb, kBlock, 0, stream>>>(n, cX, vec3(w.guess), act);
        k3Fill<T><<<nb, kBlock, 0, stream>>>(n, X, 0.0, act);
    } else {
        k3CopyNorm<T><<<rb, kBlock, 0, stream>>>(n, B, vec3(w.r), act, part);
    }
    kReduceTail<3><<<1, kTail, 0, stream>>>(rb, part, Tail3Sqrt{sc3.p, S_DP, act});
    BSIM_LAUNCH_CHECK();
    syncScalars3();
    double dp[3] = {0.0, 0.0, 0.0};
    bool loop[3] = {false, false, false}, atStart[3] = {false, false, false};
    auto test = [&](int c, double vv) {
        if (!std::isfinite(vv)) { reason[c] = -9; return true; }
        if (vv <= target[c]) { reason[c] = 2; return true; }
        return false;
    };
    for (int c = 0; c < 3; ++c) {
        if (!on[c]) continue;
        dp[c] = hsc3[c * S_COUNT + S_DP];
        its[c] = 0;
        reason[c] = 0;
        h[c].push_back(dp[c]);
        if (test(c, dp[c])) atStart[c] = true;
        else loop[c] = true;
    }
    if (haveGuess && (atStart[0] || atStart[1] || atStart[2])) {  // x = guess for those that stop at once
        upload(atStart);
        k3Axpy<T><<<nb, kBlock, 0, stream>>>(n, 1.0, cvec3(w.guess), X, act);
    }
    bool active[3] = {loop[0], loop[1], loop[2]};
    upload(active);
    k3Copy<T><<<nb, kBlock, 0, stream>>>(n, cvec3(w.r), vec3(w.rp), act);
    k3Fill<T><<<nb, kBlock, 0, stream>>>(n, vec3(w.p), 0.0, act);
    k3Fill<T><<<nb, kBlock, 0, stream>>>(n, vec3(w.v), 0.0, act);
    int iter[3] = {0, 0, 0};
    while (active[0] || active[1] || active[2]) {
        k3Dot<T><<<rb, kBlock, 0, stream>>>(n, cvec3(w.r), cvec3(w.rp), act, part);
        kReduceTail<3><<<1, kTail, 0, stream>>>(rb, part, Tail3Rho{sc3.p, act});
        k3BcgsPJacobi<T><<<nb, kBlock, 0, stream>>>(n, vec3(w.p), cvec3(w.r), cvec3(w.v), sc3.p, jv, vec3(w.tmp), act);
        k3SpmvD1<T><<<rb, kBlock, 0, stream>>>(A, cvec3(w.tmp), vec3(w.v), cvec3(w.rp), act, part);
        kReduceTail<3><<<1, kTail, 0, stream>>>(rb, part, Tail3Alpha{sc3.p, fl3.p, act});
        k3BcgsSJacobi<T><<<nb, kBlock, 0, stream>>>(n, vec3(w.s), cvec3(w.v), cvec3(w.r), sc3.p, fl3.p, jv, vec3(w.r), act);
        k3SpmvD2<T><<<rb, kBlock, 0, stream>>>(A, cvec3(w.r), vec3(w.t), cvec3(w.s), fl3.p, act, part);
        kReduceTail<6><<<1, kTail, 0, stream>>>(rb, part, Tail3Omega{sc3.p, fl3.p, act});
        k3BcgsUpdate<T><<<rb, kBlock, 0, stream>>>(n, X, cvec3(w.p), cvec3(w.s), vec3(w.r), cvec3(w.t), sc3.p, fl3.p,
                                                   act, part);
        kReduceTail<3><<<1, kTail, 0, stream>>>(rb, part, Tail3Norm{sc3.p, fl3.p, act});
        BSIM_LAUNCH_CHECK();
        syncScalars3();
        bool changed = false;
        for (int c = 0; c < 3; ++c) {
            if (!active[c]) continue;
            dl[c].spmvs += 2;
            dl[c].pcApplications += 2;
            const double* hs = hsc3 + c * S_COUNT;
            const int fl = reinterpret_cast<const int*>(hsc3 + 3 * S_COUNT)[c];
            bool stop = false;
            if (fl & F_NONFINITE) { reason[c] = -9; stop = true; }
            else if (fl & F_D1_ZERO) { reason[c] = -5; stop = true; }
            else if (fl & F_D2_ZERO) {
                stop = true;
                const bool only[3] = {c == 0, c == 1, c == 2};
                upload(only);
                k3Dot<T><<<rb, kBlock, 0, stream>>>(n, cvec3(w.s), cvec3(w.s), act, part);
                kReduceTail<3><<<1, kTail, 0, stream>>>(rb, part, Tail3Sqrt{sc3.p, S_TMP, act});
                BSIM_LAUNCH_CHECK();
                syncScalars3();
                const double ss = hsc3[c * S_COUNT + S_TMP];
                if (ss != 0.0) reason[c] = -5;
                else {
                    k3Axpy<T><<<nb, kBlock, 0, stream>>>(n, hs[S_ALPHA], cvec3(w.p), X, act);
                    ++its[c];
                    reason[c] = 2;
                    h[c].push_back(dp[c]);
                }
                changed = true;  // the mask was replaced
            } else {
                dp[c] = hs[S_DP];
                ++its[c];
                h[c].push_back(dp[c]);
                if (test(c, dp[c])) stop = true;
                else if (hs[S_RHO] == 0.0) { reason[c] = -5; stop = true; }
                else if (++iter[c] >= maxIt) { reason[c] = -3; stop = true; }
            }
            if (stop) { active[c] = false; changed = true; }
        }
        if (changed) upload(active);
    }
    // Unwind the right preconditioning for the components that iterated.
    upload(loop);
    k3Jacobi<T><<<nb, kBlock, 0, stream>>>(n, cX, jv, vec3(w.tmp), act);
    k3Copy<T><<<nb, kBlock, 0, stream>>>(n, cvec3(w.tmp), X, act);
    if (haveGuess) k3Axpy<T><<<nb, kBlock, 0, stream>>>(n, 1.0, cvec3(w.guess), X, act);
    BSIM_LAUNCH_CHECK();
    for (int c = 0; c < 3; ++c) if (loop[c]) ++dl[c].pcApplications;
}

void DeviceSystem::solve3(const double* const b[3], double* const x[3], SolveRes res[3], std::vector<double> hists[3]) {
    if (borrowed) throw std::logic_error("the lockstep vector solve allocates; not available in a fixed workspace");
    if (!ready || !configured || setupPending)
        throw std::logic_error("linear system must update and prepare before solving a right-hand side");
    if (cfg.solver == 1 && cfg.pc != 2 && cfg.precision == 1) {
        solve3Mixed(b, x, res, hists, 3);
        return;
    }
    const bool lockstep = cfg.solver == 1 && cfg.pc <= 1 && cfg.precision == 0;
    if (!lockstep) {
        for (int c = 0; c < 3; ++c) {
            res[c] = solve(b[c], x[c], c);
            hists[c] = hist;
        }
        return;
    }
    w64.ensure(n);
    for (int c = 0; c < 3; ++c) saved3[c].ensure(n);
    const int rb = reduceBlocks(n), nb = blocksFor(n);
    const CVec3p<double> B{{b[0], b[1], b[2]}};
    const Vec3p<double> X{{x[0], x[1], x[2]}};
    const CVec3p<double> cX{{x[0], x[1], x[2]}};
    // Prelude of every component (solveFrom): keep the guess, reset, norms.
    uploadAct({true, true, true});
    k3Copy<double><<<nb, kBlock, 0, stream>>>(n, cX, vec3(saved3), act3.p);
    if (!cfg.warmStart) k3Fill<double><<<nb, kBlock, 0, stream>>>(n, X, 0.0, act3.p);
    k3CopyNorm<double><<<rb, kBlock, 0, stream>>>(n, B, Vec3p<double>{{nullptr, nullptr, nullptr}}, act3.p, partials.p);
    kReduceTail<3><<<1, kTail, 0, stream>>>(rb, partials.p, Tail3Sqrt{sc3.p, S_BN, act3.p});
    if (cfg.warmStart) {
        k3ResidualNorm<double><<<rb, kBlock, 0, stream>>>(view(), B, cX, vec3(w64.tmp), act3.p, partials.p);
        kReduceTail<3><<<1, kTail, 0, stream>>>(rb, partials.p, Tail3Sqrt{sc3.p, S_TR, act3.p});
    }
    BSIM_LAUNCH_CHECK();
    syncScalars3();
    FirstRun fr[3];
    bool run[3];
    double target[3];
    for (int c = 0; c < 3; ++c) {
        fr[c].rhsNorm = hsc3[c * S_COUNT + S_BN];
        fr[c].initial = cfg.warmStart ? hsc3[c * S_COUNT + S_TR] : fr[c].rhsNorm;
        if (cfg.warmStart) ++fr[c].delta.spmvs;  // the initial residual is not a check
        const double scale = std::max(fr[c].initial, fr[c].rhsNorm) > 0.0 ? std::max(fr[c].initial, fr[c].rhsNorm) : 1.0;
        target[c] = std::max(cfg.atol, cfg.rtol * scale);
        run[c] = !(std::isfinite(fr[c].initial) && fr[c].initial <= target[c]);
    }
    // First Krylov call of attempt 0 for every component that needs one.
    int its[3] = {0, 0, 0}, reason[3] = {0, 0, 0};
    std::vector<double> h[3];
    SystemCounters dl[3];
    if (run[0] || run[1] || run[2]) {
        const double* jv = cfg.pc == 1 ? jinv.p : nullptr;
        const auto tKsp = Clock::now();
        bcgs3<double>(w64, view(), jv, b, x, run, !cfg.warmStart, cfg.maxIter, target, its, reason, h, dl);
        counters.solveSeconds += since(tKsp);
        uploadAct(run);
        k3ResidualNorm<double><<<rb, kBlock, 0, stream>>>(view(), B, cX, vec3(w64.tmp), act3.p, partials.p);
        kReduceTail<3><<<1, kTail, 0, stream>>>(rb, partials.p, Tail3Sqrt{sc3.p, S_TR, act3.p});
        BSIM_LAUNCH_CHECK();
        syncScalars3();
    }
    for (int c = 0; c < 3; ++c) {
        if (!run[c]) continue;
        fr[c].its = its[c];
        fr[c].reason = reason[c];
        fr[c].fin = hsc3[c * S_COUNT + S_TR];
        fr[c].hist = std::move(h[c]);
        fr[c].delta.spmvs += dl[c].spmvs + 1;
        fr[c].delta.pcApplications += dl[c].pcApplications;
        ++fr[c].delta.trueResidualChecks;
    }
    // Commit in component order.  A preconditioner rebuild during one
    // component's retry changes what the later components would have seen,
    // so they are solved again from their initial guesses.
    bool diverged = false;
    for (int c = 0; c < 3; ++c) {
        if (diverged) {
            kCopy<<<nb, kBlock, 0, stream>>>(n, saved3[c].p, x[c]);
            res[c] = solve(b[c], x[c], c);
        } else {
            kCopy<<<nb, kBlock, 0, stream>>>(n, saved3[c].p, saved.p);
            const std::uint64_t retries = counters.retries;
            res[c] = solveFrom(b[c], x[c], c, &fr[c]);
            diverged = counters.retries != retries;
        }
        hists[c] = hist;
    }
    BSIM_LAUNCH_CHECK();
}

// Mixed precision for a vector equation: FP64 iterative refinement around an
// FP32 lockstep BiCGSTAB (FP32 matrix, Jacobi preconditioner and vectors,
// FP64 inner products).  Each sweep forms the true residual r = b - A x in
// FP64, solves A d = r in FP32 to inner_rtol * ||r|| and adds d in FP64; a
// component stops when its true residual meets the target of the solve
// contract (max(atol, rtol * max(||r0||, ||b||))).  The history holds the
// true residual before every sweep.
void DeviceSystem::solve3Mixed(const double* const b[3], double* const x[3], SolveRes res[3],
                               std::vector<double> hists[3], int ncomp) {
    const auto t0 = Clock::now();
    w64.ensure(n);
    w32.ensure(n);
    for (int c = 0; c < 3; ++c) { saved3[c].ensure(n); d32[c].ensure(n); r32[c].ensure(n); }
    const int rb = reduceBlocks(n), nb = blocksFor(n);
    // FP32 copies of the effective matrix and of the Jacobi inverse.
    const std::uint64_t stamp = counters.matrixUpdates * 1000003ull + counters.pcSetups;
    if (stamp != f32Stamp) {
        diag32.ensure(n);
        val32.ensure(std::max(nnz, 1));
        kToFloat<<<nb, kBlock, 0, stream>>>(n, diag(), diag32.p);
        if (nnz) kToFloat<<<blocksFor(nnz), kBlock, 0, stream>>>(nnz, val(), val32.p);
        if (cfg.pc == 1) {
            jinv32.ensure(n);
            kToFloat<<<nb, kBlock, 0, stream>>>(n, jinv.p, jinv32.p);
        }
        f32Stamp = stamp;
    }
    const CsrView<float> A32{n, pat.ptr, pat.col, diag32.p, val32.p, pat.sliceOff, pat.K, nullptr};
    const float* jv32 = cfg.pc == 1 ? jinv32.p : nullptr;
    const CVec3p<double> B{{b[0], b[1], b[2]}};
    const Vec3p<double> X{{x[0], x[1], x[2]}};
    const CVec3p<double> cX{{x[0], x[1], x[2]}};
    counters.solves += ncomp;
    if (cfg.pc != 0 && pcMatrixUpdate < counters.matrixUpdates) counters.pcReuses += ncomp;
    const bool use[3] = {true, ncomp > 1, ncomp > 2};
    uploadAct(use);
    if (!cfg.warmStart) k3Fill<double><<<nb, kBlock, 0, stream>>>(n, X, 0.0, act3.p);
    k3CopyNorm<double><<<rb, kBlock, 0, stream>>>(n, B, Vec3p<double>{{nullptr, nullptr, nullptr}}, act3.p, partials.p);
    kReduceTail<3><<<1, kTail, 0, stream>>>(rb, partials.p, Tail3Sqrt{sc3.p, S_BN, act3.p});
    BSIM_LAUNCH_CHECK();
    syncScalars3();
    double rhsNorm[3], initial[3] = {0, 0, 0}, target[3], scale[3], dp[3] = {0, 0, 0}, prev[3] = {0, 0, 0};
    int total[3] = {0, 0, 0}, stalled[3] = {0, 0, 0};
    bool done[3] = {!use[0], !use[1], !use[2]}, failed[3] = {false, false, false};
    for (int c = 0; c < 3; ++c) { rhsNorm[c] = hsc3[c * S_COUNT + S_BN]; hists[c].clear(); }
    for (int sweep = 0;; ++sweep) {
        bool live[3];
        for (int c = 0; c < 3; ++c) live[c] = !done[c];
        uploadAct(live);
        k3ResidualNorm<double><<<rb, kBlock, 0, stream>>>(view(), B, cX, vec3(w64.r), act3.p, partials.p);
        kReduceTail<3><<<1, kTail, 0, stream>>>(rb, partials.p, Tail3Sqrt{sc3.p, S_TR, act3.p});
        BSIM_LAUNCH_CHECK();
        syncScalars3();
        bool run[3] = {false, false, false};
        double inner[3] = {0, 0, 0};
        for (int c = 0; c < 3; ++c) {
            if (done[c]) continue;
            ++counters.spmvs;
            ++counters.trueResidualChecks;
            dp[c] = hsc3[c * S_COUNT + S_TR];
            if (sweep == 0) {
                initial[c] = cfg.warmStart ? dp[c] : rhsNorm[c];
                scale[c] = std::max(initial[c], rhsNorm[c]) > 0.0 ? std::max(initial[c], rhsNorm[c]) : 1.0;
                target[c] = std::max(cfg.atol, cfg.rtol * scale[c]);
            }
            hists[c].push_back(dp[c]);
            if (!std::isfinite(dp[c])) { failed[c] = true; done[c] = true; continue; }
            // Refinement that stops reducing the residual (two sweeps without a
            // tenth of reduction) ends as a numerical failure.
            if (sweep > 0) stalled[c] = dp[c] > 0.9 * prev[c] ? stalled[c] + 1 : 0;
            prev[c] = dp[c];
            if (stalled[c] >= 2 && dp[c] > target[c]) { failed[c] = true; done[c] = true; continue; }
            if (dp[c] <= target[c] || total[c] >= cfg.maxIter || failed[c]) { done[c] = true; continue; }
            run[c] = true;
            inner[c] = std::max(cfg.innerRtol * dp[c], 0.5 * target[c]);
        }
        if (!run[0] && !run[1] && !run[2]) break;
        // d = A32^{-1} float(r) from a zero guess, then x += d.
        uploadAct(run);
        k3Narrow<<<nb, kBlock, 0, stream>>>(n, cvec3(w64.r), vec3(r32), act3.p);
        k3Fill<float><<<nb, kBlock, 0, stream>>>(n, vec3(d32), 0.0, act3.p);
        const float* rb32[3] = {r32[0].p, r32[1].p, r32[2].p};
        float* db32[3] = {d32[0].p, d32[1].p, d32[2].p};
        int maxIt = cfg.maxIter;
        for (int c = 0; c < 3; ++c) if (run[c]) maxIt = std::min(maxIt, cfg.maxIter - total[c]);
        int its[3] = {0, 0, 0}, reason[3] = {0, 0, 0};
        std::vector<double> h[3];
        SystemCounters dl[3];
        bcgs3<float>(w32, A32, jv32, rb32, db32, run, true, std::max(maxIt, 1), inner, its, reason, h, dl);
        uploadAct(run);
        k3AddWide<<<nb, kBlock, 0, stream>>>(n, cvec3(d32), X, act3.p);
        BSIM_LAUNCH_CHECK();
        for (int c = 0; c < 3; ++c) {
            if (!run[c]) continue;
            total[c] += its[c];
            counters.spmvs += dl[c].spmvs;
            counters.pcApplications += dl[c].pcApplications;
            if (reason[c] < 0 && reason[c] != -3 && its[c] == 0) failed[c] = true;  // breakdown without progress
        }
    }
    for (int c = 0; c < ncomp; ++c) {
        SolveRes& r = res[c];
        r.iterations = total[c];
        r.initial = initial[c];
        r.final_ = dp[c];
        r.relative = dp[c] / scale[c];
        if (std::isfinite(dp[c]) && dp[c] <= target[c]) r.status = kConverged;
        else if (!failed[c] && total[c] >= cfg.maxIter) r.status = kMaxIterations;
        else r.status = kNumericalFailure;
        if (total[c] > 0) ++pcRhsSolves;
        counters.krylovIterations += total[c];
        counters.last(r.status, r.initial, r.final_, r.relative);
    }
    counters.solveSeconds += since(t0);
}

}  // namespace bsim
