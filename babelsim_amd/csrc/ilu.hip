// Level-scheduled incomplete factorisations with no fill: IC(0) for PETSc's
// icc (petsc_linear_system.cpp:204-208; PCICC factors the upper triangle) and
// ILU(0) for asm on one rank (:224-226; one subdomain covering every row, the
// sub-PC PETSc's default ILU(0)).
//
// The rows are eliminated in the caller's numbering (the system's own rows
// keep the renumbered storage; each row's entries are sorted by their
// columns' positions in the caller's order), so the factors are those PETSc
// forms from the caller's matrix in its natural order.
//
// The symbolic analysis runs once per pattern on the host: the factor pattern
// (row-major, entries in elimination order), the stored entries each factor entry sums,
// the (i, j) <- (k, j) update pairs of each lower entry, and the levels of the
// lower and upper dependency graphs.  Rows of one level are independent; each
// is one thread's sequential work (IKJ elimination, or one row of a
// substitution), so a row's arithmetic is in the same order whatever the
// schedule.  Consecutive levels of at most kIluRunRows rows run in one
// workgroup with a barrier between levels; a larger level is a grid launch.
#include "ilu.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace bsim {

namespace {

constexpr int kIluRunThreads = 1024;
constexpr int kIluRunRows = 2048;

// Factor values: each entry sums its sources in order (a source p >= 0 is the
// stored off-diagonal value val[p], p < 0 the diagonal diag[-p - 1]).
__global__ void __launch_bounds__(kBlock) kIluGather(int nf, const int* srcPtr, const int* src, const double* diag,
                                                     const double* val, double* f) {
    const int t = globalThread();
    if (t >= nf) return;
    double s = 0.0;
    for (int q = srcPtr[t]; q < srcPtr[t + 1]; ++q) {
        const int p = src[q];
        const double v = p >= 0 ? val[p] : diag[-p - 1];
        s = q == srcPtr[t] ? v : s + v;
    }
    f[t] = s;
}

struct IluUpdates {
    const int* ptr;
    const int* dst;
    const int* src;
};

// IKJ elimination of row i: for each lower entry (i, k) in column order,
// l = a_ik / u_kk, then a_ij -= l u_kj for each j > k in both rows.
__device__ inline void factorRow(const IluView& V, const IluUpdates& U, int i, bool chol, int* flag) {
    double* F = V.f;
    const int d = V.dpos[i];
    for (int e = V.ptr[i]; e < d; ++e) {
        const double l = F[e] / F[V.dpos[V.col[e]]];
        F[e] = l;
        for (int q = U.ptr[e]; q < U.ptr[e + 1]; ++q) F[U.dst[q]] -= l * F[U.src[q]];
    }
    const double p = F[d];
    const bool ok = chol ? (p > 0.0 && __builtin_isfinite(p)) : (p != 0.0 && __builtin_isfinite(p));
    if (!ok) *flag = 1;
}
// y -= sum over entries [e0, e1) of f[e] z[col[e]], in entry order.  The
// entries are taken kIluChunk at a time: their columns and factors load
// together, then the z values together, then the products are subtracted in
// order.
constexpr int kIluChunk = 8;
__device__ inline double subtractRow(const IluView& V, int e0, int e1, const double* z, double y) {
    for (int e = e0; e < e1; e += kIluChunk) {
        int c[kIluChunk];
        double fv[kIluChunk], zv[kIluChunk];
#pragma unroll
        for (int k = 0; k < kIluChunk; ++k) {
            const bool in = e + k < e1;
            c[k] = in ? V.col[e + k] : 0;
            fv[k] = in ? V.f[e + k] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kIluChunk; ++k) zv[k] = e + k < e1 ? z[c[k]] : 0.0;
#pragma unroll
        for (int k = 0; k < kIluChunk; ++k)
            if (e + k < e1) y -= fv[k] * zv[k];
    }
    return y;
}
__device__ inline void forwardRow(const IluView& V, int i, const double* r, double* z) {
    z[i] = subtractRow(V, V.ptr[i], V.dpos[i], z, r[i]);
}
__device__ inline void backwardRow(const IluView& V, int i, double* z) {
    const int d = V.dpos[i];
    z[i] = subtractRow(V, d + 1, V.ptr[i + 1], z, z[i]) / V.f[d];
}

__global__ void __launch_bounds__(kBlock) kIluFactorLevel(IluView V, IluUpdates U, const int* rows, int count,
                                                          bool chol, int* flag) {
    const int t = globalThread();
    if (t < count) factorRow(V, U, rows[t], chol, flag);
}
__global__ void __launch_bounds__(kIluRunThreads) kIluFactorRun(IluView V, IluUpdates U, const int* lptr,
                                                                const int* rows, int l0, int l1, bool chol,
                                                                int* flag) {
    for (int l = l0; l < l1; ++l) {
        for (int t = lptr[l] + threadIdx.x; t < lptr[l + 1]; t += kIluRunThreads) factorRow(V, U, rows[t], chol, flag);
        __syncthreads();
    }
}
__global__ void __launch_bounds__(kBlock) kIluForwardLevel(IluView V, const int* rows, int count, const double* r,
                                                           double* z) {
    const int t = globalThread();
    if (t < count) forwardRow(V, rows[t], r, z);
}
__global__ void __launch_bounds__(kIluRunThreads) kIluForwardRun(IluView V, const int* lptr, const int* rows, int l0,
                                                                 int l1, const double* r, double* z) {
    for (int l = l0; l < l1; ++l) {
        for (int t = lptr[l] + threadIdx.x; t < lptr[l + 1]; t += kIluRunThreads) forwardRow(V, rows[t], r, z);
        __syncthreads();
    }
}
__global__ void __launch_bounds__(kBlock) kIluBackwardLevel(IluView V, const int* rows, int count, double* z) {
    const int t = globalThread();
    if (t < count) backwardRow(V, rows[t], z);
}
__global__ void __launch_bounds__(kIluRunThreads) kIluBackwardRun(IluView V, const int* lptr, const int* rows, int l0,
                                                                  int l1, double* z) {
    for (int l = l0; l < l1; ++l) {
        for (int t = lptr[l] + threadIdx.x; t < lptr[l + 1]; t += kIluRunThreads) backwardRow(V, rows[t], z);
        __syncthreads();
    }
}
__global__ void __launch_bounds__(kBlock) kIluSetInt(int* p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}
__global__ void __launch_bounds__(kBlock) kIluInf(int n, double* z) {
    const int i = globalThread();
    if (i < n) z[i] = HUGE_VAL;
}

}  // namespace

void LevelIlu::schedule(Schedule& S, const std::vector<int>& level, int levels) {
    S.hptr.assign(levels + 1, 0);
    for (int i = 0; i < n; ++i) ++S.hptr[level[i] + 1];
    for (int l = 0; l < levels; ++l) S.hptr[l + 1] += S.hptr[l];
    S.hrows.assign(n, 0);
    std::vector<int> next(S.hptr.begin(), S.hptr.end() - 1);
    for (int i = 0; i < n; ++i) S.hrows[next[level[i]]++] = i;
    S.segs.clear();
    for (int l = 0; l < levels; ++l) {
        const int cnt = S.hptr[l + 1] - S.hptr[l];
        if (cnt > kIluRunRows) S.segs.push_back({l, l + 1, true, cnt});
        else if (!S.segs.empty() && !S.segs.back().grid) { S.segs.back().l1 = l + 1; S.segs.back().rows += cnt; }
        else S.segs.push_back({l, l + 1, false, cnt});
    }
    S.ptr.upload(S.hptr, stream);
    S.rows.upload(S.hrows, stream);
}

void LevelIlu::analyse(const Pattern& P, int nn) {
    n = nn;
    const std::vector<int>& rp = P.hptr;
    const std::vector<int>& rc = P.hcol;
    if (static_cast<int>(rp.size()) != n + 1) throw std::invalid_argument("incomplete factorisation: pattern size");
    auto stored = [&](int q) { return P.csrToStore.empty() ? q : P.csrToStore[q]; };
    // Elimination order: each row's position in the caller's numbering, so
    // that the factors are those of the caller's matrix in its natural order.
    std::vector<int> pos(n), at(n, -1);
    for (int i = 0; i < n; ++i) {
        pos[i] = P.order.empty() ? i : P.order[i];
        if (pos[i] < 0 || pos[i] >= n || at[pos[i]] >= 0)
            throw std::invalid_argument("incomplete factorisation: the row order is not a permutation");
        at[pos[i]] = i;
    }
    auto before = [&](int a, int b) { return pos[a] < pos[b]; };
    // Factor pattern: the diagonal and the off-diagonal columns of each row;
    // for IC(0) the upper triangle's columns (later in the order) and their
    // transposes.
    hptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        ++hptr[i + 1];
        for (int q = rp[i]; q < rp[i + 1]; ++q) {
            const int j = rc[q];
            if (j < 0 || j >= n) throw std::invalid_argument("incomplete factorisation: column out of range");
            if (j == i) continue;
            if (!chol) ++hptr[i + 1];
            else if (before(i, j)) { ++hptr[i + 1]; ++hptr[j + 1]; }
        }
    }
    for (int i = 0; i < n; ++i) hptr[i + 1] += hptr[i];
    std::vector<int> cols(hptr[n]);
    {
        std::vector<int> fill(hptr.begin(), hptr.end() - 1);
        for (int i = 0; i < n; ++i) {
            cols[fill[i]++] = i;
            for (int q = rp[i]; q < rp[i + 1]; ++q) {
                const int j = rc[q];
                if (j == i) continue;
                if (!chol) cols[fill[i]++] = j;
                else if (before(i, j)) { cols[fill[i]++] = j; cols[fill[j]++] = i; }
            }
        }
    }
    // Entries in elimination order; repeated columns merged.
    std::vector<int> cptr(n + 1, 0);
    hcol.clear();
    hcol.reserve(cols.size());
    for (int i = 0; i < n; ++i) {
        auto b = cols.begin() + hptr[i], e = cols.begin() + hptr[i + 1];
        std::sort(b, e, before);
        e = std::unique(b, e);
        hcol.insert(hcol.end(), b, e);
        cptr[i + 1] = static_cast<int>(hcol.size());
    }
    hptr.swap(cptr);
    const int nf = hptr[n];
    auto find = [&](int i, int j) {
        return static_cast<int>(std::lower_bound(hcol.begin() + hptr[i], hcol.begin() + hptr[i + 1], j, before) -
                                hcol.begin());
    };
    hdpos.assign(n, 0);
    for (int i = 0; i < n; ++i) hdpos[i] = find(i, i);
    // Sources of each factor entry: the diagonal first, then the stored
    // entries in row-major order.
    std::vector<int> scount(nf + 1, 0);
    for (int i = 0; i < n; ++i) {
        ++scount[hdpos[i] + 1];
        for (int q = rp[i]; q < rp[i + 1]; ++q) {
            const int j = rc[q];
            if (j == i) ++scount[hdpos[i] + 1];
            else if (!chol) ++scount[find(i, j) + 1];
            else if (before(i, j)) { ++scount[find(i, j) + 1]; ++scount[find(j, i) + 1]; }
        }
    }
    for (int t = 0; t < nf; ++t) scount[t + 1] += scount[t];
    std::vector<int> hsrc(scount[nf]);
    {
        std::vector<int> fill(scount.begin(), scount.end() - 1);
        for (int i = 0; i < n; ++i) hsrc[fill[hdpos[i]]++] = -(i + 1);
        for (int i = 0; i < n; ++i)
            for (int q = rp[i]; q < rp[i + 1]; ++q) {
                const int j = rc[q];
                const int s = stored(q);
                if (j == i) hsrc[fill[hdpos[i]]++] = s;
                else if (!chol) { const int t = find(i, j); hsrc[fill[t]++] = s; }
                else if (before(i, j)) {
                    const int t = find(i, j), tt = find(j, i);
                    hsrc[fill[t]++] = s;
                    hsrc[fill[tt]++] = s;
                }
            }
    }
    // Update pairs of each lower entry (i, k): (i, j) and (k, j) for the
    // columns j after k in both rows, in elimination order.
    std::vector<int> uptr(nf + 1, 0), udst, usrc;
    for (int i = 0; i < n; ++i) {
        for (int e = hptr[i]; e < hptr[i + 1]; ++e) {
            if (e < hdpos[i]) {
                const int k = hcol[e];
                int a = e + 1, b = hdpos[k] + 1;
                while (a < hptr[i + 1] && b < hptr[k + 1]) {
                    if (hcol[a] == hcol[b]) { udst.push_back(a); usrc.push_back(b); ++a; ++b; }
                    else if (before(hcol[a], hcol[b])) ++a;
                    else ++b;
                }
            }
            uptr[e + 1] = static_cast<int>(udst.size());
        }
    }
    // Levels: a row follows every row its lower (forward) or upper (backward)
    // entries refer to.
    std::vector<int> lf(n, 0), lb(n, 0);
    int nlf = 0, nlb = 0;
    for (int p = 0; p < n; ++p) {
        const int i = at[p];
        int l = 0;
        for (int e = hptr[i]; e < hdpos[i]; ++e) l = std::max(l, lf[hcol[e]] + 1);
        lf[i] = l;
        nlf = std::max(nlf, l + 1);
    }
    for (int p = n - 1; p >= 0; --p) {
        const int i = at[p];
        int l = 0;
        for (int e = hdpos[i] + 1; e < hptr[i + 1]; ++e) l = std::max(l, lb[hcol[e]] + 1);
        lb[i] = l;
        nlb = std::max(nlb, l + 1);
    }
    fptr.upload(hptr, stream);
    fcol.upload(hcol, stream);
    dpos.upload(hdpos, stream);
    srcPtr.upload(scount, stream);
    src.upload(hsrc, stream);
    updPtr.upload(uptr, stream);
    if (udst.empty()) { udst.push_back(0); usrc.push_back(0); }
    updDst.upload(udst, stream);
    updSrc.upload(usrc, stream);
    f.ensure(std::max(nf, 1));
    flag.ensure(1);
    schedule(fwd, lf, nlf);
    schedule(bwd, lb, nlb);
    // The uploads read host vectors that go out of scope on return.
    BSIM_HIP(hipStreamSynchronize(stream));
}

bool LevelIlu::factor(const CsrView<double>& A) {
    if (A.n != n) throw std::logic_error("incomplete factorisation: matrix size differs from the analysed pattern");
    const int nf = hptr[n];
    kIluGather<<<blocksFor(nf), kBlock, 0, stream>>>(nf, srcPtr.p, src.p, A.diag, A.val, f.p);
    kIluSetInt<<<1, kBlock, 0, stream>>>(flag.p, 0);
    const IluUpdates U{updPtr.p, updDst.p, updSrc.p};
    for (const Seg& s : fwd.segs) {
        if (s.grid)
            kIluFactorLevel<<<blocksFor(s.rows), kBlock, 0, stream>>>(view(), U, fwd.rows.p + fwd.hptr[s.l0], s.rows,
                                                                       chol, flag.p);
        else
            kIluFactorRun<<<1, kIluRunThreads, 0, stream>>>(view(), U, fwd.ptr.p, fwd.rows.p, s.l0, s.l1, chol, flag.p);
    }
    BSIM_LAUNCH_CHECK();
    int h = 0;
    BSIM_HIP(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    BSIM_HIP(hipStreamSynchronize(stream));
    fail = h != 0;
    return !fail;
}

void LevelIlu::apply(const double* r, double* z) {
    if (fail) {
        kIluInf<<<blocksFor(n), kBlock, 0, stream>>>(n, z);
        BSIM_LAUNCH_CHECK();
        return;
    }
    for (const Seg& s : fwd.segs) {
        if (s.grid)
            kIluForwardLevel<<<blocksFor(s.rows), kBlock, 0, stream>>>(view(), fwd.rows.p + fwd.hptr[s.l0], s.rows, r, z);
        else
            kIluForwardRun<<<1, kIluRunThreads, 0, stream>>>(view(), fwd.ptr.p, fwd.rows.p, s.l0, s.l1, r, z);
    }
    for (const Seg& s : bwd.segs) {
        if (s.grid)
            kIluBackwardLevel<<<blocksFor(s.rows), kBlock, 0, stream>>>(view(), bwd.rows.p + bwd.hptr[s.l0], s.rows, z);
        else
            kIluBackwardRun<<<1, kIluRunThreads, 0, stream>>>(view(), bwd.ptr.p, bwd.rows.p, s.l0, s.l1, z);
    }
    BSIM_LAUNCH_CHECK();
}

}  // namespace bsim
