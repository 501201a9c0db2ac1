// Aggregation AMG setup on the host (once per linear system).
#include "amg_host.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace bsim {

namespace {
// Strengths are compared on a 1/4096 grid, so two assemblies of the same
// matrix that differ in the last bits build the same aggregates.
constexpr double kGrid = 4096.0;
}  // namespace

void aggregate(const HostCsr& A, double theta, std::vector<int>& agg, int& nc) {
    const int n = A.n;
    const long long thr = static_cast<long long>(std::ceil(theta * kGrid));
    std::vector<long long> q(A.val.size(), 0);
    for (int i = 0; i < n; ++i) {
        double mx = 0.0;
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) mx = std::max(mx, std::fabs(A.val[k]));
        if (!(mx > 0.0)) continue;
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
            q[k] = static_cast<long long>(std::floor(std::fabs(A.val[k]) / mx * kGrid + 0.5));
    }
    agg.assign(n, -1);
    nc = 0;
    // Pass 1: a node whose strong neighbourhood is untouched becomes a root.
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        bool free = true;
        for (int k = A.ptr[i]; k < A.ptr[i + 1] && free; ++k)
            if (q[k] >= thr && agg[A.col[k]] >= 0) free = false;
        if (!free) continue;
        agg[i] = nc;
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
            if (q[k] >= thr) agg[A.col[k]] = nc;
        ++nc;
    }
    // Pass 2: join the strongest pass-1 aggregate (ties: lowest column).
    const std::vector<int> first = agg;
    for (int i = 0; i < n; ++i) {
        if (first[i] >= 0) continue;
        long long best = -1;
        int bestCol = -1;
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
            if (q[k] < thr || first[A.col[k]] < 0) continue;
            if (q[k] > best || (q[k] == best && A.col[k] < bestCol)) { best = q[k]; bestCol = A.col[k]; }
        }
        if (bestCol >= 0) agg[i] = first[bestCol];
    }
    // Pass 3: remaining nodes group with their unassigned strong neighbours.
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        agg[i] = nc;
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
            if (q[k] >= thr && agg[A.col[k]] < 0) agg[A.col[k]] = nc;
        ++nc;
    }
}

void galerkinSymbolic(const HostCsr& A, const std::vector<int>& agg, int nc, HostCsr& Ac, std::vector<int>& gPtr,
                      std::vector<int>& gSrc, std::vector<int>& memPtr, std::vector<int>& memIdx) {
    const int n = A.n;
    memPtr.assign(nc + 1, 0);
    for (int i = 0; i < n; ++i) ++memPtr[agg[i] + 1];
    for (int I = 0; I < nc; ++I) memPtr[I + 1] += memPtr[I];
    memIdx.assign(n, 0);
    {
        std::vector<int> fill(memPtr.begin(), memPtr.end() - 1);
        for (int i = 0; i < n; ++i) memIdx[fill[agg[i]]++] = i;
    }
    Ac.n = nc;
    Ac.ptr.assign(nc + 1, 0);
    Ac.col.clear();
    std::vector<std::vector<int>> lists;  // per coarse off-diagonal entry
    std::vector<std::vector<int>> dlists(nc);
    std::vector<int> cols, pos;
    for (int I = 0; I < nc; ++I) {
        cols.clear();
        for (int m = memPtr[I]; m < memPtr[I + 1]; ++m) {
            const int i = memIdx[m];
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
                if (agg[A.col[k]] != I) cols.push_back(agg[A.col[k]]);
        }
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        const std::size_t base = lists.size();
        lists.resize(base + cols.size());
        for (int m = memPtr[I]; m < memPtr[I + 1]; ++m) {
            const int i = memIdx[m];
            dlists[I].push_back(i);
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
                const int J = agg[A.col[k]];
                if (J == I) dlists[I].push_back(n + k);
                else lists[base + (std::lower_bound(cols.begin(), cols.end(), J) - cols.begin())].push_back(n + k);
            }
        }
        for (int J : cols) Ac.col.push_back(J);
        Ac.ptr[I + 1] = static_cast<int>(Ac.col.size());
    }
    gPtr.assign(1, 0);
    gSrc.clear();
    for (int I = 0; I < nc; ++I) {
        gSrc.insert(gSrc.end(), dlists[I].begin(), dlists[I].end());
        gPtr.push_back(static_cast<int>(gSrc.size()));
    }
    for (auto& l : lists) {
        gSrc.insert(gSrc.end(), l.begin(), l.end());
        gPtr.push_back(static_cast<int>(gSrc.size()));
    }
    // Numeric values in the same order (needed to aggregate the next level).
    Ac.diag.assign(nc, 0.0);
    Ac.val.assign(Ac.col.size(), 0.0);
    for (int e = 0; e < nc + static_cast<int>(Ac.col.size()); ++e) {
        double s = 0.0;
        for (int t = gPtr[e]; t < gPtr[e + 1]; ++t) s += gSrc[t] < n ? A.diag[gSrc[t]] : A.val[gSrc[t] - n];
        if (e < nc) Ac.diag[e] = s; else Ac.val[e - nc] = s;
    }
}

std::vector<HostLevel> buildHierarchy(const HostCsr& A0, double theta, int maxLevels, int coarseSize) {
    std::vector<HostLevel> L(1);
    L[0].A = A0;
    while (static_cast<int>(L.size()) < maxLevels && L.back().A.n > coarseSize) {
        HostLevel& cur = L.back();
        int nc = 0;
        aggregate(cur.A, theta, cur.agg, nc);
        if (nc == 0 || nc >= cur.A.n) { cur.agg.clear(); break; }
        cur.nc = nc;
        HostLevel next;
        galerkinSymbolic(cur.A, cur.agg, nc, next.A, cur.gPtr, cur.gSrc, cur.memPtr, cur.memIdx);
        L.push_back(std::move(next));
    }
    if (L.back().A.n > 4096) throw std::runtime_error("AMG coarsest level is too large for the dense solve");
    return L;
}

SlicedPattern sliceRows(const HostCsr& A) {
    SlicedPattern S;
    const int n = A.n, nslices = (n + 63) / 64;
    for (int i = 0; i < n; ++i) S.K = std::max(S.K, A.ptr[i + 1] - A.ptr[i]);
    S.perm.resize(n);
    for (int i = 0; i < n; ++i) S.perm[i] = i;
    auto len = [&](int i) { return A.ptr[i + 1] - A.ptr[i]; };
    for (int s = 0; s < nslices; ++s)
        std::stable_sort(S.perm.begin() + 64 * s, S.perm.begin() + std::min(n, 64 * s + 64),
                         [&](int a, int b) { return len(a) > len(b); });
    S.sliceOff.assign(static_cast<std::size_t>(nslices) * (S.K + 1), 0);
    int base = 0;
    for (int s = 0; s < nslices; ++s) {
        int* so = &S.sliceOff[static_cast<std::size_t>(s) * (S.K + 1)];
        const int lanes = std::min(64, n - 64 * s);
        so[0] = base;
        for (int k = 0; k < S.K; ++k) {
            int width = 0;
            while (width < lanes && len(S.perm[64 * s + width]) > k) ++width;
            so[k + 1] = so[k] + width;
        }
        base = so[S.K];
    }
    S.nnz = base;
    S.col.assign(std::max(S.nnz, 1), -1);
    S.csrToStore.assign(A.ptr[n], 0);
    for (int q = 0; q < n; ++q) {
        const int i = S.perm[q];
        const int* so = &S.sliceOff[static_cast<std::size_t>(q / 64) * (S.K + 1)];
        for (int k = 0; k < len(i); ++k) {
            const int e = so[k] + q % 64;
            S.csrToStore[A.ptr[i] + k] = e;
            S.col[e] = A.col[A.ptr[i] + k];
        }
    }
    return S;
}

void remapGatherDest(std::vector<int>& gPtr, std::vector<int>& gSrc, int nc, const std::vector<int>& csrToStore) {
    const int ne = static_cast<int>(gPtr.size()) - 1;
    std::vector<int> owner(ne);  // destination -> original list
    for (int e = 0; e < nc; ++e) owner[e] = e;
    for (int k = 0; k + nc < ne; ++k) owner[nc + csrToStore[k]] = nc + k;
    std::vector<int> np(1, 0), ns;
    ns.reserve(gSrc.size());
    for (int e = 0; e < ne; ++e) {
        const int o = owner[e];
        ns.insert(ns.end(), gSrc.begin() + gPtr[o], gSrc.begin() + gPtr[o + 1]);
        np.push_back(static_cast<int>(ns.size()));
    }
    gPtr.swap(np);
    gSrc.swap(ns);
}

void remapGatherSrc(std::vector<int>& gSrc, int n, const std::vector<int>& csrToStore) {
    for (int& src : gSrc)
        if (src >= n) src = n + csrToStore[src - n];
}

}  // namespace bsim
