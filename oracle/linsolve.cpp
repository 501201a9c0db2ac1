// Oracle linear algebra: CSR assembly from LDU, the PETSc CG/BiCGSTAB
// recurrences (dependencies/petsc/src/ksp/ksp/impls/cg/cg.c KSPSolve_CG,
// bcgs/bcgs.c:14-140), the residual/reuse contract of
// src/backend/petsc_linear_system.cpp:257-481 and the build's own aggregation
// AMG.  Test infrastructure only.
#include "internal.hpp"

#include <algorithm>
#include <numeric>

namespace orc {

// Off-diagonal entries of row i: interior faces of cell i sorted by
// (neighbour, face).  Split sub-faces that couple the same pair stay separate
// entries with the same column; their products sum like PETSc's COO
// accumulation (petsc_assembly_plan.cpp:33-73).
Csr csrPattern(const Mesh& m) {
    Csr A;
    A.n = m.nc;
    A.ptr = m.csrPtr;
    A.col = m.csrCol;
    A.diag.assign(m.nc, 0.0);
    A.val.assign(A.col.size(), 0.0);
    return A;
}

void lduToCsr(const Mesh& m, const std::vector<double>& diag, const std::vector<double>& upper,
              const std::vector<double>& lower, Csr& A) {
    // Row owner holds upper[f] (column neighbour); row neighbour holds lower[f].
    A.n = m.nc;
    A.ptr = m.csrPtr;
    A.col = m.csrCol;
    A.diag = diag;
    A.val.resize(m.csrCol.size());
    const Index ne = static_cast<Index>(m.csrCol.size());
    parallelFor(ne, [&](Index e) {
        const Index f = m.csrFace[e];
        A.val[e] = m.csrOwn[e] ? upper[f] : lower[f];
    });
}

void spmv(const Csr& A, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (Index i = 0; i < A.n; ++i) {
        double s = A.diag[i] * x[i];
        for (Index k = A.ptr[i]; k < A.ptr[i + 1]; ++k) s += A.val[k] * x[A.col[k]];
        y[i] = s;
    }
}

namespace {

// Inner products are summed in a fixed blocked order: 256 strided partial sums
// per block over at most 32768 blocks, combined inside each block as four
// wavefronts of 64, then 1024 strided sums of the block sums combined as
// sixteen wavefronts of 64.  A wavefront's 64 values are added by a pairwise
// tree (lane t takes lane t + s for s = 32, 16, ..., 1: the xor butterfly's
// result) and the wavefront totals in order.  Both sides of the parity check
// use this order (it is the deterministic reduction of the device library),
// so FP64 Krylov histories can be compared bit for bit.
double wavefrontSum(double* acc, long count) {
    double total = 0.0;
    for (long w = 0; w < count / 64; ++w) {
        double* a = acc + 64 * w;
        for (long s = 32; s > 0; s >>= 1)
            for (long t = 0; t < s; ++t) a[t] += a[t + s];
        total = w == 0 ? a[0] : total + a[0];
    }
    return total;
}

double dotv(const std::vector<double>& a, const std::vector<double>& b) {
    constexpr long B = 256, Gmax = 32768;
    const long n = static_cast<long>(a.size());
    long G = (n + B - 1) / B;
    G = G < 1 ? 1 : (G > Gmax ? Gmax : G);
    std::vector<double> partial(G);
#pragma omp parallel for schedule(static)
    for (long g = 0; g < G; ++g) {
        double acc[B];
        for (long t = 0; t < B; ++t) {
            double s = 0.0;
            for (long i = g * B + t; i < n; i += G * B) s += a[i] * b[i];
            acc[t] = s;
        }
        partial[g] = wavefrontSum(acc, B);
    }
    constexpr long T = 1024;  // threads of the device's reduction tail
    double acc[T];
    for (long t = 0; t < T; ++t) {
        double s = 0.0;
        for (long g = t; g < G; g += T) s += partial[g];
        acc[t] = s;
    }
    return wavefrontSum(acc, T);
}
double nrm(const std::vector<double>& a) { return std::sqrt(dotv(a, a)); }

// Quantised relative strength: robust to last-bit differences between two
// assemblies of the same matrix, so both sides build the same aggregates.
constexpr double kStrengthGrid = 4096.0;

}  // namespace

// Greedy three-pass aggregation on the quantised strength graph.
void aggregate(const Csr& A, double theta, std::vector<Index>& agg, Index& nc) {
    const Index n = A.n;
    const long long thr = static_cast<long long>(std::ceil(theta * kStrengthGrid));
    std::vector<long long> q(A.val.size(), 0);
    for (Index i = 0; i < n; ++i) {
        double mx = 0.0;
        for (Index k = A.ptr[i]; k < A.ptr[i + 1]; ++k) mx = std::max(mx, std::abs(A.val[k]));
        if (!(mx > 0.0)) continue;
        for (Index k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
            q[k] = static_cast<long long>(std::floor(std::abs(A.val[k]) / mx * kStrengthGrid + 0.5));
    }
    auto strong = [&](Index k) { return q[k] >= thr && A.col[k] != -1; };
    agg.assign(n, -1);
    nc = 0;
    // Pass 1: roots whose strong neighbourhood is untouched.
    for (Index i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        bool free = true;
        for (Index k = A.ptr[i]; k < A.ptr[i + 1] && free; ++k)
            if (strong(k) && agg[A.col[k]] >= 0) free = false;
        if (!free) continue;
        agg[i] = nc;
        for (Index k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
            if (strong(k)) agg[A.col[k]] = nc;
        ++nc;
    }
    // Pass 2: attach to the strongest pass-1 aggregate (ties: lowest column).
    const std::vector<Index> first = agg;
    for (Index i = 0; i < n; ++i) {
        if (first[i] >= 0) continue;
        long long best = -1;
        Index bestCol = -1;
        for (Index k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
            if (!strong(k) || first[A.col[k]] < 0) continue;
            if (q[k] > best || (q[k] == best && A.col[k] < bestCol)) { best = q[k]; bestCol = A.col[k]; }
        }
        if (bestCol >= 0) agg[i] = first[bestCol];
    }
    // Pass 3: leftovers aggregate with their unassigned strong neighbours.
    for (Index i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        agg[i] = nc;
        for (Index k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
            if (strong(k) && agg[A.col[k]] < 0) agg[A.col[k]] = nc;
        ++nc;
    }
}

// Galerkin product with piecewise-constant prolongation.  Each coarse value is
// the sum of its fine contributions in (fine row, diagonal first, entry) order.
void galerkin(const Csr& A, const std::vector<Index>& agg, Index nc, Csr& Ac) {
    std::vector<std::vector<Index>> members(nc);
    for (Index i = 0; i < A.n; ++i) members[agg[i]].push_back(i);
    Ac.n = nc;
    Ac.ptr.assign(nc + 1, 0);
    Ac.col.clear(); Ac.val.clear();
    Ac.diag.assign(nc, 0.0);
    // Coarse rows are independent: each is formed on its own thread, then the
    // rows are concatenated in order.
    std::vector<std::vector<Index>> rowCols(nc);
    std::vector<std::vector<double>> rowVals(nc);
    parallelFor(nc, [&](Index I) {
        std::vector<Index> cols;
        for (Index i : members[I])
            for (Index k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
                if (agg[A.col[k]] != I) cols.push_back(agg[A.col[k]]);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        std::vector<double> vals(cols.size(), 0.0);
        double d = 0.0;
        for (Index i : members[I]) {
            d += A.diag[i];
            for (Index k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
                const Index J = agg[A.col[k]];
                if (J == I) d += A.val[k];
                else vals[std::lower_bound(cols.begin(), cols.end(), J) - cols.begin()] += A.val[k];
            }
        }
        Ac.diag[I] = d;
        rowCols[I] = std::move(cols);
        rowVals[I] = std::move(vals);
    });
    for (Index I = 0; I < nc; ++I) {
        Ac.col.insert(Ac.col.end(), rowCols[I].begin(), rowCols[I].end());
        Ac.val.insert(Ac.val.end(), rowVals[I].begin(), rowVals[I].end());
        Ac.ptr[I + 1] = static_cast<Index>(Ac.col.size());
    }
}

void Amg::setup(const Csr& A, const SolverConfig& c) {
    // The first setup aggregates every level; later setups (preconditioner
    // rebuilds) keep the aggregates and recompute the Galerkin values.
    cfg = c;
    if (!aggregated) {
        levels.clear();
        levels.push_back(AmgLevel{A, {}, 0, {}, {}});
        while (static_cast<int>(levels.size()) < cfg.amgMaxLevels && levels.back().A.n > cfg.amgCoarseSize) {
            AmgLevel& L = levels.back();
            Index nc = 0;
            aggregate(L.A, cfg.amgStrength, L.agg, nc);
            if (nc == 0 || nc >= L.A.n) { L.agg.clear(); break; }
            L.nc = nc;
            L.memPtr.assign(nc + 1, 0);
            for (Index i = 0; i < L.A.n; ++i) ++L.memPtr[L.agg[i] + 1];
            for (Index I = 0; I < nc; ++I) L.memPtr[I + 1] += L.memPtr[I];
            L.memIdx.resize(L.A.n);
            {
                std::vector<Index> fill(L.memPtr.begin(), L.memPtr.end() - 1);
                for (Index i = 0; i < L.A.n; ++i) L.memIdx[fill[L.agg[i]]++] = i;
            }
            Csr Ac;
            galerkin(L.A, L.agg, nc, Ac);
            levels.push_back(AmgLevel{std::move(Ac), {}, 0, {}, {}});
        }
        aggregated = true;
    } else {
        levels[0].A = A;
        for (std::size_t l = 0; l + 1 < levels.size(); ++l)
            galerkin(levels[l].A, levels[l].agg, levels[l].nc, levels[l + 1].A);
    }
    const Csr& Cc = levels.back().A;
    const Index n = Cc.n;
    if (n > 4096) throw std::runtime_error("AMG coarsest level is too large for the dense solve");
    coarseLU.assign(static_cast<std::size_t>(n) * n, 0.0);
    for (Index i = 0; i < n; ++i) {
        coarseLU[static_cast<std::size_t>(i) * n + i] = Cc.diag[i];
        for (Index k = Cc.ptr[i]; k < Cc.ptr[i + 1]; ++k) coarseLU[static_cast<std::size_t>(i) * n + Cc.col[k]] += Cc.val[k];
    }
    // LU with partial pivoting, row-major, in place.
    coarsePiv.resize(n);
    for (Index k = 0; k < n; ++k) {
        Index p = k;
        for (Index i = k + 1; i < n; ++i)
            if (std::abs(coarseLU[static_cast<std::size_t>(i) * n + k]) > std::abs(coarseLU[static_cast<std::size_t>(p) * n + k])) p = i;
        coarsePiv[k] = p;
        if (p != k)
            for (Index j = 0; j < n; ++j) std::swap(coarseLU[static_cast<std::size_t>(k) * n + j], coarseLU[static_cast<std::size_t>(p) * n + j]);
        const double piv = coarseLU[static_cast<std::size_t>(k) * n + k];
        if (!(std::abs(piv) > 0.0)) throw std::runtime_error("AMG coarsest matrix is singular");
        for (Index i = k + 1; i < n; ++i) {
            const double l = coarseLU[static_cast<std::size_t>(i) * n + k] / piv;
            coarseLU[static_cast<std::size_t>(i) * n + k] = l;
            for (Index j = k + 1; j < n; ++j) coarseLU[static_cast<std::size_t>(i) * n + j] -= l * coarseLU[static_cast<std::size_t>(k) * n + j];
        }
    }
    // Explicit inverse: column j is the LU solve of e_j.
    coarseInv.assign(static_cast<std::size_t>(n) * n, 0.0);
    std::vector<double> w(n);
    for (Index j = 0; j < n; ++j) {
        for (Index i = 0; i < n; ++i) w[i] = i == j ? 1.0 : 0.0;
        for (Index k = 0; k < n; ++k) std::swap(w[k], w[coarsePiv[k]]);
        for (Index i = 0; i < n; ++i)
            for (Index q = 0; q < i; ++q) w[i] -= coarseLU[static_cast<std::size_t>(i) * n + q] * w[q];
        for (Index i = n - 1; i >= 0; --i) {
            for (Index q = i + 1; q < n; ++q) w[i] -= coarseLU[static_cast<std::size_t>(i) * n + q] * w[q];
            w[i] /= coarseLU[static_cast<std::size_t>(i) * n + i];
        }
        for (Index i = 0; i < n; ++i) coarseInv[static_cast<std::size_t>(i) * n + j] = w[i];
    }
    wx.assign(levels.size(), {}); wb.assign(levels.size(), {}); wr.assign(levels.size(), {});
    for (std::size_t l = 0; l < levels.size(); ++l) {
        wx[l].assign(levels[l].A.n, 0.0); wb[l].assign(levels[l].A.n, 0.0); wr[l].assign(levels[l].A.n, 0.0);
    }
}

void Amg::vcycle(int l) {
    const Csr& A = levels[l].A;
    std::vector<double>& x = wx[l];
    std::vector<double>& b = wb[l];
    std::vector<double>& r = wr[l];
    const Index n = A.n;
    if (l == static_cast<int>(levels.size()) - 1) {  // dense solve with the explicit inverse
        for (Index i = 0; i < n; ++i) {
            double s = 0.0;
            for (Index j = 0; j < n; ++j) s += coarseInv[static_cast<std::size_t>(i) * n + j] * b[j];
            x[i] = s;
        }
        return;
    }
    const double w = cfg.amgOmega;
    auto sweep = [&]() {  // x <- x + w (b - A x) / d
        spmv(A, x.data(), r.data());
#pragma omp parallel for schedule(static)
        for (Index i = 0; i < n; ++i) r[i] = x[i] + w * (b[i] - r[i]) / A.diag[i];
        x.swap(r);
    };
#pragma omp parallel for schedule(static)
    for (Index i = 0; i < n; ++i) x[i] = w * b[i] / A.diag[i];
    for (int s = 1; s < cfg.amgPreSweeps; ++s) sweep();
    spmv(A, x.data(), r.data());
#pragma omp parallel for schedule(static)
    for (Index i = 0; i < n; ++i) r[i] = b[i] - r[i];
    std::vector<double>& bc = wb[l + 1];
    // Restriction as a gather over each aggregate's members in ascending row
    // order: the sums of a serial scatter over the rows, term for term.
    const AmgLevel& lv = levels[l];
    parallelFor(lv.nc, [&](Index I) {
        double s = 0.0;
        for (Index k = lv.memPtr[I]; k < lv.memPtr[I + 1]; ++k) s += r[lv.memIdx[k]];
        bc[I] = s;
    });
    vcycle(l + 1);
    const std::vector<double>& xc = wx[l + 1];
#pragma omp parallel for schedule(static)
    for (Index i = 0; i < n; ++i) x[i] += cfg.amgCorrection * xc[levels[l].agg[i]];
    for (int s = 0; s < cfg.amgPostSweeps; ++s) sweep();
}

void Amg::apply(const double* r, double* z) {
    std::copy(r, r + levels[0].A.n, wb[0].begin());
    vcycle(0);
    std::copy(wx[0].begin(), wx[0].end(), z);
}

LinearSystem::~LinearSystem() { delete amg; }

void LinearSystem::update(const std::vector<double>& diag, const std::vector<double>& val) {
    // petsc_linear_system.cpp:257-282: exact comparison decides a value update.
    const bool same = ready && diag == prevDiag && val == prevVal;
    if (!same) {
        prevDiag = diag; prevVal = val;
        A.diag = diag; A.val = val;
        ready = true; setupPending = true; ++matrixUpdates;
    } else {
        ++rhsOnly;
    }
}

static bool sameConfig(const SolverConfig& a, const SolverConfig& b) {
    return a.solver == b.solver && a.pc == b.pc && a.atol == b.atol && a.rtol == b.rtol && a.maxIter == b.maxIter &&
           a.warmStart == b.warmStart && a.rebuildInterval == b.rebuildInterval && a.reuseGrowth == b.reuseGrowth &&
           a.amgMaxLevels == b.amgMaxLevels && a.amgCoarseSize == b.amgCoarseSize &&
           a.amgPreSweeps == b.amgPreSweeps && a.amgPostSweeps == b.amgPostSweeps && a.amgOmega == b.amgOmega &&
           a.amgStrength == b.amgStrength && a.amgCorrection == b.amgCorrection;
}

// petsc_linear_system.cpp:284-301
void LinearSystem::prepare(const SolverConfig& c) {
    if (!configured || !sameConfig(c, cfg)) {
        cfg = c; configured = true; setupPending = true; forceRebuild = true;
    }
    if (!setupPending && !forceRebuild) return;
    const bool reuse = cfg.pc != 0 && pcSetups != 0 && !forceRebuild &&
                       matrixUpdates - pcMatrixUpdate < static_cast<std::uint64_t>(cfg.rebuildInterval) &&
                       pcRhsSolves < static_cast<std::uint64_t>(cfg.rebuildInterval);
    setup(reuse);
}

// Block Jacobi with ILU(0) in each block of 64 consecutive rows: PETSc's
// bjacobi (petsc_backend.cpp:46) with ceil(n/64) blocks.  The factorisation
// runs k = 0..62 over the rows r > k with (r, k) in the pattern: l = a_rk / a_kk,
// then a_rj -= l * a_kj for each j > k present in rows r and k (no fill).  A
// repeated column keeps the factor in its first entry and 0 in the others.
// The substitutions take the columns in order and skip zero factors.  This is
// the order of the device library's kernels, so the factors and every
// preconditioned vector agree bit for bit.
namespace {
constexpr Index kIlu = 64;
}

void blockIluFactor(const Csr& A, std::vector<double>& fd, std::vector<double>& fv) {
    const Index n = A.n, nb = (n + kIlu - 1) / kIlu;
    fd.assign(n, 0.0);
    fv.assign(A.val.size(), 0.0);
#pragma omp parallel for schedule(static)
    for (Index b = 0; b < nb; ++b) {
        const Index base = b * kIlu;
        std::vector<double> T(kIlu * kIlu, 0.0);
        std::uint64_t mask[kIlu];
        for (Index r = 0; r < kIlu; ++r) {
            const Index i = base + r;
            mask[r] = 1ull << r;
            T[r * kIlu + r] = i < n ? A.diag[i] : 1.0;
            if (i >= n) continue;
            for (Index e = A.ptr[i]; e < A.ptr[i + 1]; ++e) {
                const Index c = A.col[e] - base;
                if (c >= 0 && c < kIlu) { T[r * kIlu + c] += A.val[e]; mask[r] |= 1ull << c; }
            }
        }
        for (Index k = 0; k < kIlu - 1; ++k)
            for (Index r = k + 1; r < kIlu; ++r) {
                if (!((mask[r] >> k) & 1ull)) continue;
                const double l = T[r * kIlu + k] / T[k * kIlu + k];
                T[r * kIlu + k] = l;
                for (Index j = k + 1; j < kIlu; ++j)
                    if (((mask[r] & mask[k]) >> j) & 1ull) T[r * kIlu + j] -= l * T[k * kIlu + j];
            }
        for (Index r = 0; r < kIlu && base + r < n; ++r) {
            const Index i = base + r;
            fd[i] = T[r * kIlu + r];
            std::uint64_t written = 1ull << r;
            for (Index e = A.ptr[i]; e < A.ptr[i + 1]; ++e) {
                const Index c = A.col[e] - base;
                if (c < 0 || c >= kIlu) continue;
                fv[e] = ((written >> c) & 1ull) ? 0.0 : T[r * kIlu + c];
                written |= 1ull << c;
            }
        }
    }
}

void blockIluApply(const Csr& A, const std::vector<double>& fd, const std::vector<double>& fv, const double* r,
                   double* z) {
    const Index n = A.n, nb = (n + kIlu - 1) / kIlu;
#pragma omp parallel for schedule(static)
    for (Index b = 0; b < nb; ++b) {
        const Index base = b * kIlu;
        std::vector<double> T(kIlu * kIlu, 0.0);
        double y[kIlu], d[kIlu];
        for (Index q = 0; q < kIlu; ++q) {
            const Index i = base + q;
            y[q] = i < n ? r[i] : 0.0;
            d[q] = i < n ? fd[i] : 1.0;
            if (i >= n) continue;
            for (Index e = A.ptr[i]; e < A.ptr[i + 1]; ++e) {
                const Index c = A.col[e] - base;
                if (c >= 0 && c < kIlu) T[q * kIlu + c] += fv[e];
            }
        }
        for (Index s = 0; s < kIlu - 1; ++s)
            for (Index q = s + 1; q < kIlu; ++q) {
                const double t = T[q * kIlu + s];
                if (t != 0.0) y[q] = y[q] - t * y[s];
            }
        for (Index s = kIlu - 1; s >= 0; --s) {
            y[s] = y[s] / d[s];
            for (Index q = 0; q < s; ++q) {
                const double t = T[q * kIlu + s];
                if (t != 0.0) y[q] = y[q] - t * y[s];
            }
        }
        for (Index q = 0; q < kIlu && base + q < n; ++q) z[base + q] = y[q];
    }
}

// IC(0) / ILU(0) of the whole matrix (PETSc's icc, petsc_linear_system.cpp:
// 204-208, factors the upper triangle; asm on one rank, :224-226, is one
// subdomain with the sub-PC's default ILU(0)), natural row order, no fill.
// Factor entry (i, j) sums its sources: the diagonal first, then the stored
// entries in row-major order; for IC(0) a lower entry (i, k) takes the upper
// entry (k, i).  Row i is eliminated in IKJ order: for each lower (i, k) in
// column order, l = a_ik / u_kk, then a_ij -= l u_kj for the columns j > k
// present in both rows.  A pivot that is zero (ILU), not positive (IC) or not
// finite fails the factorisation, after which the preconditioner returns +Inf
// in every entry (MatSolve on a failed factor).
void WholeIlu::analyse(const Csr& A, bool cholesky) {
    chol = cholesky;
    n = A.n;
    std::vector<std::vector<Index>> rows(n);
    for (Index i = 0; i < n; ++i) {
        rows[i].push_back(i);
        for (Index q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
            const Index j = A.col[q];
            if (j == i) continue;
            if (!chol) rows[i].push_back(j);
            else if (j > i) { rows[i].push_back(j); rows[j].push_back(i); }
        }
    }
    ptr.assign(n + 1, 0);
    col.clear();
    for (Index i = 0; i < n; ++i) {
        std::sort(rows[i].begin(), rows[i].end());
        rows[i].erase(std::unique(rows[i].begin(), rows[i].end()), rows[i].end());
        col.insert(col.end(), rows[i].begin(), rows[i].end());
        ptr[i + 1] = static_cast<Index>(col.size());
    }
    auto at = [&](Index i, Index j) {
        for (Index e = ptr[i]; e < ptr[i + 1]; ++e)
            if (col[e] == j) return e;
        return Index(-1);
    };
    dpos.assign(n, 0);
    for (Index i = 0; i < n; ++i) dpos[i] = at(i, i);
    const Index nf = ptr[n];
    std::vector<std::vector<Index>> sources(nf);
    for (Index i = 0; i < n; ++i) sources[dpos[i]].push_back(-(i + 1));
    for (Index i = 0; i < n; ++i)
        for (Index q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
            const Index j = A.col[q];
            if (j == i) sources[dpos[i]].push_back(q);
            else if (!chol) sources[at(i, j)].push_back(q);
            else if (j > i) { sources[at(i, j)].push_back(q); sources[at(j, i)].push_back(q); }
        }
    srcPtr.assign(nf + 1, 0);
    src.clear();
    for (Index t = 0; t < nf; ++t) {
        src.insert(src.end(), sources[t].begin(), sources[t].end());
        srcPtr[t + 1] = static_cast<Index>(src.size());
    }
    uptr.assign(nf + 1, 0);
    udst.clear();
    usrc.clear();
    for (Index i = 0; i < n; ++i)
        for (Index e = ptr[i]; e < ptr[i + 1]; ++e) {
            if (e < dpos[i]) {
                const Index k = col[e];
                for (Index a = e + 1; a < ptr[i + 1]; ++a) {
                    const Index b = at(k, col[a]);
                    if (b >= 0 && b > dpos[k]) { udst.push_back(a); usrc.push_back(b); }
                }
            }
            uptr[e + 1] = static_cast<Index>(udst.size());
        }
    f.assign(nf, 0.0);
    analysed = true;
}

void WholeIlu::factor(const Csr& A) {
    for (Index t = 0; t + 1 < static_cast<Index>(srcPtr.size()); ++t) {
        double s = 0.0;
        for (Index q = srcPtr[t]; q < srcPtr[t + 1]; ++q) {
            const Index p = src[q];
            const double v = p >= 0 ? A.val[p] : A.diag[-p - 1];
            s = q == srcPtr[t] ? v : s + v;
        }
        f[t] = s;
    }
    fail = false;
    for (Index i = 0; i < n; ++i) {
        for (Index e = ptr[i]; e < dpos[i]; ++e) {
            const double l = f[e] / f[dpos[col[e]]];
            f[e] = l;
            for (Index q = uptr[e]; q < uptr[e + 1]; ++q) f[udst[q]] -= l * f[usrc[q]];
        }
        const double p = f[dpos[i]];
        if (!(std::isfinite(p) && (chol ? p > 0.0 : p != 0.0))) fail = true;
    }
}

void WholeIlu::apply(const double* r, double* z) const {
    if (fail) {
        for (Index i = 0; i < n; ++i) z[i] = HUGE_VAL;
        return;
    }
    for (Index i = 0; i < n; ++i) {
        double y = r[i];
        for (Index e = ptr[i]; e < dpos[i]; ++e) y -= f[e] * z[col[e]];
        z[i] = y;
    }
    for (Index i = n - 1; i >= 0; --i) {
        double y = z[i];
        for (Index e = dpos[i] + 1; e < ptr[i + 1]; ++e) y -= f[e] * z[col[e]];
        z[i] = y / f[dpos[i]];
    }
}

// petsc_linear_system.cpp:303-324; the PC keeps the matrix it was built from.
static std::vector<double> jacobiInv;
void LinearSystem::setup(bool reuse) {
    if (!reuse) {
        if (cfg.pc == 2) {
            if (!amg) amg = new Amg();
            amg->setup(A, cfg);
        } else if (cfg.pc == 1) {
            if (!amg) amg = new Amg();
            amg->levels.assign(1, AmgLevel{A, {}, 0, {}, {}});
        } else if (cfg.pc == 3) {
            if (!amg) amg = new Amg();
            amg->levels.assign(1, AmgLevel{A, {}, 0, {}, {}});
            blockIluFactor(A, amg->iluDiag, amg->iluVal);
        } else if (cfg.pc == 4 || cfg.pc == 5) {
            if (!amg) amg = new Amg();
            // The analysis depends on the pattern and the kind only.
            if (!amg->whole.analysed || amg->whole.chol != (cfg.pc == 4)) amg->whole.analyse(A, cfg.pc == 4);
            amg->whole.factor(A);
        }
        if (cfg.pc != 0) ++pcSetups;
        pcMatrixUpdate = matrixUpdates;
        pcRhsSolves = 0;
        freshIterations[0] = freshIterations[1] = freshIterations[2] = 0;
    }
    setupPending = false;
    forceRebuild = false;
}

namespace {

struct Ksp {
    const Csr& A;
    const SolverConfig& cfg;
    Amg* pc;
    double target;
    std::vector<double>* hist;
    int its = 0;
    int reason = 0;  // >0 converged, <0 diverged (-3 = max iterations)
    void applyPC(const std::vector<double>& r, std::vector<double>& z) {
        const Index n = A.n;
        if (cfg.pc == 0) { z = r; return; }
        if (cfg.pc == 1) {
            const Csr& P = pc->levels[0].A;
#pragma omp parallel for schedule(static)
            for (Index i = 0; i < n; ++i) z[i] = r[i] * (1.0 / P.diag[i]);
            return;
        }
        if (cfg.pc == 3) {
            blockIluApply(pc->levels[0].A, pc->iluDiag, pc->iluVal, r.data(), z.data());
            return;
        }
        if (cfg.pc >= 4) {
            pc->whole.apply(r.data(), z.data());
            return;
        }
        pc->apply(r.data(), z.data());
    }
    bool test(double dp) {  // petsc_linear_system.cpp:164-175
        if (!std::isfinite(dp)) { reason = -9; return true; }
        if (dp <= target) { reason = 2; return true; }
        return false;
    }
    void cg(const std::vector<double>& b, std::vector<double>& x, bool guessZero, int maxIt) {
        const Index n = A.n;
        std::vector<double> r(n), z(n), p(n), w(n);
        if (!guessZero) { spmv(A, x.data(), r.data()); for (Index i = 0; i < n; ++i) r[i] = b[i] - r[i]; }
        else r = b;
        double dp = nrm(r);
        its = 0;
        if (hist) hist->push_back(dp);
        if (test(dp)) return;
        applyPC(r, z);
        double beta = dotv(z, r), betaold = 1.0, dpi = 0.0, dpiold;
        if (!std::isfinite(beta)) { reason = -9; return; }
        int i = 0;
        do {
            its = i + 1;
            if (beta == 0.0) { reason = 3; break; }
            if (i > 0 && beta * betaold < 0.0) { reason = -8; break; }
            if (i == 0) p = z;
            else {
                const double bb = beta / betaold;
#pragma omp parallel for schedule(static)
                for (Index k = 0; k < n; ++k) p[k] = z[k] + bb * p[k];
            }
            dpiold = dpi;
            spmv(A, p.data(), w.data());
            dpi = dotv(p, w);
            if (!std::isfinite(dpi)) { reason = -9; break; }
            betaold = beta;
            if (dpi == 0.0 || (i > 0 && ((dpi > 0) - (dpi < 0)) * ((dpiold > 0) - (dpiold < 0)) < 0)) { reason = -10; break; }
            const double a = beta / dpi;
#pragma omp parallel for schedule(static)
            for (Index k = 0; k < n; ++k) x[k] += a * p[k];
#pragma omp parallel for schedule(static)
            for (Index k = 0; k < n; ++k) r[k] += -a * w[k];
            dp = nrm(r);
            if (hist) hist->push_back(dp);
            if (test(dp)) break;
            applyPC(r, z);
            beta = dotv(z, r);
            if (!std::isfinite(beta)) { reason = -9; break; }
            ++i;
        } while (i < maxIt);
        if (i >= maxIt) reason = -3;
    }
    void bcgs(const std::vector<double>& b, std::vector<double>& x, bool guessZero, int maxIt) {
        const Index n = A.n;
        std::vector<double> r(n), rp(n), v(n, 0.0), t(n), s(n), p(n, 0.0), tmp(n), guess;
        if (!guessZero) {
            spmv(A, x.data(), r.data());
            for (Index i = 0; i < n; ++i) r[i] = b[i] - r[i];
            guess = x;
            std::fill(x.begin(), x.end(), 0.0);
        } else {
            r = b;
        }
        double dp = nrm(r);
        its = 0;
        if (hist) hist->push_back(dp);
        auto finish = [&]() {
            applyPC(x, tmp);  // KSPUnwindPreconditioner for right preconditioning
            x = tmp;
            if (!guess.empty()) for (Index i = 0; i < n; ++i) x[i] += guess[i];
        };
        if (test(dp)) {
            if (!guess.empty()) for (Index i = 0; i < n; ++i) x[i] += guess[i];
            return;
        }
        rp = r;
        double rhoold = 1.0, alpha = 1.0, omegaold = 1.0, rho, omega, d1, d2;
        int i = 0;
        do {
            rho = dotv(r, rp);
            const double beta = (rho / rhoold) * (alpha / omegaold);
#pragma omp parallel for schedule(static)
            for (Index k = 0; k < n; ++k) p[k] = r[k] + (-omegaold * beta) * v[k] + beta * p[k];
            applyPC(p, tmp); spmv(A, tmp.data(), v.data());
            d1 = dotv(v, rp);
            if (!std::isfinite(d1)) { reason = -9; break; }
            if (d1 == 0.0) { reason = -5; break; }
            alpha = rho / d1;
#pragma omp parallel for schedule(static)
            for (Index k = 0; k < n; ++k) s[k] = -alpha * v[k] + r[k];
            applyPC(s, r); spmv(A, r.data(), t.data());
            d1 = dotv(s, t); d2 = dotv(t, t);
            if (d2 == 0.0) {
                d1 = dotv(s, s);
                if (d1 != 0.0) { reason = -5; break; }
                for (Index k = 0; k < n; ++k) x[k] += alpha * p[k];
                ++its;
                reason = 2;
                if (hist) hist->push_back(dp);
                break;
            }
            omega = d1 / d2;
#pragma omp parallel for schedule(static)
            for (Index k = 0; k < n; ++k) x[k] = alpha * p[k] + omega * s[k] + x[k];
#pragma omp parallel for schedule(static)
            for (Index k = 0; k < n; ++k) r[k] = -omega * t[k] + s[k];
            dp = nrm(r);
            rhoold = rho; omegaold = omega;
            ++its;
            if (hist) hist->push_back(dp);
            if (test(dp)) break;
            if (rho == 0.0) { reason = -5; break; }
            ++i;
        } while (i < maxIt);
        if (i >= maxIt) reason = -3;
        finish();
    }
    // KSPSolve_GMRES (gmres.c:88-235, KSPGMRESUpdateHessenberg, KSPGMRESBuildSoln)
    // and KSPSolve_FGMRES (fgmres/fgmres.c:56-300) as the reference sets them
    // up: KSP_NORM_UNPRECONDITIONED (petsc_linear_system.cpp:181) makes PETSc
    // precondition GMRES from the right, so both test the Arnoldi estimate of
    // ||b - A x||.  Restart 30, classical Gram-Schmidt without refinement
    // (borthog2.c), happy-breakdown tolerance 1e-30, and for GMRES the restart
    // breakdown test with tolerance 0.1.  Vector arithmetic in a fixed order:
    // the Gram-Schmidt update v += (-h_j) V_j and the combinations sum_j c_j V_j
    // take j in order from the left.
    void gmres(const std::vector<double>& b, std::vector<double>& x, bool guessZero, int maxIt, bool flexible) {
        const Index n = A.n;
        constexpr int m = 30;
        constexpr double hapTol = 1e-30, breakdownTol = 0.1;
        std::vector<std::vector<double>> V(m + 1, std::vector<double>(n)), Z(flexible ? m : 0, std::vector<double>(n));
        std::vector<double> tmp(n), temp(n), h(m + 1);
        std::vector<double> hh(static_cast<std::size_t>(m + 1) * (m + 1), 0.0), cc(m + 1, 0.0), ss(m + 1, 0.0),
            grs(m + 1, 0.0), nrs(m + 1, 0.0);
        auto H = [&](int i, int j) -> double& { return hh[static_cast<std::size_t>(i) * (m + 1) + j]; };
        auto scale = [&](std::vector<double>& v, double a) {
#pragma omp parallel for schedule(static)
            for (Index i = 0; i < n; ++i) v[i] *= a;
        };
        its = 0;
        reason = 0;
        double rnorm = -1.0, rnorm0 = 0.0;
        bool zero = guessZero;
        while (true) {
            if (zero) V[0] = b;
            else { spmv(A, x.data(), V[0].data()); for (Index i = 0; i < n; ++i) V[0][i] = b[i] - V[0][i]; }
            double res = nrm(V[0]);
            zero = false;
            if (!std::isfinite(res)) { reason = -9; return; }
            if (!flexible) {
                if (res != 0.0) scale(V[0], 1.0 / res);
                if (rnorm > 0.0 && std::fabs(res - rnorm) > breakdownTol * rnorm0) { reason = -5; return; }
            }
            grs[0] = rnorm0 = res;
            rnorm = res;
            if (hist) hist->push_back(res);
            if (!flexible && res == 0.0) { reason = 3; return; }
            test(res);
            int it = 0;
            bool hapend = false, built = true;
            if (!reason && flexible) scale(V[0], 1.0 / res);
            while (!reason && it < m && its < maxIt) {
                std::vector<double>& w = V[it + 1];
                if (flexible) { applyPC(V[it], Z[it]); spmv(A, Z[it].data(), w.data()); }
                else { applyPC(V[it], tmp); spmv(A, tmp.data(), w.data()); }
                bool dotFailed = false;
                for (int j = 0; j <= it; ++j) {
                    h[j] = dotv(w, V[j]);
                    if (!std::isfinite(h[j])) dotFailed = true;
                }
#pragma omp parallel for schedule(static)
                for (Index i = 0; i < n; ++i) {
                    double sum = w[i];
                    for (int j = 0; j <= it; ++j) sum += -h[j] * V[j][i];
                    w[i] = sum;
                }
                const double tt = nrm(w);
                if (dotFailed) { reason = -9; break; }
                for (int j = 0; j <= it; ++j) H(j, it) = 0.0 - (-h[j]);
                if (!std::isfinite(tt)) { reason = -9; built = false; break; }
                H(it + 1, it) = tt;
                const double hapbnd = std::min(std::fabs(tt / grs[it]), hapTol);
                if (flexible) {
                    if (tt > hapbnd) scale(w, 1.0 / tt);
                    else hapend = true;
                } else {
                    if (tt != 0.0) scale(w, 1.0 / tt);
                    if (tt < hapbnd) hapend = true;
                }
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
                    reason = -2;
                    if (flexible) break;
                }
                ++it;
                ++its;
                rnorm = res;
                if (reason) break;
                if (hist) hist->push_back(res);
                test(res);
                if (hapend && !reason) { reason = -5; break; }
            }
            if (built && it > 0) {
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
                    const std::vector<std::vector<double>>& B = flexible ? Z : V;
#pragma omp parallel for schedule(static)
                    for (Index i = 0; i < n; ++i) {
                        double sum = 0.0;
                        for (int j = 0; j < it; ++j) sum += nrs[j] * B[j][i];
                        temp[i] = sum;
                    }
                    if (!flexible) { applyPC(temp, tmp); temp.swap(tmp); }  // KSPUnwindPreconditioner
#pragma omp parallel for schedule(static)
                    for (Index i = 0; i < n; ++i) x[i] += 1.0 * temp[i];
                }
            }
            if (!built) return;
            if (reason == 0 && its >= maxIt) reason = -3;
            if (reason) return;
        }
    }
};

}  // namespace

// petsc_linear_system.cpp:355-481 with the contract of solver_control.h:196-210.
SolveResult LinearSystem::solve(const std::vector<double>& b, std::vector<double>& x, int rhsIndex) {
    const Index n = A.n;
    const std::vector<double> guess = x;
    auto resetGuess = [&]() { if (cfg.warmStart) x = guess; else std::fill(x.begin(), x.end(), 0.0); };
    resetGuess();
    hist.clear();
    const double rhsNorm = nrm(b);
    auto trueResidual = [&]() {
        std::vector<double> r(n);
        spmv(A, x.data(), r.data());
        for (Index i = 0; i < n; ++i) r[i] = b[i] - r[i];
        return nrm(r);
    };
    double initial = rhsNorm;
    if (cfg.warmStart) initial = trueResidual();
    const double scale = std::max(initial, rhsNorm) > 0.0 ? std::max(initial, rhsNorm) : 1.0;
    const double target = std::max(cfg.atol, cfg.rtol * scale);
    const bool reused = cfg.pc != 0 && pcMatrixUpdate < matrixUpdates;
    ++solves;
    if (reused) ++pcReuses;
    SolveResult res;
    auto record = [&]() {
        lastStatus = res.status; lastInitial = res.initial; lastFinal = res.final_; lastRelative = res.relative;
    };
    if (std::isfinite(initial) && initial <= target) {
        res.status = Converged; res.iterations = 0; res.initial = initial; res.final_ = initial;
        res.relative = initial / scale;
        ++trueResidualChecks;
        record();
        return res;
    }
    int total = 0;
    bool pcReused = reused;
    for (int attempt = 0; attempt < 2; ++attempt) {
        int attemptIts = 0, reason = 0;
        double fin = 0.0, internalTarget = target;
        bool guessZero = !cfg.warmStart;
        for (int restart = 0; restart < 3; ++restart) {
            Ksp k{A, cfg, amg, internalTarget, &hist};
            if (cfg.solver == 0) k.cg(b, x, guessZero, cfg.maxIter - attemptIts);
            else if (cfg.solver == 1) k.bcgs(b, x, guessZero, cfg.maxIter - attemptIts);
            else k.gmres(b, x, guessZero, cfg.maxIter - attemptIts, cfg.solver == 3);
            attemptIts += k.its; total += k.its; reason = k.reason;
            fin = trueResidual();
            ++trueResidualChecks;
            if ((std::isfinite(fin) && fin <= target) || reason <= 0 || !std::isfinite(fin) || attemptIts >= cfg.maxIter) break;
            internalTarget *= 0.5;
            guessZero = false;
        }
        int status = NumericalFailure;
        if (std::isfinite(fin) && fin <= target) status = Converged;
        else if (reason == -3 || attemptIts >= cfg.maxIter) status = MaxIterations;
        res.status = status; res.iterations = total; res.initial = initial; res.final_ = fin; res.relative = fin / scale;
        if (status != Converged && attempt == 0 && pcReused) {
            setup(false);
            ++retries;
            resetGuess();
            pcReused = false;
            continue;
        }
        if (status == Converged) {
            int& fresh = freshIterations[rhsIndex];
            if (pcMatrixUpdate == matrixUpdates) fresh = std::max(fresh, attemptIts);
            else if (attemptIts > cfg.reuseGrowth * static_cast<double>(std::max(fresh, 1))) forceRebuild = true;
        }
        break;
    }
    if (total > 0) ++pcRhsSolves;
    krylovIterations += total;
    record();
    return res;
}

}  // namespace orc
