// Mesh setup on the GPU: polygon triangulation, face area vectors and centres,
// tetrahedral cell volumes and centroids, and the per-face FVM factors of
// BabelSim's Mesh::polyhedral (src/core/mesh.cpp:321-578), one thread per face
// or per cell.  Runs once per case; its outputs are the inputs of every
// face-loop kernel.
#include "dmesh.hpp"

#include <algorithm>
#include <chrono>
#include <numeric>

namespace bsim {
namespace {

constexpr int kMaxRing = 32;

enum GeomError : int { kErrDegenerate = 1, kErrOrientation = 2, kErrVolume = 4, kErrClosure = 8, kErrRing = 16 };

struct P2 {
    double x, y;
};
__device__ inline P2 project(D3 p, int dropped) {
    if (dropped == 0) return {p.y, p.z};
    if (dropped == 1) return {p.z, p.x};
    return {p.x, p.y};
}
__device__ inline double cross2(P2 a, P2 b, P2 c) { return (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x); }

// Triangulates ring positions 0..n-1: fan from 0 when convex, ear clipping
// otherwise (the rule of mesh.cpp:62-149).  Writes 3*(n-2) local indices.
__device__ bool triangulate(const D3* p, int n, unsigned char* tri) {
    D3 nrm = d3(0, 0, 0);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < n; ++i) {
        nrm = nrm + 0.5 * cross(p[i], p[(i + 1) % n]);
        const double c[3] = {p[i].x, p[i].y, p[i].z};
        for (int d = 0; d < 3; ++d) { lo[d] = fmin(lo[d], c[d]); hi[d] = fmax(hi[d], c[d]); }
    }
    const double ax = fabs(nrm.x), ay = fabs(nrm.y), az = fabs(nrm.z);
    const int dropped = ay > ax ? (az > ay ? 2 : 1) : (az > ax ? 2 : 0);
    const double scale = fmax(fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]), 1.0);
    const double tol = 1e-13 * scale * scale;
    P2 pr[kMaxRing];
    for (int i = 0; i < n; ++i) pr[i] = project(p[i], dropped);
    double sa = 0.0;
    for (int i = 0; i < n; ++i) sa += pr[i].x * pr[(i + 1) % n].y - pr[(i + 1) % n].x * pr[i].y;
    if (fabs(sa) <= tol) return false;
    const double o = sa > 0.0 ? 1.0 : -1.0;
    bool convex = true, turned = false;
    for (int i = 0; i < n; ++i) {
        const double t = o * cross2(pr[(i + n - 1) % n], pr[i], pr[(i + 1) % n]);
        turned = turned || fabs(t) > tol;
        if (t < -tol) { convex = false; break; }
    }
    if (!turned) return false;
    if (convex) {
        for (int i = 1; i + 1 < n; ++i) { tri[3 * (i - 1)] = 0; tri[3 * (i - 1) + 1] = i; tri[3 * (i - 1) + 2] = i + 1; }
        return true;
    }
    unsigned char rem[kMaxRing];
    int m = n, nt = 0;
    for (int i = 0; i < n; ++i) rem[i] = i;
    while (m > 3) {
        bool clipped = false;
        for (int pos = 0; pos < m; ++pos) {
            const int pv = rem[(pos + m - 1) % m], cu = rem[pos], nx = rem[(pos + 1) % m];
            if (o * cross2(pr[pv], pr[cu], pr[nx]) <= tol) continue;
            bool contains = false;
            for (int q = 0; q < m && !contains; ++q) {
                const int c = rem[q];
                if (c == pv || c == cu || c == nx) continue;
                contains = o * cross2(pr[pv], pr[cu], pr[c]) >= -tol && o * cross2(pr[cu], pr[nx], pr[c]) >= -tol &&
                           o * cross2(pr[nx], pr[pv], pr[c]) >= -tol;
            }
            if (contains) continue;
            tri[3 * nt] = pv; tri[3 * nt + 1] = cu; tri[3 * nt + 2] = nx;
            ++nt;
            for (int q = pos; q + 1 < m; ++q) rem[q] = rem[q + 1];
            --m;
            clipped = true;
            break;
        }
        if (!clipped) return false;
    }
    tri[3 * nt] = rem[0]; tri[3 * nt + 1] = rem[1]; tri[3 * nt + 2] = rem[2];
    return true;
}

__device__ D3 areaOf(const D3* p, int n, const unsigned char* tri) {
    D3 a = d3(0, 0, 0);
    for (int t = 0; t < n - 2; ++t) {
        const D3 A = p[tri[3 * t]], B = p[tri[3 * t + 1]], Cc = p[tri[3 * t + 2]];
        a = a + 0.5 * cross(B - A, Cc - A);
    }
    return a;
}

__global__ void __launch_bounds__(kBlock) kFacePoints(int nf, const int* fpo, const int* fpi, const double* vx, double* psum, int* pcnt) {
    const int f = globalThread();
    if (f >= nf) return;
    D3 s = d3(0, 0, 0);
    for (int k = fpo[f]; k < fpo[f + 1]; ++k) s = s + ld3(vx, fpi[k]);
    st3(psum, f, s);
    pcnt[f] = fpo[f + 1] - fpo[f];
}

__global__ void __launch_bounds__(kBlock) kCellReference(MeshView m, const double* psum, const int* pcnt, double* ref) {
    const int c = globalThread();
    if (c >= m.nc) return;
    D3 s = d3(0, 0, 0);
    int cnt = 0;
    const RowRange R = m.row(c);
    for (int k = 0; R.valid(k); ++k) {
        const int f = m.mface[R.at(k)];
        if (f < 0) continue;
        s = s + ld3(psum, f);
        cnt += pcnt[f];
    }
    for (int k = m.bptr[c]; k < m.bptr[c + 1]; ++k) { s = s + ld3(psum, m.bface[k]); cnt += pcnt[m.bface[k]]; }
    st3(ref, c, s / static_cast<double>(cnt > 1 ? cnt : 1));
}

// Orientation from owner toward exterior, area vector, area-weighted centre.
__global__ void __launch_bounds__(kBlock) kFaceGeometry(int nf, const int* fpo, int* fpi, const double* vx, const int* own, const double* psum,
                              const double* ref, double* Cf, double* Sf, double* Af, double* nrm, int* err) {
    const int f = globalThread();
    if (f >= nf) return;
    const int b = fpo[f], n = fpo[f + 1] - b;
    if (n < 3 || n > kMaxRing) { atomicOr(err, kErrRing); return; }
    D3 p[kMaxRing];
    int ids[kMaxRing];
    for (int i = 0; i < n; ++i) { ids[i] = fpi[b + i]; p[i] = ld3(vx, ids[i]); }
    unsigned char tri[3 * (kMaxRing - 2)];
    if (!triangulate(p, n, tri)) { atomicOr(err, kErrDegenerate); return; }
    D3 area = areaOf(p, n, tri);
    const D3 avg = ld3(psum, f) / static_cast<double>(n);
    if (dot(area, avg - ld3(ref, own[f])) < 0.0) {
        for (int i = 1, j = n - 1; i < j; ++i, --j) {
            const int t = ids[i]; ids[i] = ids[j]; ids[j] = t;
            const D3 q = p[i]; p[i] = p[j]; p[j] = q;
        }
        for (int i = 0; i < n; ++i) fpi[b + i] = ids[i];
        if (!triangulate(p, n, tri)) { atomicOr(err, kErrDegenerate); return; }
        area = areaOf(p, n, tri);
    }
    D3 c = d3(0, 0, 0);
    double total = 0.0;
    for (int t = 0; t < n - 2; ++t) {
        const D3 A = p[tri[3 * t]], B = p[tri[3 * t + 1]], Cc = p[tri[3 * t + 2]];
        const D3 a = 0.5 * cross(B - A, Cc - A);
        const double wgt = sqrt(dot(a, a));
        if (!(wgt > 0.0)) atomicOr(err, kErrDegenerate);
        c = c + wgt * ((A + B + Cc) / 3.0);
        total += wgt;
    }
    if (!(sqrt(dot(area, area)) > 0.0) || !(total > 0.0)) atomicOr(err, kErrDegenerate);
    st3(Cf, f, c / total);
    st3(Sf, f, area);
    Af[f] = total;
    st3(nrm, f, area / total);
}

__device__ void cellFace(MeshView m, const int* fpo, const int* fpi, const double* vx, int c, int f, D3 ref,
                         double& vol, D3& mom, D3& closure, double& atot, int* err) {
    const int sign = m.own[f] == c ? 1 : -1;
    const int b = fpo[f], n = fpo[f + 1] - b;
    D3 p[kMaxRing];
    for (int i = 0; i < n; ++i) p[i] = ld3(vx, fpi[b + i]);
    unsigned char tri[3 * (kMaxRing - 2)];
    if (!triangulate(p, n, tri)) { atomicOr(err, kErrDegenerate); return; }
    for (int t = 0; t < n - 2; ++t) {
        const D3 A = p[tri[3 * t]], B = p[tri[3 * t + 1]], Dd = p[tri[3 * t + 2]];
        const double tet = sign * dot(A - ref, cross(B - ref, Dd - ref)) / 6.0;
        vol += tet;
        mom = mom + tet * ((ref + A + B + Dd) / 4.0);
    }
    closure = closure + static_cast<double>(sign) * ld3(m.Sf, f);
    atot += m.Af[f];
}

__global__ void __launch_bounds__(kBlock) kCellGeometry(MeshView m, const int* fpo, const int* fpi, const double* vx, const double* ref,
                              double* C, double* V, double* invV, int* err) {
    const int c = globalThread();
    if (c >= m.nc) return;
    const D3 r = ld3(ref, c);
    double vol = 0.0, atot = 0.0;
    D3 mom = d3(0, 0, 0), closure = d3(0, 0, 0);
    const RowRange R = m.row(c);
    for (int k = 0; R.valid(k); ++k) {
        const int f = m.mface[R.at(k)];
        if (f >= 0) cellFace(m, fpo, fpi, vx, c, f, r, vol, mom, closure, atot, err);
    }
    for (int k = m.bptr[c]; k < m.bptr[c + 1]; ++k) cellFace(m, fpo, fpi, vx, c, m.bface[k], r, vol, mom, closure, atot, err);
    if (!(vol > 0.0) || !isfinite(vol)) atomicOr(err, kErrVolume);
    if (!(sqrt(dot(closure, closure)) <= 1e-9 * fmax(atot, 1.0))) atomicOr(err, kErrClosure);
    V[c] = vol;
    invV[c] = 1.0 / vol;
    st3(C, c, mom / vol);
}

__global__ void __launch_bounds__(kBlock) kFaceFactors(MeshView m, double* w, double* orth, double* kf, double* skew, int* nonOrth, int* err) {
    const int f = globalThread();
    if (f >= m.nf) return;
    const int o = m.own[f], n = m.nei[f];
    const D3 Cf = ld3(m.Cf, f), Sf = ld3(m.Sf, f), nr = ld3(m.nrm, f), Co = ld3(m.C, o);
    const D3 d = n < 0 ? Cf - Co : ld3(m.C, n) - Co;
    const double d2 = dot(d, d), proj = dot(Sf, d);
    if (!(d2 > 0.0) || !(proj > 0.0)) atomicOr(err, kErrOrientation);
    const double oc = proj / d2;
    const D3 k = Sf - oc * d;
    orth[f] = oc;
    st3(kf, f, k);
    bool non = sqrt(dot(k, k)) > 1e-12 * m.Af[f];
    double wo = 1.0;
    D3 sk = d3(0, 0, 0);
    if (n >= 0) {
        const double od = fabs(dot(Cf - Co, nr));
        const double nd = fabs(dot(ld3(m.C, n) - Cf, nr));
        wo = nd / (od + nd);
        const double den = dot(d, nr);
        if (!(den > 0.0)) atomicOr(err, kErrOrientation);
        const double frac = dot(Cf - Co, nr) / den;
        sk = Cf - (Co + frac * d);
        non = non || sqrt(dot(sk, sk)) > 1e-12 * sqrt(d2);
    }
    w[f] = wo;
    st3(skew, f, sk);
    if (non) atomicOr(nonOrth, 1);
}

}  // namespace

DeviceMesh::DeviceMesh(const MeshDesc& caller, bool renumber, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    prepared = prepare(caller, renumber, cellNewToOld, faceNewToOld);
    const MeshDesc& d = prepared;
    nc = d.ncells();
    nf = d.nfaces();
    np = d.npatch;
    if (nc <= 0 || nf <= 0 || np <= 0) throw std::invalid_argument("mesh is empty");
    nfi = 0;
    while (nfi < nf && d.nei[nfi] >= 0) ++nfi;
    for (int f = nfi; f < nf; ++f)
        if (d.nei[f] >= 0) throw std::logic_error("face order is not interior-first");
    for (int f = 0; f < nf; ++f) {
        if (d.own[f] < 0 || d.own[f] == d.nei[f]) throw std::invalid_argument("face has invalid owner/neighbour");
        if (d.nei[f] < 0 && (d.patch[f] < 0 || d.patch[f] >= np)) throw std::invalid_argument("boundary face has invalid patch");
        if (d.nei[f] >= 0 && d.patch[f] >= 0) throw std::invalid_argument("internal face cannot belong to a patch");
    }
    hown = d.own;
    hnei = d.nei;
    // Row structure: interior half-faces sorted by (column, face); the prepared
    // face order already sorts faces by their lower cell.
    std::vector<int> cnt(nc, 0), bcnt(nc, 0);
    for (int f = 0; f < nfi; ++f) { ++cnt[d.own[f]]; ++cnt[d.nei[f]]; }
    for (int f = nfi; f < nf; ++f) ++bcnt[d.own[f]];
    hmptr.assign(nc + 1, 0);
    std::vector<int> hbptr(nc + 1, 0);
    for (int i = 0; i < nc; ++i) { hmptr[i + 1] = hmptr[i] + cnt[i]; hbptr[i + 1] = hbptr[i] + bcnt[i]; }
    nnzOff = hmptr[nc];
    std::vector<std::pair<long long, int>> entries(nnzOff);
    std::vector<int> fill(hmptr.begin(), hmptr.end() - 1);
    for (int f = 0; f < nfi; ++f) {
        entries[fill[d.own[f]]++] = {static_cast<long long>(d.nei[f]) * nf + f, f};
        entries[fill[d.nei[f]]++] = {static_cast<long long>(d.own[f]) * nf + f, f};
    }
    hmcol.resize(nnzOff);
    hmface.resize(nnzOff);
    for (int i = 0; i < nc; ++i) {
        std::sort(entries.begin() + hmptr[i], entries.begin() + hmptr[i + 1]);
        for (int k = hmptr[i]; k < hmptr[i + 1]; ++k) {
            hmface[k] = entries[k].second;
            hmcol[k] = static_cast<int>(entries[k].first / nf);
        }
    }
    // Sliced layout: level k of slice s holds the k-th entries of the slice's
    // rows; the level's width is one past the last lane that still has one.
    K = 0;
    for (int i = 0; i < nc; ++i) K = std::max(K, cnt[i]);
    const int nslices = (nc + 63) / 64;
    hsliceOff.assign(static_cast<std::size_t>(nslices) * (K + 1), 0);
    int base = 0;
    for (int s0 = 0; s0 < nslices; ++s0) {
        int* so = &hsliceOff[static_cast<std::size_t>(s0) * (K + 1)];
        so[0] = base;
        for (int k = 0; k < K; ++k) {
            int width = 0;
            for (int l = 0; l < 64 && s0 * 64 + l < nc; ++l)
                if (cnt[s0 * 64 + l] > k) width = l + 1;
            so[k + 1] = so[k] + width;
        }
        base = so[K];
    }
    nnzSliced = base;
    std::vector<int> scol(std::max(nnzSliced, 1), -1), sface(std::max(nnzSliced, 1), -1);
    csrToSliced.assign(nnzOff, 0);
    std::vector<int> hslotOwn(nfi), hslotNei(nfi);
    for (int i = 0; i < nc; ++i) {
        const int* so = &hsliceOff[static_cast<std::size_t>(i / 64) * (K + 1)];
        for (int k = 0; k < cnt[i]; ++k) {
            const int e = so[k] + (i % 64), q = hmptr[i] + k;
            csrToSliced[q] = e;
            scol[e] = hmcol[q];
            sface[e] = hmface[q];
            const int f = hmface[q];
            if (d.own[f] == i) hslotOwn[f] = e; else hslotNei[f] = e;
        }
    }
    std::vector<int> hbface(nf - nfi);
    std::vector<int> bfill(hbptr.begin(), hbptr.end() - 1);
    for (int f = nfi; f < nf; ++f) hbface[bfill[d.own[f]]++] = f;
    hpatchStart.assign(np + 1, 0);
    for (int f = nfi; f < nf; ++f) ++hpatchStart[d.patch[f] + 1];
    for (int p = 0; p < np; ++p) hpatchStart[p + 1] += hpatchStart[p];
    for (int f = nfi + 1; f < nf; ++f)
        if (d.patch[f] < d.patch[f - 1]) throw std::logic_error("boundary faces are not grouped by patch");

    own.upload(d.own, s); nei.upload(d.nei, s); patch.upload(d.patch, s);
    sliceOff.upload(hsliceOff, s); mcol.upload(scol, s); mface.upload(sface, s);
    bptr.upload(hbptr, s); bface.upload(hbface, s);
    slotOwn.upload(hslotOwn, s); slotNei.upload(hslotNei, s);
    patchStart.upload(hpatchStart, s);
    cellPerm.upload(cellNewToOld, s); facePerm.upload(faceNewToOld, s);
    Cf.alloc(3 * nf); Sf.alloc(3 * nf); nrm.alloc(3 * nf); Af.alloc(nf); w.alloc(nf); orth.alloc(nf);
    kf.alloc(3 * nf); skew.alloc(3 * nf); C.alloc(3 * nc); V.alloc(nc); invV.alloc(nc);

    DBuf<double> vx, psum(3 * static_cast<std::size_t>(nf)), ref(3 * static_cast<std::size_t>(nc));
    DBuf<int> fpo, fpi, pcnt(nf), flags(2);
    vx.upload(d.vert, s); fpo.upload(d.fpo, s); fpi.upload(d.fpi, s);
    flags.zero(s);
    const MeshView mv = view();
    kFacePoints<<<blocksFor(nf), kBlock, 0, s>>>(nf, fpo.p, fpi.p, vx.p, psum.p, pcnt.p);
    kCellReference<<<blocksFor(nc), kBlock, 0, s>>>(mv, psum.p, pcnt.p, ref.p);
    kFaceGeometry<<<blocksFor(nf), kBlock, 0, s>>>(nf, fpo.p, fpi.p, vx.p, own.p, psum.p, ref.p, Cf.p, Sf.p, Af.p,
                                                    nrm.p, flags.p);
    kCellGeometry<<<blocksFor(nc), kBlock, 0, s>>>(mv, fpo.p, fpi.p, vx.p, ref.p, C.p, V.p, invV.p, flags.p);
    kFaceFactors<<<blocksFor(nf), kBlock, 0, s>>>(mv, w.p, orth.p, kf.p, skew.p, flags.p + 1, flags.p);
    BSIM_LAUNCH_CHECK();
    const std::vector<int> f = flags.toHost(s);
    if (f[0] & kErrRing) throw std::invalid_argument("polyhedral face ring has fewer than 3 or more than 32 vertices");
    if (f[0] & kErrDegenerate) throw std::invalid_argument("polyhedral mesh contains a degenerate face");
    if (f[0] & kErrVolume) throw std::invalid_argument("polyhedral cell has non-positive signed volume");
    if (f[0] & kErrClosure) throw std::runtime_error("polyhedral cell face area vectors do not close");
    if (f[0] & kErrOrientation) throw std::invalid_argument("polyhedral face geometry has invalid owner/neighbour direction");
    orthogonal = f[1] == 0;
    geometrySeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

MeshView DeviceMesh::view() const {
    MeshView m;
    m.nc = nc; m.nf = nf; m.nfi = nfi; m.np = np; m.K = K;
    m.own = own.p; m.nei = nei.p; m.patch = patch.p;
    m.Cf = Cf.p; m.Sf = Sf.p; m.nrm = nrm.p; m.Af = Af.p; m.w = w.p; m.orth = orth.p; m.kf = kf.p; m.skew = skew.p;
    m.C = C.p; m.V = V.p; m.invV = invV.p;
    m.sliceOff = sliceOff.p; m.mcol = mcol.p; m.mface = mface.p; m.bptr = bptr.p; m.bface = bface.p;
    m.slotOwn = slotOwn.p; m.slotNei = slotNei.p; m.patchStart = patchStart.p;
    return m;
}

}  // namespace bsim
