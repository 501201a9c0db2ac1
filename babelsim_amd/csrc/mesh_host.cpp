// Host-side mesh generation and renumbering (setup only, never on the timed
// path).  The generator produces the synthetic polyhedral cavities used by the
// benchmark and tests; prepare() fixes the cell and face order the device
// kernels rely on for coalesced face and row access.
#include "mesh_host.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace bsim {

std::string MeshDesc::topologyError() const {
    const long long nv = static_cast<long long>(vert.size() / 3);
    const int nf = nfaces();
    if (nv <= 0 || nf <= 0 || npatch <= 0) return "mesh arrays are empty";
    if (vert.size() % 3 != 0) return "vertex array is not a list of points";
    for (double v : vert)
        if (!std::isfinite(v)) return "mesh has a non-finite vertex";
    if (static_cast<int>(fpo.size()) != nf + 1 || nei.size() != own.size() || patch.size() != own.size())
        return "face arrays differ in length";
    if (fpo[0] != 0) return "face point offsets must start at 0";
    for (int f = 0; f < nf; ++f)
        if (fpo[f + 1] < fpo[f]) return "face point offsets must be non-decreasing";
    if (static_cast<std::size_t>(fpo[nf]) != fpi.size()) return "face point offsets do not end at the point count";
    for (int f = 0; f < nf; ++f) {
        if (own[f] < 0) return "face owner out of range";
        if (nei[f] >= 0) {
            if (nei[f] == own[f]) return "face neighbour out of range";
            if (patch[f] != -1) return "interior face with a patch";
        } else if (nei[f] != -1 || patch[f] < 0 || patch[f] >= npatch) {
            return "boundary face without a valid patch";
        }
        if (fpo[f + 1] - fpo[f] < 3) return "face with fewer than three vertices";
        std::vector<int> ring(fpi.begin() + fpo[f], fpi.begin() + fpo[f + 1]);
        for (int v : ring)
            if (v < 0 || v >= nv) return "face vertex index out of range";
        std::sort(ring.begin(), ring.end());
        if (std::adjacent_find(ring.begin(), ring.end()) != ring.end()) return "face repeats a vertex";
    }
    const int nc = ncells();
    std::vector<char> hasFace(nc, 0);
    for (int f = 0; f < nf; ++f) {
        hasFace[own[f]] = 1;
        if (nei[f] >= 0) hasFace[nei[f]] = 1;
    }
    if (std::find(hasFace.begin(), hasFace.end(), 0) != hasFace.end()) return "cell with no faces";
    return {};
}

int MeshDesc::ncells() const {
    int n = 0;
    for (std::size_t f = 0; f < own.size(); ++f) {
        n = std::max(n, own[f] + 1);
        if (nei[f] >= 0) n = std::max(n, nei[f] + 1);
    }
    return n;
}

namespace {

std::uint64_t mix(std::uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
double unit(std::uint64_t seed, std::uint64_t a, std::uint64_t b) {
    return static_cast<double>(mix(seed ^ mix(a * 8 + b)) >> 11) * 0x1.0p-53;
}

// Local hex vertices: v0=(0,0,0) v1=(1,0,0) v2=(1,1,0) v3=(0,1,0) and +4 at dz=1.
constexpr int kLocal[2][2][2] = {{{0, 1}, {3, 2}}, {{4, 5}, {7, 6}}};  // [dz][dy][dx]
constexpr int kOff[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
constexpr int kFace[6][4] = {{0, 4, 7, 3}, {1, 2, 6, 5}, {0, 1, 5, 4}, {3, 7, 6, 2}, {0, 3, 2, 1}, {4, 5, 6, 7}};
// Diagonal plane per family: prism A holds vertices with d[first] >= d[second].
constexpr int kPlane[3][2] = {{1, 2}, {0, 2}, {0, 1}};
constexpr int kInterior[3][4] = {{0, 1, 6, 7}, {0, 3, 6, 5}, {0, 2, 6, 4}};

}  // namespace

MeshDesc generateHex(int nx, int ny, int nz, const double lo[3], const double hi[3], const double* vertices,
                     double jitter, double splitFraction, std::uint64_t seed) {
    if (nx <= 0 || ny <= 0 || nz <= 0) throw std::invalid_argument("hex counts must be positive");
    if (!(jitter >= 0.0 && jitter < 0.5)) throw std::invalid_argument("jitter must lie in [0, 0.5)");
    if (!(splitFraction >= 0.0 && splitFraction <= 1.0)) throw std::invalid_argument("split fraction must lie in [0, 1]");
    const long long nvx = nx + 1, nvy = ny + 1, nvz = nz + 1;
    const long long nv = nvx * nvy * nvz;
    const long long nh = static_cast<long long>(nx) * ny * nz;
    if (nh * 2 + 1 > 0x7fffffffLL || nv > 0x7fffffffLL) throw std::invalid_argument("mesh exceeds 32-bit indexing");
    MeshDesc m;
    m.npatch = 6;
    m.vert.resize(3 * nv);
    const double h[3] = {(hi[0] - lo[0]) / nx, (hi[1] - lo[1]) / ny, (hi[2] - lo[2]) / nz};
    for (long long k = 0; k < nvz; ++k)
        for (long long j = 0; j < nvy; ++j)
            for (long long i = 0; i < nvx; ++i) {
                const long long v = i + nvx * (j + nvy * k);
                double p[3];
                if (vertices) { p[0] = vertices[3 * v]; p[1] = vertices[3 * v + 1]; p[2] = vertices[3 * v + 2]; }
                else {
                    p[0] = lo[0] + (hi[0] - lo[0]) * static_cast<double>(i) / nx;
                    p[1] = lo[1] + (hi[1] - lo[1]) * static_cast<double>(j) / ny;
                    p[2] = lo[2] + (hi[2] - lo[2]) * static_cast<double>(k) / nz;
                }
                const bool interior = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
                if (interior && jitter > 0.0)
                    for (int d = 0; d < 3; ++d) p[d] += (2.0 * unit(seed, static_cast<std::uint64_t>(v), d) - 1.0) * jitter * h[d];
                m.vert[3 * v] = p[0]; m.vert[3 * v + 1] = p[1]; m.vert[3 * v + 2] = p[2];
            }
    // Split decisions and cell numbering: a split hex becomes two consecutive cells.
    std::vector<signed char> family(nh, -1);
    std::vector<int> base(nh);
    const std::uint64_t splitSeed = mix(seed ^ 0x5b1175eedULL);
    for (long long hx = 0; hx < nh; ++hx)
        if (splitFraction > 0.0 && unit(splitSeed, static_cast<std::uint64_t>(hx), 0) < splitFraction)
            family[hx] = static_cast<signed char>(mix(splitSeed + static_cast<std::uint64_t>(hx)) % 3);
    auto hexId = [&](int a, int b, int c) { return static_cast<long long>(a) + nx * (static_cast<long long>(b) + static_cast<long long>(ny) * c); };
    auto vid = [&](int a, int b, int c, int lv) {
        return static_cast<int>((a + kOff[lv][0]) + nvx * ((b + kOff[lv][1]) + nvy * (c + kOff[lv][2])));
    };
    // A split is kept only if every face of the two prisms keeps its area
    // vector within ~84 degrees of the owner-to-neighbour direction (vertex
    // averages stand in for centroids); heavy jitter can otherwise produce a
    // face with S_f . d <= 0, which Mesh::polyhedral rejects (mesh.cpp:515-520).
    {
        auto pt = [&](int v) { return &m.vert[3 * static_cast<std::size_t>(v)]; };
        auto partCentre = [&](int a, int b, int c, int part, double* out) {  // part -1: whole hex
            const long long hx = hexId(a, b, c);
            out[0] = out[1] = out[2] = 0.0;
            int cnt = 0;
            for (int lv = 0; lv < 8; ++lv) {
                if (part >= 0) {
                    const int f = family[hx], first = kPlane[f][0], second = kPlane[f][1];
                    const bool inA = kOff[lv][first] >= kOff[lv][second];
                    const bool inB = kOff[lv][first] <= kOff[lv][second];
                    if ((part == 0 && !inA) || (part == 1 && !inB)) continue;
                }
                const double* p = pt(vid(a, b, c, lv));
                out[0] += p[0]; out[1] += p[1]; out[2] += p[2];
                ++cnt;
            }
            out[0] /= cnt; out[1] /= cnt; out[2] /= cnt;
        };
        // Which part of hex (a,b,c) holds the lattice vertices g.
        auto partOf = [&](int a, int b, int c, const int* g, int n) {
            const long long hx = hexId(a, b, c);
            if (family[hx] < 0) return -1;
            const int first = kPlane[family[hx]][0], second = kPlane[family[hx]][1];
            for (int t = 0; t < n; ++t) {
                const long long v = g[t];
                const int d[3] = {static_cast<int>(v % nvx) - a, static_cast<int>((v / nvx) % nvy) - b, static_cast<int>(v / (nvx * nvy)) - c};
                if (d[first] < d[second]) return 1;
            }
            return 0;
        };
        auto faceOk = [&](const int* ring, int n, const double* co, const double* cn) {
            double S[3] = {0, 0, 0}, avg[3] = {0, 0, 0};
            for (int t = 0; t < n; ++t) {
                const double* p = pt(ring[t]);
                const double* q = pt(ring[(t + 1) % n]);
                S[0] += 0.5 * (p[1] * q[2] - p[2] * q[1]);
                S[1] += 0.5 * (p[2] * q[0] - p[0] * q[2]);
                S[2] += 0.5 * (p[0] * q[1] - p[1] * q[0]);
                for (int k = 0; k < 3; ++k) avg[k] += p[k] / n;
            }
            double sgn = 0.0;
            for (int k = 0; k < 3; ++k) sgn += S[k] * (avg[k] - co[k]);
            const double d[3] = {cn[0] - co[0], cn[1] - co[1], cn[2] - co[2]};
            const double sd = (sgn < 0 ? -1.0 : 1.0) * (S[0] * d[0] + S[1] * d[1] + S[2] * d[2]);
            const double ns = std::sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
            const double nd = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            return sd > 0.1 * ns * nd;
        };
        auto hexOk = [&](int a, int b, int c) {
            const long long hx = hexId(a, b, c);
            double cA[3], cB[3], cn[3];
            partCentre(a, b, c, 0, cA);
            partCentre(a, b, c, 1, cB);
            int ring[4];
            for (int t = 0; t < 4; ++t) ring[t] = vid(a, b, c, kInterior[family[hx]][t]);
            if (!faceOk(ring, 4, cA, cB)) return false;
            const int cnt[3] = {nx, ny, nz};
            for (int lf = 0; lf < 6; ++lf) {
                const int dir = lf / 2;
                int nb[3] = {a, b, c};
                nb[dir] += lf % 2 == 1 ? 1 : -1;
                if (nb[dir] < 0 || nb[dir] >= cnt[dir]) continue;
                for (int t = 0; t < 4; ++t) ring[t] = vid(a, b, c, kFace[lf][t]);
                const bool cut = family[hx] == dir || family[hexId(nb[0], nb[1], nb[2])] == dir;
                int subs[2][4], sn[2], ns = 0;
                if (!cut) { for (int t = 0; t < 4; ++t) subs[0][t] = ring[t]; sn[0] = 4; ns = 1; }
                else {
                    int p0 = 0;
                    for (int t = 0; t < 4; ++t) {
                        const int lv = kFace[lf][t];
                        if (kOff[lv][(dir + 1) % 3] == 0 && kOff[lv][(dir + 2) % 3] == 0) { p0 = t; break; }
                    }
                    subs[0][0] = ring[p0]; subs[0][1] = ring[(p0 + 1) % 4]; subs[0][2] = ring[(p0 + 2) % 4];
                    subs[1][0] = ring[(p0 + 2) % 4]; subs[1][1] = ring[(p0 + 3) % 4]; subs[1][2] = ring[p0];
                    sn[0] = sn[1] = 3;
                    ns = 2;
                }
                for (int k = 0; k < ns; ++k) {
                    const int po = partOf(a, b, c, subs[k], sn[k]);
                    const int pn = partOf(nb[0], nb[1], nb[2], subs[k], sn[k]);
                    partCentre(nb[0], nb[1], nb[2], pn, cn);
                    if (!faceOk(subs[k], sn[k], po == 1 ? cB : cA, cn)) return false;
                }
            }
            return true;
        };
        for (int pass = 0; pass < 8; ++pass) {
            bool changed = false;
            for (int c = 0; c < nz; ++c)
                for (int b = 0; b < ny; ++b)
                    for (int a = 0; a < nx; ++a) {
                        const long long hx = hexId(a, b, c);
                        if (family[hx] >= 0 && !hexOk(a, b, c)) { family[hx] = -1; changed = true; }
                    }
            if (!changed) break;
        }
    }
    int next = 0;
    for (long long hx = 0; hx < nh; ++hx) {
        base[hx] = next;
        next += family[hx] >= 0 ? 2 : 1;
    }
    // Cell of hex (a,b,c) containing the lattice vertices `g` (global ids).
    auto cellOf = [&](int a, int b, int c, const int* g, int n) {
        const long long hx = hexId(a, b, c);
        if (family[hx] < 0) return base[hx];
        const int first = kPlane[family[hx]][0], second = kPlane[family[hx]][1];
        bool inA = true;
        for (int t = 0; t < n; ++t) {
            const long long v = g[t];
            const int d[3] = {static_cast<int>(v % nvx) - a, static_cast<int>((v / nvx) % nvy) - b, static_cast<int>(v / (nvx * nvy)) - c};
            if (d[first] < d[second]) { inA = false; break; }
        }
        return base[hx] + (inA ? 0 : 1);
    };
    m.fpo.push_back(0);
    auto emit = [&](const int* ring, int n, int owner, int neighbour, int patch) {
        for (int t = 0; t < n; ++t) m.fpi.push_back(ring[t]);
        m.fpo.push_back(static_cast<int>(m.fpi.size()));
        m.own.push_back(owner); m.nei.push_back(neighbour); m.patch.push_back(patch);
    };
    for (int c = 0; c < nz; ++c)
        for (int b = 0; b < ny; ++b)
            for (int a = 0; a < nx; ++a) {
                const long long hx = hexId(a, b, c);
                if (family[hx] >= 0) {
                    int ring[4];
                    for (int t = 0; t < 4; ++t) ring[t] = vid(a, b, c, kInterior[family[hx]][t]);
                    emit(ring, 4, base[hx], base[hx] + 1, -1);
                }
                for (int lf = 0; lf < 6; ++lf) {
                    const int dir = lf / 2;
                    const bool plus = lf % 2 == 1;
                    int nb[3] = {a, b, c};
                    nb[dir] += plus ? 1 : -1;
                    const int cnt[3] = {nx, ny, nz};
                    const bool boundary = nb[dir] < 0 || nb[dir] >= cnt[dir];
                    if (!boundary && !plus) continue;  // the lower hex emits shared faces
                    bool cut = family[hx] == dir;
                    if (!boundary && family[hexId(nb[0], nb[1], nb[2])] == dir) cut = true;
                    int ring[4];
                    for (int t = 0; t < 4; ++t) ring[t] = vid(a, b, c, kFace[lf][t]);
                    auto place = [&](const int* r, int n) {
                        const int owner = cellOf(a, b, c, r, n);
                        if (boundary) emit(r, n, owner, -1, lf);
                        else emit(r, n, owner, cellOf(nb[0], nb[1], nb[2], r, n), -1);
                    };
                    if (!cut) { place(ring, 4); continue; }
                    // Main diagonal: the ring vertices whose two in-plane offsets agree.
                    int p0 = -1;
                    for (int t = 0; t < 4; ++t) {
                        const int lv = kFace[lf][t];
                        const int u = kOff[lv][(dir + 1) % 3], w = kOff[lv][(dir + 2) % 3];
                        if (u == 0 && w == 0) { p0 = t; break; }
                    }
                    const int t0[3] = {ring[p0], ring[(p0 + 1) % 4], ring[(p0 + 2) % 4]};
                    const int t1[3] = {ring[(p0 + 2) % 4], ring[(p0 + 3) % 4], ring[p0]};
                    place(t0, 3);
                    place(t1, 3);
                }
            }
    (void)kLocal;
    return m;
}

MeshDesc prepare(const MeshDesc& in, bool renumber, std::vector<int>& cellNewToOld, std::vector<int>& faceNewToOld) {
    const std::string bad = in.topologyError();
    if (!bad.empty()) throw std::invalid_argument("invalid polyhedral mesh: " + bad);
    const int nc = in.ncells();
    const int nf = in.nfaces();
    std::vector<int> oldToNew(nc);
    cellNewToOld.resize(nc);
    if (renumber) {
        // Cell graph (unique neighbours).
        std::vector<int> deg(nc, 0);
        for (int f = 0; f < nf; ++f) if (in.nei[f] >= 0) { ++deg[in.own[f]]; ++deg[in.nei[f]]; }
        std::vector<int> ptr(nc + 1, 0);
        for (int i = 0; i < nc; ++i) ptr[i + 1] = ptr[i] + deg[i];
        std::vector<int> adj(ptr[nc]), fill(ptr.begin(), ptr.end() - 1);
        for (int f = 0; f < nf; ++f)
            if (in.nei[f] >= 0) { adj[fill[in.own[f]]++] = in.nei[f]; adj[fill[in.nei[f]]++] = in.own[f]; }
        for (int i = 0; i < nc; ++i) {
            std::sort(adj.begin() + ptr[i], adj.begin() + ptr[i + 1]);
            deg[i] = static_cast<int>(std::unique(adj.begin() + ptr[i], adj.begin() + ptr[i + 1]) - (adj.begin() + ptr[i]));
        }
        std::vector<int> level(nc, -1), order;
        order.reserve(nc);
        auto bfs = [&](int root, std::vector<int>& lv) {  // returns last-level node of minimum degree
            std::vector<int> q{root};
            lv.assign(nc, -1);
            lv[root] = 0;
            for (std::size_t h = 0; h < q.size(); ++h) {
                const int u = q[h];
                for (int k = ptr[u]; k < ptr[u] + deg[u]; ++k)
                    if (lv[adj[k]] < 0) { lv[adj[k]] = lv[u] + 1; q.push_back(adj[k]); }
            }
            int far = q.back();
            for (int u : q) if (lv[u] == lv[far] && (deg[u] < deg[far] || (deg[u] == deg[far] && u < far))) far = u;
            return std::make_pair(far, lv[q.back()]);
        };
        std::vector<char> seen(nc, 0);
        std::vector<int> lv;
        for (int start = 0; start < nc; ++start) {
            if (seen[start]) continue;
            // Pseudo-peripheral root of this component (George-Liu).
            int root = start;
            auto r = bfs(root, lv);
            for (int it = 0; it < 4; ++it) {
                auto r2 = bfs(r.first, lv);
                if (r2.second <= r.second) break;
                root = r.first;
                r = r2;
            }
            root = r.first;
            std::size_t head = order.size();
            order.push_back(root);
            seen[root] = 1;
            std::vector<int> nb;
            while (head < order.size()) {
                const int u = order[head++];
                nb.clear();
                for (int k = ptr[u]; k < ptr[u] + deg[u]; ++k) if (!seen[adj[k]]) { nb.push_back(adj[k]); seen[adj[k]] = 1; }
                std::sort(nb.begin(), nb.end(), [&](int x, int y) { return deg[x] != deg[y] ? deg[x] < deg[y] : x < y; });
                order.insert(order.end(), nb.begin(), nb.end());
            }
        }
        std::reverse(order.begin(), order.end());
        // Within each 64-cell slice (one wavefront of row threads) order the
        // rows by decreasing number of interior faces, so the k-th neighbours
        // of a slice's rows form one contiguous, unpadded run in HBM.
        std::vector<int> nfaces(nc, 0);
        for (int f = 0; f < nf; ++f) if (in.nei[f] >= 0) { ++nfaces[in.own[f]]; ++nfaces[in.nei[f]]; }
        for (int s0 = 0; s0 < nc; s0 += 64)
            std::stable_sort(order.begin() + s0, order.begin() + std::min(nc, s0 + 64),
                             [&](int x, int y) { return nfaces[x] > nfaces[y]; });
        for (int i = 0; i < nc; ++i) { cellNewToOld[i] = order[i]; oldToNew[order[i]] = i; }
    } else {
        std::iota(cellNewToOld.begin(), cellNewToOld.end(), 0);
        std::iota(oldToNew.begin(), oldToNew.end(), 0);
    }
    std::vector<long long> key(nf);
    std::vector<int> idx(nf);
    std::iota(idx.begin(), idx.end(), 0);
    for (int f = 0; f < nf; ++f) {
        const int o = oldToNew[in.own[f]];
        if (in.nei[f] >= 0) {
            const int n = oldToNew[in.nei[f]];
            key[f] = static_cast<long long>(std::min(o, n)) * (static_cast<long long>(nc) + 1) + std::max(o, n);
        } else {
            key[f] = static_cast<long long>(nc + 1) * (nc + 1) + static_cast<long long>(in.patch[f]) * (nc + 1) + o;
        }
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return key[x] < key[y]; });
    MeshDesc out;
    out.vert = in.vert;
    out.npatch = in.npatch;
    out.patchName = in.patchName;
    out.patchKind = in.patchKind;
    out.fpo.push_back(0);
    faceNewToOld = idx;
    for (int nfc = 0; nfc < nf; ++nfc) {
        const int f = idx[nfc];
        for (int k = in.fpo[f]; k < in.fpo[f + 1]; ++k) out.fpi.push_back(in.fpi[k]);
        out.fpo.push_back(static_cast<int>(out.fpi.size()));
        out.own.push_back(oldToNew[in.own[f]]);
        out.nei.push_back(in.nei[f] >= 0 ? oldToNew[in.nei[f]] : -1);
        out.patch.push_back(in.patch[f]);
    }
    return out;
}

}  // namespace bsim
