// Oracle mesh: restates Mesh::polyhedral (src/core/mesh.cpp:321-578) and its
// polygon helpers (src/core/mesh.cpp:22-178).  Test infrastructure only.
#include "internal.hpp"

#include <algorithm>
#include <set>

namespace orc {
namespace {

struct Tri { Index a, b, c; };
struct P2 { double x, y; };

P2 project(const V3& p, int dropped) {  // mesh.cpp:37-41
    if (dropped == 0) return {p.y, p.z};
    if (dropped == 1) return {p.z, p.x};
    return {p.x, p.y};
}
double cross2(const P2& a, const P2& b, const P2& c) {  // mesh.cpp:43-45
    return (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
}
bool insideOrOn(const P2& p, const P2& a, const P2& b, const P2& c, double o, double tol) {  // mesh.cpp:47-60
    return o * cross2(a, b, p) >= -tol && o * cross2(b, c, p) >= -tol && o * cross2(c, a, p) >= -tol;
}

// Fan triangulation for convex rings, ear clipping otherwise (mesh.cpp:62-149).
std::vector<Tri> triangles(const std::vector<Index>& ids, const std::vector<V3>& vx) {
    const std::size_t n = ids.size();
    if (n < 3) throw std::invalid_argument("polyhedral face has fewer than three vertices");
    V3 nrm{};
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (std::size_t i = 0; i < n; ++i) {
        const V3& cur = vx[ids[i]];
        const V3& nxt = vx[ids[(i + 1) % n]];
        nrm += 0.5 * cross(cur, nxt);
        for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], cur[d]); hi[d] = std::max(hi[d], cur[d]); }
    }
    const double ax = std::abs(nrm.x), ay = std::abs(nrm.y), az = std::abs(nrm.z);
    const int dropped = ay > ax ? (az > ay ? 2 : 1) : (az > ax ? 2 : 0);
    const double scale = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1.0});
    const double tol = 1e-13 * scale * scale;
    std::vector<P2> pr(n);
    for (std::size_t i = 0; i < n; ++i) pr[i] = project(vx[ids[i]], dropped);
    double sa = 0.0;
    for (std::size_t i = 0; i < n; ++i) {
        const P2& c = pr[i]; const P2& d = pr[(i + 1) % n];
        sa += c.x * d.y - d.x * c.y;
    }
    if (std::abs(sa) <= tol) throw std::invalid_argument("polyhedral face projects to zero area");
    const double o = sa > 0.0 ? 1.0 : -1.0;
    bool convex = true, turned = false;
    for (std::size_t i = 0; i < n; ++i) {
        const double t = o * cross2(pr[(i + n - 1) % n], pr[i], pr[(i + 1) % n]);
        turned = turned || std::abs(t) > tol;
        if (t < -tol) { convex = false; break; }
    }
    if (!turned) throw std::invalid_argument("polyhedral face is collinear");
    std::vector<Tri> out;
    if (convex) {
        for (std::size_t i = 1; i + 1 < n; ++i) out.push_back({ids[0], ids[i], ids[i + 1]});
        return out;
    }
    std::vector<std::size_t> rem(n);
    for (std::size_t i = 0; i < n; ++i) rem[i] = i;
    while (rem.size() > 3) {
        bool clipped = false;
        for (std::size_t pos = 0; pos < rem.size(); ++pos) {
            const std::size_t pv = rem[(pos + rem.size() - 1) % rem.size()], cu = rem[pos],
                              nx = rem[(pos + 1) % rem.size()];
            if (o * cross2(pr[pv], pr[cu], pr[nx]) <= tol) continue;
            bool contains = false;
            for (std::size_t c : rem) {
                if (c == pv || c == cu || c == nx) continue;
                if (insideOrOn(pr[c], pr[pv], pr[cu], pr[nx], o, tol)) { contains = true; break; }
            }
            if (contains) continue;
            out.push_back({ids[pv], ids[cu], ids[nx]});
            rem.erase(rem.begin() + static_cast<std::ptrdiff_t>(pos));
            clipped = true;
            break;
        }
        if (!clipped) throw std::invalid_argument("polyhedral face is self-intersecting or not simple");
    }
    out.push_back({ids[rem[0]], ids[rem[1]], ids[rem[2]]});
    return out;
}

V3 areaVector(const std::vector<Index>& ids, const std::vector<V3>& vx) {  // mesh.cpp:151-160
    V3 a{};
    for (const Tri& t : triangles(ids, vx)) {
        const V3 &p = vx[t.a], &q = vx[t.b], &r = vx[t.c];
        a += 0.5 * cross(q - p, r - p);
    }
    return a;
}

void polygonGeometry(const std::vector<Index>& ids, const std::vector<V3>& vx, V3& centre, double& total) {
    // mesh.cpp:162-178: centre weighted by |triangle area|; area = sum of |triangle areas|.
    V3 c{};
    double t = 0.0;
    for (const Tri& tr : triangles(ids, vx)) {
        const V3 &p = vx[tr.a], &q = vx[tr.b], &r = vx[tr.c];
        const V3 a = 0.5 * cross(q - p, r - p);
        const double wgt = norm(a);
        if (!(wgt > 0.0) || !std::isfinite(wgt))
            throw std::invalid_argument("polyhedral mesh contains a degenerate face triangle");
        c += wgt * ((p + q + r) / 3.0);
        t += wgt;
    }
    if (!(t > 0.0) || !std::isfinite(t)) throw std::invalid_argument("polyhedral mesh contains a degenerate face");
    centre = c / t;
    total = t;
}

}  // namespace

Mesh buildMesh(const std::vector<V3>& vertices, const std::vector<Index>& fpo,
               const std::vector<Index>& fpi, const std::vector<Index>& owner,
               const std::vector<Index>& neighbour, const std::vector<Index>& patch, Index npatch) {
    Mesh m;
    m.vert = vertices;
    m.nv = static_cast<Index>(vertices.size());
    m.nf = static_cast<Index>(owner.size());
    m.np = npatch;
    if (m.nf == 0 || m.nv == 0 || npatch <= 0) throw std::invalid_argument("empty mesh");
    Index nc = 0;
    for (Index f = 0; f < m.nf; ++f) {
        if (owner[f] < 0 || neighbour[f] < kNone || neighbour[f] == owner[f])
            throw std::invalid_argument("polyhedral face has invalid owner/neighbour");
        nc = std::max(nc, owner[f] + 1);
        if (neighbour[f] != kNone) nc = std::max(nc, neighbour[f] + 1);
    }
    m.nc = nc;
    // Orientation reference: averaged face points per cell (mesh.cpp:370-386).
    std::vector<V3> psum(nc);
    std::vector<Index> pcnt(nc, 0);
    for (Index f = 0; f < m.nf; ++f) {
        V3 s{};
        for (Index k = fpo[f]; k < fpo[f + 1]; ++k) {
            if (fpi[k] < 0 || fpi[k] >= m.nv) throw std::invalid_argument("face references an invalid vertex");
            s += vertices[fpi[k]];
        }
        const Index cnt = fpo[f + 1] - fpo[f];
        psum[owner[f]] += s; pcnt[owner[f]] += cnt;
        if (neighbour[f] != kNone) { psum[neighbour[f]] += s; pcnt[neighbour[f]] += cnt; }
    }
    m.fpo.assign(1, 0);
    m.own = owner; m.nei = neighbour; m.patch = patch;
    m.Cf.resize(m.nf); m.Sf.resize(m.nf); m.nrm.resize(m.nf); m.Af.resize(m.nf);
    m.w.assign(m.nf, 1.0); m.orth.assign(m.nf, 0.0); m.kf.assign(m.nf, V3{}); m.skew.assign(m.nf, V3{});
    m.patchFaces.assign(npatch, {});
    std::vector<std::vector<Index>> cf(nc), cs(nc);
    for (Index f = 0; f < m.nf; ++f) {  // mesh.cpp:395-461
        std::vector<Index> ring(fpi.begin() + fpo[f], fpi.begin() + fpo[f + 1]);
        if (ring.size() < 3) throw std::invalid_argument("polyhedral face has fewer than three vertices");
        std::set<Index> uniq;
        V3 s{};
        for (Index v : ring) {
            if (!uniq.insert(v).second) throw std::invalid_argument("polyhedral face repeats a vertex");
            s += vertices[v];
        }
        const V3 avg = s / static_cast<double>(ring.size());
        cf[owner[f]].push_back(f); cs[owner[f]].push_back(1);
        if (neighbour[f] != kNone) {
            cf[neighbour[f]].push_back(f); cs[neighbour[f]].push_back(-1);
            if (patch[f] != kNone) throw std::invalid_argument("internal face cannot belong to a patch");
        } else if (patch[f] < 0 || patch[f] >= npatch) {
            throw std::invalid_argument("boundary face has invalid patch");
        }
        V3 area = areaVector(ring, vertices);
        const V3 est = psum[owner[f]] / static_cast<double>(std::max<Index>(1, pcnt[owner[f]]));
        if (dot(area, avg - est) < 0.0) {
            std::reverse(ring.begin() + 1, ring.end());
            area = areaVector(ring, vertices);
        }
        if (!(norm(area) > 0.0)) throw std::invalid_argument("polyhedral face has zero area");
        V3 centre; double asz;
        polygonGeometry(ring, vertices, centre, asz);
        for (Index v : ring) m.fpi.push_back(v);
        m.fpo.push_back(static_cast<Index>(m.fpi.size()));
        m.Cf[f] = centre; m.Sf[f] = area; m.Af[f] = asz; m.nrm[f] = area / asz;
        if (patch[f] != kNone) m.patchFaces[patch[f]].push_back(f);
    }
    m.cfo.assign(1, 0);
    for (Index c = 0; c < nc; ++c) {
        for (std::size_t i = 0; i < cf[c].size(); ++i) { m.cfi.push_back(cf[c][i]); m.cfs.push_back(cs[c][i]); }
        m.cfo.push_back(static_cast<Index>(m.cfi.size()));
    }
    // Tetrahedral decomposition about the averaged face points (mesh.cpp:484-513).
    m.C.resize(nc); m.V.resize(nc); m.invV.resize(nc);
    for (Index c = 0; c < nc; ++c) {
        if (cf[c].empty()) throw std::invalid_argument("polyhedral cell has no faces");
        const V3 ref = psum[c] / static_cast<double>(std::max<Index>(1, pcnt[c]));
        double vol = 0.0;
        V3 mom{};
        for (Index f : cf[c]) {
            const Index sign = m.own[f] == c ? 1 : -1;
            std::vector<Index> ids(m.fpi.begin() + m.fpo[f], m.fpi.begin() + m.fpo[f + 1]);
            for (const Tri& t : triangles(ids, vertices)) {
                const V3 &a = vertices[t.a], &b = vertices[t.b], &d = vertices[t.c];
                const double tet = sign * dot(a - ref, cross(b - ref, d - ref)) / 6.0;
                vol += tet;
                mom += tet * ((ref + a + b + d) / 4.0);
            }
        }
        if (!(vol > 0.0) || !std::isfinite(vol)) throw std::invalid_argument("polyhedral cell has non-positive signed volume");
        m.V[c] = vol; m.invV[c] = 1.0 / vol; m.C[c] = mom / vol;
    }
    // Face FVM factors (mesh.cpp:515-565).
    m.orthogonal = true;
    for (Index f = 0; f < m.nf; ++f) {
        const Index o = m.own[f], n = m.nei[f];
        const V3 d = n == kNone ? m.Cf[f] - m.C[o] : m.C[n] - m.C[o];
        const double d2 = norm2(d), proj = dot(m.Sf[f], d);
        if (!(d2 > 0.0) || !(proj > 0.0))
            throw std::invalid_argument("polyhedral face geometry has invalid owner/neighbour direction");
        const V3 nr = m.nrm[f];
        m.orth[f] = proj / d2;
        m.kf[f] = m.Sf[f] - m.orth[f] * d;
        if (norm(m.kf[f]) > 1e-12 * m.Af[f]) m.orthogonal = false;
        if (n != kNone) {
            const double od = std::abs(dot(m.Cf[f] - m.C[o], nr));
            const double nd = std::abs(dot(m.C[n] - m.Cf[f], nr));
            const double sum = od + nd;
            if (!(sum > 0.0)) throw std::invalid_argument("polyhedral interpolation distance is invalid");
            m.w[f] = nd / sum;
            const double den = dot(d, nr);
            if (!(den > 0.0) || !std::isfinite(den)) throw std::invalid_argument("interpolation direction invalid");
            const double frac = dot(m.Cf[f] - m.C[o], nr) / den;
            m.skew[f] = m.Cf[f] - (m.C[o] + frac * d);
            if (norm(m.skew[f]) > 1e-12 * std::sqrt(d2)) m.orthogonal = false;
        }
    }
    // The sorted off-diagonal pattern, built once (linsolve.cpp fills values into it).
    {
        std::vector<std::vector<std::pair<Index, Index>>> rows(nc);
        for (Index f = 0; f < m.nf; ++f) {
            if (m.nei[f] == kNone) continue;
            rows[m.own[f]].push_back({m.nei[f], f});
            rows[m.nei[f]].push_back({m.own[f], f});
        }
        m.csrPtr.assign(nc + 1, 0);
        for (Index i = 0; i < nc; ++i) {
            std::sort(rows[i].begin(), rows[i].end());
            m.csrPtr[i + 1] = m.csrPtr[i] + static_cast<Index>(rows[i].size());
            for (auto& e : rows[i]) {
                m.csrCol.push_back(e.first);
                m.csrFace.push_back(e.second);
                m.csrOwn.push_back(m.own[e.second] == i ? 1 : 0);
            }
        }
    }
    return m;
}

}  // namespace orc
