// C ABI: host mesh lists (generation, wrapping, renumbering).
#include "../../include/bsim.h"
#include "errors.hpp"
#include "io_host.hpp"
#include "mesh_host.hpp"

#include <algorithm>
#include <cstring>
#include <memory>

struct bsim_mesh {
    bsim::MeshDesc d;
};

namespace bsim {
thread_local std::string g_lastError;
void setError(const std::string& message) { g_lastError = message; }
}  // namespace bsim

extern "C" {

const char* bsim_last_error(void) { return bsim::g_lastError.c_str(); }
int bsim_version(void) { return 1; }

int bsim_mesh_generate_hex(int nx, int ny, int nz, const double* lo, const double* hi, const double* vertices,
                           double jitter, double split_fraction, uint64_t seed, bsim_mesh** out) {
    return bsim::guarded([&] {
        auto* m = new bsim_mesh;
        try { m->d = bsim::generateHex(nx, ny, nz, lo, hi, vertices, jitter, split_fraction, seed); }
        catch (...) { delete m; throw; }
        *out = m;
    });
}

int bsim_mesh_from_arrays(int nv, const double* vertices, int nf, const int* fpo, const int* fpi, const int* owner,
                          const int* neighbour, const int* patch, int npatch, bsim_mesh** out) {
    return bsim::guarded([&] {
        if (nv <= 0 || nf <= 0 || npatch <= 0) throw std::invalid_argument("mesh arrays are empty");
        if (!vertices || !fpo || !fpi || !owner || !neighbour || !patch || !out)
            throw std::invalid_argument("null argument");
        // The ring offsets decide how much of face_point_ids is read.
        if (fpo[0] != 0) throw std::invalid_argument("face point offsets must start at 0");
        for (int f = 0; f < nf; ++f)
            if (fpo[f + 1] < fpo[f]) throw std::invalid_argument("face point offsets must be non-decreasing");
        auto m = std::make_unique<bsim_mesh>();
        m->d.vert.assign(vertices, vertices + 3 * static_cast<long long>(nv));
        m->d.fpo.assign(fpo, fpo + nf + 1);
        m->d.fpi.assign(fpi, fpi + fpo[nf]);
        m->d.own.assign(owner, owner + nf);
        m->d.nei.assign(neighbour, neighbour + nf);
        m->d.patch.assign(patch, patch + nf);
        m->d.npatch = npatch;
        const std::string bad = m->d.topologyError();
        if (!bad.empty()) throw std::invalid_argument("invalid polyhedral mesh: " + bad);
        *out = m.release();
    });
}

int bsim_mesh_prepare(const bsim_mesh* in, int renumber, bsim_mesh** out, int* cell_new_to_old,
                      int* face_new_to_old) {
    return bsim::guarded([&] {
        std::vector<int> c, f;
        auto* m = new bsim_mesh;
        m->d = bsim::prepare(in->d, renumber != 0, c, f);
        if (cell_new_to_old) std::memcpy(cell_new_to_old, c.data(), sizeof(int) * c.size());
        if (face_new_to_old) std::memcpy(face_new_to_old, f.data(), sizeof(int) * f.size());
        *out = m;
    });
}

void bsim_mesh_sizes(const bsim_mesh* m, int* nv, int* nf, int* nfp, int* npatch, int* ncells) {
    if (nv) *nv = static_cast<int>(m->d.vert.size() / 3);
    if (nf) *nf = m->d.nfaces();
    if (nfp) *nfp = static_cast<int>(m->d.fpi.size());
    if (npatch) *npatch = m->d.npatch;
    if (ncells) *ncells = m->d.ncells();
}

void bsim_mesh_arrays(const bsim_mesh* m, double* vertices, int* fpo, int* fpi, int* owner, int* neighbour,
                      int* patch) {
    const auto& d = m->d;
    if (vertices) std::memcpy(vertices, d.vert.data(), sizeof(double) * d.vert.size());
    if (fpo) std::memcpy(fpo, d.fpo.data(), sizeof(int) * d.fpo.size());
    if (fpi) std::memcpy(fpi, d.fpi.data(), sizeof(int) * d.fpi.size());
    if (owner) std::memcpy(owner, d.own.data(), sizeof(int) * d.own.size());
    if (neighbour) std::memcpy(neighbour, d.nei.data(), sizeof(int) * d.nei.size());
    if (patch) std::memcpy(patch, d.patch.data(), sizeof(int) * d.patch.size());
}

void bsim_mesh_free(bsim_mesh* m) { delete m; }

int bsim_mesh_read(const char* path, bsim_mesh** out) {
    return bsim::guarded([&] {
        if (!path || !out) throw std::invalid_argument("null argument");
        auto* m = new bsim_mesh;
        try { m->d = bsim::readMeshFile(path); }
        catch (...) { delete m; throw; }
        *out = m;
    });
}

int bsim_mesh_write(const bsim_mesh* m, const char* path) {
    return bsim::guarded([&] {
        if (!m || !path) throw std::invalid_argument("null argument");
        bsim::writeMeshFile(m->d, path);
    });
}

int bsim_mesh_patch(const bsim_mesh* m, int patch, char* name, int cap, int* kind) {
    return bsim::guarded([&] {
        if (!m || patch < 0 || patch >= m->d.npatch) throw std::invalid_argument("patch index out of range");
        const bool named = patch < static_cast<int>(m->d.patchName.size());
        const std::string n = named ? m->d.patchName[patch] : "patch" + std::to_string(patch);
        if (name && cap > 0) {
            const std::size_t k = std::min<std::size_t>(n.size(), static_cast<std::size_t>(cap - 1));
            std::memcpy(name, n.data(), k);
            name[k] = '\0';
        }
        if (kind) *kind = named ? m->d.patchKind[patch] : bsim::kPatchGeneric;
    });
}

int bsim_write_field_csv(const char* time_dir, int rank, const char* name, int ncells, const int* global_ids,
                         const double* centres, int ncomp, const double* values) {
    return bsim::guarded([&] {
        if (!time_dir || !name || !centres || !values || ncells < 0 || rank < 0)
            throw std::invalid_argument("invalid field output arguments");
        bsim::writeFieldCsv(time_dir, rank, name, ncells, global_ids, centres, ncomp, values);
    });
}

int bsim_write_result_metadata(const char* time_dir, int rank, int ranks, const bsim_mesh* m, const int* global_ids,
                               long long global_cells, const char* time_name, int nfields, const char* const* names,
                               const char* const* types, const int* locations) {
    return bsim::guarded([&] {
        if (!time_dir || !m || !time_name || nfields <= 0 || !names || !types || !locations || rank < 0 ||
            ranks <= rank)
            throw std::invalid_argument("invalid result metadata arguments");
        std::vector<std::string> n, t;
        std::vector<int> l(locations, locations + nfields);
        for (int k = 0; k < nfields; ++k) {
            n.emplace_back(names[k] ? names[k] : "");
            t.emplace_back(types[k] ? types[k] : "");
        }
        bsim::writeResultMetadata(time_dir, rank, ranks, m->d, global_ids, global_cells, time_name, n, t, l);
    });
}

int bsim_read_field_values(const char* path, const char* name, int ncells, int ncomp, double* out) {
    return bsim::guarded([&] {
        if (!path || !out) throw std::invalid_argument("null argument");
        bsim::readFieldValues(path, name ? name : "", ncells, ncomp, out);
    });
}

}  // extern "C"
