// Host-side polyhedral face lists: synthetic workload generators and the
// bandwidth-reducing renumbering applied before the mesh goes to HBM.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bsim {

// A polyhedral mesh as BabelSim's Mesh::polyhedral consumes it
// (include/babelsim/mesh.h:116-121, PolyhedralFaceSpec): vertex rings, owner,
// neighbour (-1 on boundaries) and patch (-1 on interior faces).
// Patch roles of the mesh file (include/babelsim/mesh.h:79-86).
enum PatchKindId : int { kPatchGeneric = 0, kPatchWall, kPatchInlet, kPatchOutlet, kPatchSymmetry, kPatchProcessor };

struct MeshDesc {
    std::vector<double> vert;     // 3 * nv
    std::vector<int> fpo, fpi;    // face ring CSR
    std::vector<int> own, nei, patch;
    int npatch = 0;
    std::vector<std::string> patchName;  // empty unless read from a file
    std::vector<int> patchKind;
    int ncells() const;
    int nfaces() const { return static_cast<int>(own.size()); }
    // The topology checks of Mesh::polyhedral (src/core/mesh.cpp:321-367):
    // finite vertices, face rings inside [0, nv) with at least three distinct
    // vertices, owners in range, neighbours in range and not the owner,
    // interior faces without a patch and boundary faces with one in
    // [0, npatch), every cell with a face.  Returns the first failure, or an
    // empty string.
    std::string topologyError() const;
};

// Hex lattice nx*ny*nz over [lo,hi] with six patches (-x,+x,-y,+y,-z,+z).
// Optional explicit lattice vertices ((nx+1)(ny+1)(nz+1), x fastest).
// Interior vertices are jittered uniformly by +-jitter*spacing per axis and a
// seeded fraction of hexes is cut into two triangular prisms along a diagonal
// plane (one of three orientation families per hex).  Every shared lattice
// face is cut along the same global diagonal by both sides, so the cut faces
// are conforming triangle pairs.
MeshDesc generateHex(int nx, int ny, int nz, const double lo[3], const double hi[3], const double* vertices,
                     double jitter, double splitFraction, std::uint64_t seed);

// Reverse Cuthill-McKee cell order (when renumber) and the face order the
// kernels assume: interior faces by (lower cell, higher cell), then boundary
// faces by (patch, owner).  Owner/neighbour roles are kept as given.
MeshDesc prepare(const MeshDesc& in, bool renumber, std::vector<int>& cellNewToOld, std::vector<int>& faceNewToOld);

}  // namespace bsim
