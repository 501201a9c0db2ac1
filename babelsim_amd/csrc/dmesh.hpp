// The mesh as it lives in HBM: face geometry in face order (interior faces
// first, sorted by lower cell; boundary faces grouped by patch), cell geometry
// in renumbered cell order, and the row structure every cell-gather kernel and
// the CSR matrices share.
#pragma once

#include "device_common.hpp"
#include "mesh_host.hpp"

namespace bsim {

// Plain view passed by value to kernels.
struct MeshView {
    int nc, nf, nfi, np, K;
    const int *own, *nei, *patch;
    const double *Cf, *Sf, *nrm, *Af, *w, *orth, *kf, *skew;  // face AoS (x3) or scalar
    const double *C, *V, *invV;                                  // cell
    const int *sliceOff, *mcol, *mface;  // interior half-faces, sliced layout (see RowRange)
    const int *bptr, *bface;             // boundary faces per cell
    const int *slotOwn, *slotNei;        // interior face -> off-diagonal slot in owner / neighbour row
    const int* patchStart;               // boundary faces of patch p: [nfi + patchStart[p], nfi + patchStart[p+1])
    __device__ RowRange row(int c) const { return RowRange(sliceOff, K, c); }
};

class DeviceMesh {
public:
    DeviceMesh(const MeshDesc& callerMesh, bool renumber, hipStream_t stream);
    MeshView view() const;
    int nc = 0, nf = 0, nfi = 0, np = 0, nnzOff = 0, nnzSliced = 0, K = 0;
    bool orthogonal = true;
    // Host copies of the structure (renumbered numbering).  hmptr/hmcol/hmface
    // are the plain row-major CSR of the off-diagonal entries (sorted by column,
    // face); csrToSliced maps a CSR entry to its sliced position.
    std::vector<int> cellNewToOld, faceNewToOld;
    std::vector<int> hmptr, hmcol, hmface, hpatchStart, hown, hnei, csrToSliced, hsliceOff;
    MeshDesc prepared;
    // Device arrays.
    DBuf<int> own, nei, patch, sliceOff, mcol, mface, bptr, bface, slotOwn, slotNei, patchStart;
    DBuf<int> cellPerm, facePerm;  // new -> old
    DBuf<double> Cf, Sf, nrm, Af, w, orth, kf, skew, C, V, invV;
    double geometrySeconds = 0.0;
};

}  // namespace bsim
