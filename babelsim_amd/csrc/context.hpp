// A case context: one device, one stream, the mesh resident in HBM and the
// persistent linear systems keyed by equation identity (the role of
// PetscBackend's per-equation systems, src/backend/petsc_backend.cpp:123-198).
#pragma once

#include "dmesh.hpp"
#include "fvm_ops.hpp"
#include "linsolve.hpp"

#include <map>
#include <memory>
#include <string>

namespace bsim {

struct Context {
    Context(const MeshDesc& caller, bool renumber, int device);
    ~Context();
    int device = 0;
    hipStream_t stream = nullptr;
    std::unique_ptr<DeviceMesh> mesh;
    DBuf<int> err;  // operator error bits
    std::vector<int> cellOldToNew;
    void checkErrors();  // throws on device-side operator errors
    // Caller <-> internal numbering.
    template <class T> std::vector<T> cellsIn(const T* a, int ncomp) const;
    template <class T> void cellsOut(const std::vector<T>& v, T* a, int ncomp) const;
    template <class T> std::vector<T> facesIn(const T* a, int ncomp) const;
    template <class T> void facesOut(const std::vector<T>& v, T* a, int ncomp) const;
};

SolverCfg toSolverCfg(const void* cfg /* bsim_solver_config* */);

// Matrix storage pattern of a cell-centred operator on the mesh (sliced rows).
Pattern meshPattern(const DeviceMesh& mesh);

}  // namespace bsim
