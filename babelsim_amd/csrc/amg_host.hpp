// Host side of the aggregation AMG setup: greedy aggregation on a quantised
// strength graph and the symbolic Galerkin product (coarse pattern plus, for
// every coarse value, the ordered list of fine values it sums).  Runs once per
// linear system; the numeric Galerkin product runs on the GPU at every
// preconditioner rebuild.
#pragma once

#include <vector>

namespace bsim {

struct HostCsr {
    int n = 0;
    std::vector<int> ptr, col;
    std::vector<double> diag, val;
};

struct HostLevel {
    HostCsr A;                 // pattern + values (values from the aggregating setup)
    std::vector<int> agg;      // fine -> coarse (empty on the coarsest level)
    int nc = 0;
    std::vector<int> memPtr, memIdx;      // coarse -> fine members, ascending
    // Galerkin gather for the next level: coarse value e (diagonals first,
    // then off-diagonal entries in CSR order) sums fine values gSrc[gPtr[e]..)
    // where fine index i < n is diagonal i and n + k is off-diagonal entry k.
    std::vector<int> gPtr, gSrc;
};

// Greedy three-pass aggregation on the quantised strength |a_ij| / max_k |a_ik|.
void aggregate(const HostCsr& A, double theta, std::vector<int>& agg, int& nc);
// Symbolic + numeric Galerkin product with piecewise-constant prolongation.
void galerkinSymbolic(const HostCsr& A, const std::vector<int>& agg, int nc, HostCsr& Ac, std::vector<int>& gPtr,
                      std::vector<int>& gSrc, std::vector<int>& memPtr, std::vector<int>& memIdx);
// Builds the level structure (aggregation on each level) from the fine matrix.
std::vector<HostLevel> buildHierarchy(const HostCsr& A0, double theta, int maxLevels, int coarseSize);

// Sliced storage of a CSR pattern for wavefront-wide row kernels: storage
// slot q (slice q / 64, lane q % 64) holds row perm[q]; the rows of a slice
// are ordered by decreasing length, and the k-th entry of the row in lane l
// sits at sliceOff[slice * (K + 1) + k] + l.  Entries keep their CSR order.
struct SlicedPattern {
    int K = 0, nnz = 0;
    std::vector<int> sliceOff, col, perm, csrToStore;
};
SlicedPattern sliceRows(const HostCsr& A);

// Rewrites a Galerkin gather whose destination level is stored sliced:
// coarse value nc + k (CSR entry k) moves to nc + csrToStore[k].
void remapGatherDest(std::vector<int>& gPtr, std::vector<int>& gSrc, int nc, const std::vector<int>& csrToStore);
// Rewrites gather sources that address a sliced level: n + k -> n + csrToStore[k].
void remapGatherSrc(std::vector<int>& gSrc, int n, const std::vector<int>& csrToStore);

}  // namespace bsim
