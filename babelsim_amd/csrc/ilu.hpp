// Incomplete factorisations of the whole matrix with no fill: PETSc's icc
// (IC(0) from the upper triangle, petsc_linear_system.cpp:204-208) and asm on
// one rank (one subdomain, sub-PC ILU(0), :224-226), level-scheduled on the
// device.
#pragma once

#include "linsolve.hpp"

#include <vector>

namespace bsim {

// Factor pattern and values, row-major with each row's entries in
// elimination order and the diagonal at dpos[i]: L (unit lower) before the
// diagonal, U from it on.
struct IluView {
    const int* ptr;
    const int* col;
    const int* dpos;
    double* f;
};

class LevelIlu {
public:
    // cholesky: IC(0) of the matrix's upper triangle (the lower factor's
    // pattern and values are the upper triangle's transpose), else ILU(0).
    LevelIlu(hipStream_t s, bool cholesky) : stream(s), chol(cholesky) {}
    // Symbolic analysis of the system's pattern (once per pattern), in the
    // elimination order P.order (empty: the row order).
    void analyse(const Pattern& P, int n);
    // Numeric factorisation of A's current values.  False when a pivot is
    // zero (ILU) or not positive (IC), or not finite; apply() then returns +Inf
    // in every entry, as MatSolve does with a failed factor.
    bool factor(const CsrView<double>& A);
    // z = (L U)^-1 r: forward then backward substitution, level by level.
    void apply(const double* r, double* z);
    int forwardLevels() const { return static_cast<int>(fwd.hptr.size()) - 1; }
    int backwardLevels() const { return static_cast<int>(bwd.hptr.size()) - 1; }
    int launchesPerApply() const { return static_cast<int>(fwd.segs.size() + bwd.segs.size()); }
    long long factorEntries() const { return static_cast<long long>(hcol.size()); }
    bool failed() const { return fail; }
    bool cholesky() const { return chol; }

private:
    // Rows grouped by level; consecutive small levels run in one workgroup
    // with a barrier between levels, a large level takes a grid launch.
    struct Seg { int l0, l1; bool grid; int rows; };
    struct Schedule {
        std::vector<int> hptr, hrows;
        std::vector<Seg> segs;
        DBuf<int> ptr, rows;
    };
    void schedule(Schedule& S, const std::vector<int>& level, int levels);
    IluView view() { return {fptr.p, fcol.p, dpos.p, f.p}; }

    hipStream_t stream;
    bool chol;
    int n = 0;
    bool fail = false;
    std::vector<int> hptr, hcol, hdpos;  // factor pattern
    Schedule fwd, bwd;
    DBuf<int> fptr, fcol, dpos, srcPtr, src, updPtr, updDst, updSrc, flag;
    DBuf<double> f;
};

}  // namespace bsim
