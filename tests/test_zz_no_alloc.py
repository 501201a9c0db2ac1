"""The allocation-free CSR solve (bsim_csr_solve_noalloc) against the
allocating one (bsim_csr_solve) on the same systems: the same kernels run in
the same order on the scratch array, so statuses, iteration counts, residual
histories and solutions agree bit for bit.  Two calls in a row, the second on
a smaller system, show that what the first call left in the shared array does
not leak into the second."""
import ctypes as C

import numpy as np
import pytest
import torch

from babelsim_amd import solver_config
from babelsim_amd._capi import SolveResult, lib
from babelsim_amd.device import csr_solve, csr_solve_noalloc

pytestmark = pytest.mark.gpu

MAX_ROWS = 524288  # BSIM_NOALLOC_MAX_ROWS


def lattice(nx, ny, nz, asym, seed):
    """7-point operator on an nx x ny x nz lattice: diagonal 6.5 apart from the
    row-major off-diagonal entries -1 (plus an antisymmetric convection part
    of size asym), and a seeded right-hand side."""
    n = nx * ny * nz
    idx = np.arange(n).reshape(nz, ny, nx)
    rows, cols, vals = [], [], []
    for axis, sign in ((2, 1), (2, -1), (1, 1), (1, -1), (0, 1), (0, -1)):
        src = idx
        dst = np.roll(idx, -sign, axis=axis)
        valid = np.ones_like(idx, bool)
        sl = [slice(None)] * 3
        sl[axis] = -1 if sign == 1 else 0
        valid[tuple(sl)] = False
        rows.append(src[valid])
        cols.append(dst[valid])
        vals.append(np.full(valid.sum(), -1.0 + sign * asym))
    rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    ptr_ = np.zeros(n + 1, np.int32)
    np.add.at(ptr_, rows + 1, 1)
    ptr_ = np.cumsum(ptr_).astype(np.int32)
    diag = np.full(n, 6.5)
    b = np.random.default_rng(seed).standard_normal(n)
    return ptr_, cols.astype(np.int32), diag, vals, b


def on_device(*arrays):
    return [torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0") for a in arrays]


@pytest.mark.parametrize("solver,pc,warm", [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 0, 0), (1, 1, 0), (1, 1, 1)])
def test_noalloc_equals_the_allocating_solve(solver, pc, warm):
    for shape, seed in (((48, 40, 36), 1), ((19, 13, 11), 2)):  # the second call is smaller
        P, Cc, D, V, b = lattice(*shape, asym=0.3 if solver == 1 else 0.0, seed=seed)
        x0 = np.random.default_rng(seed + 10).standard_normal(len(D)) if warm else np.zeros(len(D))
        cfg = solver_config(solver=solver, pc=pc, rtol=1e-10, atol=1e-30, max_iter=2000, warm_start=warm)
        x1, r1, h1 = csr_solve(P, Cc, D, V, b, x0, cfg)
        dP, dC, dD, dV, db, dx = on_device(P, Cc, D, V, b, x0)
        torch.cuda.synchronize()
        r2, h2 = csr_solve_noalloc(dP, dC, dD, dV, db, dx, cfg,
                                   stream=torch.cuda.current_stream().cuda_stream)
        x2 = dx.cpu().numpy()
        print(f"shape {shape} solver {solver} pc {pc} warm {warm}: {r2.iterations} iterations", flush=True)
        assert r1.status == r2.status == 0
        assert r1.iterations == r2.iterations > 0
        assert r1.final_residual == r2.final_residual
        np.testing.assert_array_equal(h2, h1)
        np.testing.assert_array_equal(x2, x1)


def test_noalloc_rejects_what_it_cannot_hold():
    P, Cc, D, V, b = lattice(6, 5, 4, asym=0.0, seed=3)
    dP, dC, dD, dV, db, dx = on_device(P, Cc, D, V, b, np.zeros(len(D)))
    res, hist, ln = SolveResult(), np.zeros(8), C.c_int()

    def call(n, cfg, col=dC):
        return lib().bsim_csr_solve_noalloc(0, None, n, dP.data_ptr(), col.data_ptr(), dD.data_ptr(), dV.data_ptr(),
                                            db.data_ptr(), dx.data_ptr(), C.byref(cfg), C.byref(res),
                                            hist.ctypes.data_as(C.POINTER(C.c_double)), 8, C.byref(ln))

    ok = solver_config(solver=0, pc=1, rtol=1e-10, atol=1e-30)
    # Rejected before any device array is read.
    assert call(MAX_ROWS + 1, ok) == 2
    assert b"at most 524288" in lib().bsim_last_error()
    assert call(len(D), solver_config(solver=0, pc=2)) == 2
    assert call(len(D), solver_config(solver=1, pc=3)) == 2  # block Jacobi keeps factors: not in the fixed scratch
    assert call(len(D), solver_config(solver=1, pc=1, precision=1)) == 2
    assert call(len(D), solver_config(solver=2, pc=1)) == 2  # the GMRES basis does not fit the fixed scratch
    # A column outside [0, n), found by the device check (every read in bounds).
    bad = Cc.copy()
    bad[5] = len(D) + 7
    (dbad,) = on_device(bad)
    torch.cuda.synchronize()
    assert call(len(D), ok, col=dbad) == 2
    assert b"column index out of range" in lib().bsim_last_error()
    # The entry point still works after the rejections.
    torch.cuda.synchronize()
    assert call(len(D), ok) == 0 and res.status == 0
