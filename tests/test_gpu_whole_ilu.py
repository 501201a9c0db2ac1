"""Device icc (IC(0) of the whole matrix) and asm (ILU(0) of the whole matrix,
PCASM on one rank) against the oracle (src/backend/petsc_linear_system.cpp:
204-208, 224-226).  Each row's elimination and substitution runs in the
oracle's order whatever the level schedule, so residual histories agree to
1e-10 relative, iteration counts and statuses exactly, and solutions to 1e-9
in relative L2.  The 300k-row grid has levels wider than one workgroup's run,
so it takes the grid-launched level kernels as well as the single-workgroup
runs."""
import numpy as np
import pytest

from babelsim_amd import solver_config
from babelsim_amd.device import Context, System, csr_solve
import oracle_lib as orc
from test_gpu_parity import laplace_system, rel_l2
from test_oracle_gmres import convection_diffusion, csr_of
from test_oracle_whole_ilu import csr, transport_system, tridiagonal

pytestmark = pytest.mark.gpu

CG, BCGS, GMRES, FGMRES = 0, 1, 2, 3
ICC, ASM = 4, 5


def both(P, Cc, D, Vv, b, cfg):
    o = orc.csr_solve(P, Cc, D, Vv, b, np.zeros(len(b)), cfg)
    d = csr_solve(P, Cc, D, Vv, b, np.zeros(len(b)), cfg)
    return o, d


def same(o, d, sol=1e-9):
    (x1, r1, h1), (x2, r2, h2) = o, d
    assert r1.status == r2.status
    assert r1.iterations == r2.iterations
    assert len(h1) == len(h2)
    np.testing.assert_allclose(h2, h1, rtol=1e-10, atol=0)
    if r1.status == 0:
        assert rel_l2(x2, x1) < sol


@pytest.mark.parametrize("solver", [CG, BCGS, GMRES, FGMRES])
@pytest.mark.parametrize("pc", [ICC, ASM])
def test_pressure_system_histories(solver, pc):
    """SPD diffusion on a jittered polyhedral mesh (1320 cells, oracle-assembled)."""
    arrays, om, diag, up, lo, b = laplace_system()
    P, Cc, D, Vv = orc.csr_from_ldu(arrays, diag, up, lo)
    cfg = solver_config(solver=solver, pc=pc, rtol=1e-10, atol=1e-30, max_iter=1000)
    o, d = both(P, Cc, D, Vv, b, cfg)
    print(f"solver {solver} pc {pc}: {o[1].iterations} iterations", flush=True)
    assert o[1].status == 0
    same(o, d)


@pytest.mark.parametrize("solver", [BCGS, GMRES])
def test_non_symmetric_asm(solver):
    A = convection_diffusion(m=40, peclet=4.0)
    P, Cc, D, Vv = csr_of(A)
    b = np.sin(np.arange(A.shape[0]) * 0.37)
    o, d = both(P, Cc, D, Vv, b, solver_config(solver=solver, pc=ASM, rtol=1e-10, atol=1e-30, max_iter=1000))
    assert o[1].status == 0
    same(o, d)
    np.testing.assert_allclose(d[0], np.linalg.solve(A, b), rtol=1e-7, atol=1e-9)


def grid_laplacian(nx, ny, nz, seed=1):
    """7-point operator with random positive couplings on an nx x ny x nz
    grid in natural order (levels are the planes i + j + k = const)."""
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    idx = np.arange(n).reshape(nz, ny, nx)
    rows, cols = [], []
    for a, b in ((idx[:, :, :-1], idx[:, :, 1:]), (idx[:, :-1, :], idx[:, 1:, :]), (idx[:-1, :, :], idx[1:, :, :])):
        rows.append(a.ravel())
        cols.append(b.ravel())
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    w = 0.5 + rng.random(len(r))
    R = np.concatenate([r, c])
    Cc = np.concatenate([c, r])
    V = -np.concatenate([w, w])
    order = np.lexsort((Cc, R))
    R, Cc, V = R[order], Cc[order], V[order]
    diag = np.zeros(n)
    np.add.at(diag, R, -V)
    diag += 0.01
    ptr = np.zeros(n + 1, np.int64)
    np.add.at(ptr, R + 1, 1)
    return np.cumsum(ptr).astype(np.int32), Cc.astype(np.int32), diag, V


@pytest.mark.parametrize("solver,pc", [(CG, ICC), (BCGS, ASM)])
def test_grid_launched_levels(solver, pc):
    P, Cc, D, Vv = grid_laplacian(100, 100, 30)
    b = np.cos(np.arange(len(D)) * 0.001)
    o, d = both(P, Cc, D, Vv, b, solver_config(solver=solver, pc=pc, rtol=1e-9, atol=1e-30, max_iter=2000))
    print(f"300k grid solver {solver} pc {pc}: {o[1].iterations} iterations", flush=True)
    assert o[1].status == 0
    same(o, d)


@pytest.mark.parametrize("pc", [ICC, ASM])
def test_failed_pivot(pc):
    A = tridiagonal(40, symmetric=True)
    if pc == ASM:
        A[0, 0] = 0.0
    else:
        A[5, 5] = -1.0
    for solver in (CG, BCGS, GMRES):
        o, d = both(*csr(A), np.ones(40), solver_config(solver=solver, pc=pc, max_iter=100))
        assert o[1].status == 2
        same(o, d)


@pytest.mark.parametrize("scale", [1e-40, 1.0, 1e40])
@pytest.mark.parametrize("solver,pc", [(CG, ICC), (BCGS, ASM), (GMRES, ICC)])
def test_scale_contract(scale, solver, pc):
    """numerical_contract_test.cpp:129-176 on the device: the same
    decisions as the oracle, converged to u = x."""
    A, b, xs = transport_system(scale, convective=(solver != CG))
    cfg = solver_config(solver=solver, pc=pc, atol=scale * 1e-13, rtol=1e-10, max_iter=2000)
    o, d = both(*csr(A), b, cfg)
    assert d[1].status == 0
    same(o, d)
    assert np.max(np.abs(d[0] - xs)) < 1e-8


@pytest.mark.parametrize("solver,pc", [(CG, ICC), (BCGS, ASM)])
def test_mesh_system_solve_contract(solver, pc):
    """The persistent system on the mesh layout (sliced rows, renumbered
    cells) eliminates in the caller's cell order, so its factors are those of
    the oracle's CSR system in that order: the same iteration count and
    residual history up to the renumbered inner products' rounding, the
    oracle's solution; a second right-hand side reuses the factors; the
    performance record names the preconditioner."""
    arrays, om, diag, up, lo, b = laplace_system(anchor=True)
    cfg = solver_config(solver=solver, pc=pc, rtol=1e-10, atol=1e-30)
    xo, ro, ho = orc.csr_solve(*orc.csr_from_ldu(arrays, diag, up, lo), b, np.zeros(len(b)), cfg)
    sys = System(Context(arrays, renumber=True))
    x, res = sys.solve(diag, up, lo, b, np.zeros_like(b), cfg)
    hd = sys.history()
    m = min(len(hd), len(ho))
    dev = np.abs(hd[:m] - ho[:m]) / ho[:m]
    print(f"solver {solver} pc {pc}: {res[0].iterations} / {ro.iterations} iterations, "
          f"history rel diff max {dev.max():.2e}, first 10 {dev[:10].max():.2e}", flush=True)
    assert res[0].status == 0
    if solver == CG:
        # CG's residuals are smooth in the rounding of its inner products.
        assert res[0].iterations == ro.iterations
        np.testing.assert_allclose(hd, ho, rtol=1e-8)
    else:
        # BiCGSTAB's residuals are not monotone: reordered inner products may
        # move the crossing of the tolerance by an iteration.
        assert abs(res[0].iterations - ro.iterations) <= 1
        np.testing.assert_allclose(hd[:10], ho[:10], rtol=1e-8)
    assert rel_l2(x, xo) < 1e-9
    x2, res2 = sys.solve(diag, up, lo, 2 * b, np.zeros_like(b), cfg)
    assert res2[0].status == 0
    assert rel_l2(x2, 2 * xo) < 1e-8
    c = sys.counters()
    assert c["solves"] == 2 and c["rhs_only"] == 1 and c["pc_setups"] == 1
    assert sys.performance()["pc_type"] == ("icc" if pc == ICC else "asm")
