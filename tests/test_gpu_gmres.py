"""Device GMRES and FGMRES against the oracle (restart 30, right
preconditioning, classical Gram-Schmidt; petsc_linear_system.cpp:180-195 with
KSPGMRES/KSPFGMRES).  The device forms every inner product, Gram-Schmidt
update and basis combination in the oracle's order and keeps the Hessenberg
scalars on the host, so residual histories agree to 1e-10 relative and the
solutions to 1e-9 in relative L2."""
import numpy as np
import pytest

from babelsim_amd import solver_config
from babelsim_amd.device import Context, System, csr_solve
import oracle_lib as orc
from test_gpu_parity import laplace_system, rel_l2
from test_oracle_gmres import convection_diffusion, csr_of

pytestmark = pytest.mark.gpu


def both(P, Cc, D, Vv, b, cfg):
    x1, r1, h1 = orc.csr_solve(P, Cc, D, Vv, b, np.zeros(len(b)), cfg)
    x2, r2, h2 = csr_solve(P, Cc, D, Vv, b, np.zeros(len(b)), cfg)
    return (x1, r1, h1), (x2, r2, h2)


def same(o, d, sol=1e-9):
    (x1, r1, h1), (x2, r2, h2) = o, d
    assert r1.status == r2.status
    assert r1.iterations == r2.iterations
    assert len(h1) == len(h2)
    np.testing.assert_allclose(h2, h1, rtol=1e-10, atol=0)
    assert rel_l2(x2, x1) < sol


@pytest.mark.parametrize("solver", [2, 3])
@pytest.mark.parametrize("pc", [0, 1, 2, 3])
def test_pressure_system_histories(solver, pc):
    """SPD diffusion system on a jittered polyhedral mesh (1444 cells): every
    preconditioner, through the restarts the Jacobi and block Jacobi runs need
    (about 470 and 300 iterations).  Unpreconditioned GMRES(30) stalls on this
    system; it is compared over its first 300 iterations."""
    arrays, om, diag, up, lo, b = laplace_system()
    P, Cc, D, Vv = orc.csr_from_ldu(arrays, diag, up, lo)
    cfg = solver_config(solver=solver, pc=pc, rtol=1e-10, atol=1e-30, max_iter=300 if pc == 0 else 3000)
    o, d = both(P, Cc, D, Vv, b, cfg)
    print(f"solver {solver} pc {pc}: {o[1].iterations} iterations", flush=True)
    assert o[1].status == (1 if pc == 0 else 0)
    same(o, d)


@pytest.mark.parametrize("solver", [2, 3])
@pytest.mark.parametrize("pc", [1, 3])
def test_non_symmetric_histories(solver, pc):
    """Upwinded convection-diffusion (a momentum-shaped, non-symmetric matrix)
    with Jacobi and block Jacobi."""
    A = convection_diffusion(m=48, peclet=4.0)
    P, Cc, D, Vv = csr_of(A)
    b = np.sin(np.arange(A.shape[0]) * 0.37)
    cfg = solver_config(solver=solver, pc=pc, rtol=1e-10, atol=1e-30, max_iter=2000)
    o, d = both(P, Cc, D, Vv, b, cfg)
    assert o[1].status == 0
    same(o, d)
    np.testing.assert_allclose(d[0], np.linalg.solve(A, b), rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("solver", [2, 3])
def test_iteration_cap_and_assembled_solution(solver):
    """Stopped by the iteration cap inside the second cycle: the same status,
    history and partial solution on both sides."""
    A = convection_diffusion(m=32, peclet=1.0)
    P, Cc, D, Vv = csr_of(A)
    b = np.ones(A.shape[0])
    cfg = solver_config(solver=solver, pc=0, rtol=1e-14, atol=1e-30, max_iter=45)
    o, d = both(P, Cc, D, Vv, b, cfg)
    assert o[1].status == 1 and o[1].iterations == 45
    same(o, d, sol=1e-10)


@pytest.mark.parametrize("solver", [2, 3])
def test_mesh_system_solve_contract(solver):
    """The persistent system on the mesh layout (sliced rows, renumbered cells):
    a scalar solve converges to the oracle's CSR solution; a second right-hand
    side reuses the preconditioner as the contract counts it."""
    arrays, om, diag, up, lo, b = laplace_system(anchor=True)
    cfg = solver_config(solver=solver, pc=2, rtol=1e-10, atol=1e-30)
    xo, ro, _ = orc.csr_solve(*orc.csr_from_ldu(arrays, diag, up, lo), b, np.zeros(len(b)), cfg)
    sys = System(Context(arrays, renumber=True))
    x, res = sys.solve(diag, up, lo, b, np.zeros_like(b), cfg)
    assert res[0].status == 0
    assert rel_l2(x, xo) < 1e-8
    x2, res2 = sys.solve(diag, up, lo, 2 * b, np.zeros_like(b), cfg)
    assert res2[0].status == 0
    assert rel_l2(x2, 2 * xo) < 1e-8
    c = sys.counters()
    assert c["solves"] == 2 and c["rhs_only"] == 1
    assert sys.performance()["ksp_type"] == ("gmres" if solver == 2 else "fgmres")
