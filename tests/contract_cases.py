"""The linear-system contract scenarios of the reference's own tests, written
once for any backend that offers a persistent system fed LDU arrays
(ComputeBackend::solve): the oracle (tests/test_oracle_contract.py) and the
HIP library (tests/test_gpu_contract.py).

`new_system(arrays)` returns an object with
  solve(diag, upper, lower, source, x, cfg) -> (x, [SolveResult...])
  counters() -> {matrix_updates, rhs_only, pc_setups, pc_reuses, retries,
                 solves, krylov_iterations, true_residual_checks, ...}

The expectations are the reference's: tests/petsc_backend_test.cpp:23-235,
271-374 and tests/residual_restart_test.cpp:16-84.  The reference's PETSc
preconditioners (BlockJacobi, Hypre, GAMG) are replaced by the two this
backend has (Jacobi, aggregation AMG); the counter and solution expectations
are unchanged."""
from __future__ import annotations

import numpy as np

from babelsim_amd import MeshData, solver_config

CG, BCGS = 0, 1
PC_NONE, PC_JACOBI, PC_AMG = 0, 1, 2


def box(n=(4, 2, 2), hi=(1.0, 1.0, 1.0)) -> dict:
    """makeHexBox({4,2,2}, {0,0,0}, {1,1,1}) of petsc_backend_test.cpp:245."""
    return MeshData.hex(n, (0, 0, 0), hi).arrays()


def ncells(arrays) -> int:
    return int(max(arrays["owner"].max(), arrays["neighbour"].max()) + 1)


def exact(arrays):
    """1 + 0.01 gid (petsc_backend_test.cpp:23-25)."""
    return 1.0 + 0.01 * np.arange(ncells(arrays))


def manufactured(arrays, scale=1.0, d=6.0, off=-1.0):
    """assembleManufactured (petsc_backend_test.cpp:27-45): diagonal 6,
    off-diagonals -1 on every interior face, source = A x_exact."""
    own, nei = arrays["owner"], arrays["neighbour"]
    nc, nf = ncells(arrays), len(own)
    e = exact(arrays)
    diag = np.full(nc, d)
    upper = np.zeros(nf)
    lower = np.zeros(nf)
    interior = nei >= 0
    upper[interior] = off
    lower[interior] = off
    src = d * e
    np.add.at(src, own[interior], off * e[nei[interior]])
    np.add.at(src, nei[interior], off * e[own[interior]])
    return diag, upper, lower, scale * src


def cg(pc, **kw):
    base = dict(solver=CG, pc=pc, atol=1e-13, rtol=1e-12, max_iter=200)
    base.update(kw)
    return solver_config(**base)


def check_field(x, want, tol):
    assert np.max(np.abs(x - want)) <= tol, float(np.max(np.abs(x - want)))


def preconditioner_reuse(new_system, pc):
    """verifyPreconditionerReuse (petsc_backend_test.cpp:99-181)."""
    a = box()
    e = exact(a)
    nc = len(e)
    sys = new_system(a)
    cfg = cg(pc, rebuild_interval=3)
    x = np.zeros(nc)
    for version in range(7):
        d, u, lo, b = manufactured(a)
        f = 1.0 + 0.01 * version
        x, (r,) = sys.solve(f * d, f * u, f * lo, f * b, x, cfg)
        assert r.status == 0
        check_field(x, e, 5e-11)
    c = sys.counters()
    assert c["matrix_updates"] == 7 and c["pc_setups"] == 3, c
    assert c["pc_reuses"] == 4 and c["retries"] == 0, c
    # Warm start from the exact solution: the true initial residual, no Krylov.
    warm = cg(pc, rebuild_interval=3, warm_start=1)
    d, u, lo, b = manufactured(a)
    f = 1.06
    x, (r,) = sys.solve(f * d, f * u, f * lo, f * b, e.copy(), warm)
    assert r.status == 0 and r.iterations == 0 and r.initial < 1e-10, (r.status, r.iterations, r.initial)
    # A change of the matrix (one rank's rows in the reference: all rows here).
    d2, b2 = f * d + 0.1, f * b + 0.1 * e
    x, (r,) = sys.solve(d2, f * u, f * lo, b2, x, warm)
    assert r.status == 0
    check_field(x, e, 5e-11)
    # Several right-hand sides on an unchanged matrix keep its preconditioner;
    # once the matrix changes, their accumulated work asks for a fresh one.
    sys = new_system(a)
    x = np.zeros(nc)
    for rhs in range(1, 6):
        d, u, lo, b = manufactured(a, float(rhs))
        x, (r,) = sys.solve(d, u, lo, b, np.zeros(nc), cfg)
        assert r.status == 0
        check_field(x, rhs * e, 3e-10)
    c = sys.counters()
    assert c["pc_setups"] == 1 and c["matrix_updates"] == 1, c
    d, u, lo, b = manufactured(a, 6.0)
    x, (r,) = sys.solve(d + 0.1, u, lo, b + 0.6 * e, np.zeros(nc), cfg)
    assert r.status == 0
    check_field(x, 6.0 * e, 4e-10)
    c = sys.counters()
    assert c["pc_setups"] == 2 and c["matrix_updates"] == 2 and c["pc_reuses"] == 0, c


def diagonal_system(a, values):
    nf = len(a["owner"])
    e = exact(a)
    return values, np.zeros(nf), np.zeros(nf), values * e


def retry_after_stale_preconditioner(new_system):
    """petsc_backend_test.cpp:183-209: a lagged Jacobi preconditioner fails
    within two iterations, is rebuilt and the solve retried; the failed
    attempt's iterations and true-residual check are kept."""
    a = box()
    e = exact(a)
    nc = len(e)
    sys = new_system(a)
    cfg = cg(PC_JACOBI, max_iter=2, rebuild_interval=8)
    x, (r,) = sys.solve(*diagonal_system(a, np.full(nc, 2.0)), np.zeros(nc), cfg)
    assert r.status == 0, "fresh Jacobi solve failed"
    x, (r,) = sys.solve(*diagonal_system(a, 3.0 + np.arange(nc)), np.zeros(nc), cfg)
    assert r.status == 0, "failed stale PC was not rebuilt and retried"
    check_field(x, e, 5e-11)
    c = sys.counters()
    assert c["retries"] == 1 and c["pc_setups"] == 2 and c["true_residual_checks"] == 3, c
    assert r.iterations == 3, r.iterations


def rebuild_after_iteration_growth(new_system):
    """petsc_backend_test.cpp:211-234: a lagged solve that costs more than
    twice the fresh one asks for a new preconditioner before the next
    right-hand side, with no matrix update in between."""
    a = box()
    e = exact(a)
    nc = len(e)
    sys = new_system(a)
    cfg = cg(PC_JACOBI, rebuild_interval=8)
    x, (r,) = sys.solve(*diagonal_system(a, np.full(nc, 2.0)), np.zeros(nc), cfg)
    assert r.iterations == 1, "fresh diagonal solve should require one CG iteration"
    grown = diagonal_system(a, 3.0 + np.arange(nc))
    x, (r,) = sys.solve(*grown, np.zeros(nc), cfg)
    assert r.status == 0 and r.iterations > 2, r.iterations
    x, (r,) = sys.solve(*grown, np.zeros(nc), cfg)
    assert r.status == 0 and r.iterations == 1, r.iterations
    c = sys.counters()
    assert c["pc_setups"] == 2 and c["retries"] == 0 and c["matrix_updates"] == 2, c
    check_field(x, e, 5e-11)


def scalar_and_vector_solves(new_system, pc):
    """The manufactured scalar solve, its right-hand-side-only repeat and the
    three-component vector solve (petsc_backend_test.cpp:271-336), with the
    per-system counters of :366-370 (two systems: 2 updates, 1 rhs-only
    solve, 5 true-residual checks in all)."""
    a = box()
    e = exact(a)
    nc = len(e)
    cfg = cg(pc)
    scalar = new_system(a)
    d, u, lo, b = manufactured(a)
    x, (r,) = scalar.solve(d, u, lo, b, np.zeros(nc), cfg)
    assert r.status == 0
    check_field(x, e, 2e-11)
    x, (r,) = scalar.solve(d, u, lo, 2.0 * b, x, cfg)
    assert r.status == 0
    check_field(x, 2.0 * e, 4e-11)
    vector = new_system(a)
    own, nei = a["owner"], a["neighbour"]
    inner = nei >= 0
    src = np.stack([6.0 * e, -12.0 * e, 3.0 * e], 1)
    for row, col in ((own[inner], nei[inner]), (nei[inner], own[inner])):
        np.add.at(src[:, 0], row, -e[col])
        np.add.at(src[:, 1], row, 2.0 * e[col])
        np.add.at(src[:, 2], row, -0.5 * e[col])
    X, res = vector.solve(d, u, lo, src, np.zeros((nc, 3)), cfg)
    assert all(r.status == 0 for r in res)
    assert np.max(np.abs(X[:, 0] - e)) < 2e-11
    assert np.max(np.abs(X[:, 1] + 2.0 * e)) < 4e-11
    assert np.max(np.abs(X[:, 2] - 0.5 * e)) < 1e-11
    cs, cv = scalar.counters(), vector.counters()
    assert cs["matrix_updates"] + cv["matrix_updates"] == 2
    assert cs["rhs_only"] + cv["rhs_only"] == 1
    assert cs["true_residual_checks"] + cv["true_residual_checks"] == 5


def repeated_face_pairs(new_system, split_arrays, pc):
    """Two faces joining the same cell pair add both contributions
    (petsc_backend_test.cpp:338-360): diagonal 2, off-diagonals -0.5."""
    a = split_arrays
    e = exact(a)
    d, u, lo, b = manufactured(a, d=2.0, off=-0.5)
    x, (r,) = new_system(a).solve(d, u, lo, b, np.zeros(len(e)), cg(pc))
    assert r.status == 0
    assert np.max(np.abs(x - e)) < 1e-12


def residual_restarts(new_system, seconds=None):
    """tests/residual_restart_test.cpp:16-84 on its 64x4x1 box: strict
    absolute targets are met or honestly rejected against the true residual
    with the original target and one iteration budget; an accepted warm guess
    skips the Krylov solve (solve seconds unchanged when `seconds` reads the
    backend's Krylov time)."""
    a = box((64, 4, 1))
    own, nei = a["owner"], a["neighbour"]
    nc = ncells(a)
    gid = np.arange(nc)
    ex = 1.25 + np.sin(0.13 * gid)
    inner = nei >= 0
    diag = np.full(nc, 0.02)
    np.add.at(diag, own[inner], 1.0)
    np.add.at(diag, nei[inner], 1.0)
    upper = np.where(inner, -1.0, 0.0)
    lower = upper.copy()
    src = diag * ex
    np.add.at(src, own[inner], -ex[nei[inner]])
    np.add.at(src, nei[inner], -ex[own[inner]])
    sys = new_system(a)
    repaired = rejected = 0
    for target in (1e-12, 1e-13, 6e-14, 4e-14, 2e-14, 1e-14):
        cfg = solver_config(solver=CG, pc=PC_NONE, atol=target, rtol=1e-30, max_iter=1000)
        before = sys.counters()["true_residual_checks"]
        x, (r,) = sys.solve(diag, upper, lower, src, np.zeros(nc), cfg)
        after = sys.counters()["true_residual_checks"]
        assert r.iterations <= 1000
        assert (r.status == 0) == (r.final_residual <= target)
        if target >= 1e-13:
            assert r.status == 0, "attainable manufactured target failed"
        if r.status == 0:
            repaired += after > before + 1
            assert np.max(np.abs(x - ex)) < 1e-9
        else:
            rejected += 1
            assert np.isfinite(r.final_residual)
    warm = solver_config(solver=CG, pc=PC_NONE, atol=1e-12, rtol=1e-30, max_iter=1000, warm_start=1)
    t0 = seconds(sys) if seconds else None
    before = sys.counters()["true_residual_checks"]
    x, (r,) = sys.solve(diag, upper, lower, src, ex.copy(), warm)
    assert r.status == 0 and r.iterations == 0 and r.final_residual == r.initial
    # The accepted initial residual is the one check (petsc_linear_system.cpp:418-427).
    assert sys.counters()["true_residual_checks"] == before + 1
    if seconds:
        assert seconds(sys) == t0, "accepted warm start performed a Krylov solve"
    return repaired, rejected


def reuses_count_per_right_hand_side(new_system, pc):
    """A lagged preconditioner counts one reuse per right-hand side solved
    with it (petsc_backend.cpp:262-266, petsc_linear_system.cpp:413-414): three
    for a vector equation, and again for a right-hand-side-only repeat."""
    a = box()
    e = exact(a)
    nc = len(e)
    sys = new_system(a)
    cfg = cg(pc, rebuild_interval=8)
    d, u, lo, b = manufactured(a)
    B = np.stack([b, 2 * b, -b], 1)
    X, res = sys.solve(d, u, lo, B, np.zeros((nc, 3)), cfg)
    assert all(r.status == 0 for r in res)
    assert sys.counters()["pc_reuses"] == 0
    X, res = sys.solve(1.01 * d, 1.01 * u, 1.01 * lo, 1.01 * B, X, cfg)
    assert all(r.status == 0 for r in res)
    c = sys.counters()
    assert c["pc_setups"] == 1 and c["pc_reuses"] == 3, c
    X, res = sys.solve(1.01 * d, 1.01 * u, 1.01 * lo, 2.02 * B, X, cfg)
    c = sys.counters()
    assert c["rhs_only"] == 1 and c["pc_reuses"] == 6 and c["solves"] == 9, c
    np.testing.assert_allclose(X, np.stack([2 * e, 4 * e, -2 * e], 1), atol=1e-10)
