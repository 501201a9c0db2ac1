"""The HIP library's persistent linear system (bsim_system_*) against the
reference's own solve-contract expectations, the scenarios of
tests/contract_cases.py that the oracle passes in test_oracle_contract.py:
preconditioner reuse and rebuild, retry of a stale preconditioner, rebuild
after iteration growth, right-hand-side-only and vector solves, repeated face
pairs, strict residual targets with honest rejection and the warm-start
shortcut (no Krylov time spent), and the counter record of
bsim_system_performance."""
import pytest

import contract_cases as cc
import meshes
from babelsim_amd.device import Context, System

pytestmark = pytest.mark.gpu


def new_system(arrays):
    return System(Context(arrays, renumber=True))


@pytest.mark.parametrize("pc", [cc.PC_JACOBI, cc.PC_AMG])
def test_preconditioner_reuse(pc):
    cc.preconditioner_reuse(new_system, pc)


def test_retry_after_stale_preconditioner():
    cc.retry_after_stale_preconditioner(new_system)


def test_rebuild_after_iteration_growth():
    cc.rebuild_after_iteration_growth(new_system)


@pytest.mark.parametrize("pc", [cc.PC_JACOBI, cc.PC_AMG])
def test_scalar_and_vector_solves(pc):
    cc.scalar_and_vector_solves(new_system, pc)


@pytest.mark.parametrize("pc", [cc.PC_JACOBI, cc.PC_AMG])
def test_repeated_face_pairs(pc):
    cc.repeated_face_pairs(new_system, meshes.split_interface(), pc)


def test_residual_restarts():
    # Which of the strictest targets FP64 CG reaches depends on the row order;
    # the contract is that each outcome agrees with the true residual.
    cc.residual_restarts(new_system, seconds=lambda s: s.performance()["solve_seconds"])


@pytest.mark.parametrize("pc", [cc.PC_JACOBI, cc.PC_AMG])
def test_reuses_count_per_right_hand_side(pc):
    cc.reuses_count_per_right_hand_side(new_system, pc)


def test_performance_record():
    """bsim_system_performance mirrors the counters, names the solver and
    preconditioner with the reference's words and keeps the last result."""
    a = cc.box()
    sys = new_system(a)
    d, u, lo, b = cc.manufactured(a)
    x, (r,) = sys.solve(d, u, lo, b, 0 * b, cc.cg(cc.PC_AMG))
    p = sys.performance()
    c = sys.counters()
    assert p["ksp_type"] == "cg" and p["pc_type"] == "amg" and p["last_status"] == "converged"
    assert p["linear_solves"] == c["solves"] == 1 and p["krylov_iterations"] == r.iterations
    assert p["matrix_value_updates"] == 1 and p["preconditioner_setups"] == 1 and p["true_residual_checks"] == 1
    assert p["last_final_residual"] == r.final_residual and p["last_initial_residual"] == r.initial
    assert p["solve_seconds"] > 0 and p["preconditioner_seconds"] > 0 and p["matrix_pattern_builds"] == 1


def test_simple_counter_sequences_equal_the_oracle():
    """Over several SIMPLE iterations on the jittered, split cavity, the
    momentum and pressure systems' counters (updates, right-hand-side-only
    solves, preconditioner setups, reuses and retries, solves, Krylov
    iterations, true-residual checks) equal the oracle's, and the run record
    sums them with the iteration window of RunTime::recordIteration."""
    import cases
    import oracle_lib as orc
    from babelsim_amd.device import Simple
    prepared, _, _ = cases.cavity3d_mesh(8).prepare()
    arrays = prepared.arrays()
    ub, pb = cases.cavity_bcs(two_d=False)
    cfg = cases.c2_config()
    gpu = Simple(Context(arrays, renumber=False), cfg, ub, pb)
    cpu = orc.OracleSimple(orc.oracle_mesh(arrays), cfg, ub, pb)
    gpu.start()
    cpu.start()
    n = 6
    gpu.iterate(n)
    cpu.iterate(n)
    total, systems = gpu.performance()
    assert [s["identity"] for s in systems] == ["vector:momentum:U", "scalar:pressureCorrection:pPrime"]
    for which, s in enumerate(systems):
        want = cpu.counters(which)
        assert {k: s[k] for k in want} == want, (which, s, want)
    assert systems[0]["linear_solves"] == 3 * n
    assert total["reported_iteration_intervals"] == n - 1 and total["iteration_compute_seconds"] > 0
    for key in ("linear_solves", "krylov_iterations", "preconditioner_setups", "true_residual_checks",
                "matrix_value_updates", "preconditioner_reuses"):
        assert total[key] == sum(s[key] for s in systems), key
    assert total["matrix_pattern_builds"] == 2 and total["global_reductions"] >= 2 * n
