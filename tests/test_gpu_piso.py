"""Device PISO (babelsim_amd/csrc/piso.hip) against the oracle's PISO step on
the same renumbered meshes: per-step reports (iteration counts exactly,
residual measures to 1e-8 relative), fields, the ddtPhiCorr flux and the
Krylov histories.  Plus the textbook time-integration check and the
reference's coupling-test behaviour run on the device."""
from __future__ import annotations

import numpy as np
import pytest

import cases
import oracle_lib as orc
import piso_cases as pc
from babelsim_amd import BDF2, EULER, STEP_COMPLETED, STEP_NOT_CONVERGED, MeshData, solver_config
from babelsim_amd.device import Context, Piso, TimeStepper

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def pair(mesh_data, cfg, bcs, U0=None):
    arrays, _, _ = mesh_data.prepare()
    arrays = arrays.arrays()
    ub, pb = bcs
    om = orc.oracle_mesh(arrays)
    so = orc.OraclePiso(om, cfg, ub, pb)
    ctx = Context(arrays, renumber=False)
    sg = Piso(ctx, cfg, ub, pb)
    if U0 is not None:
        U0 = U0(om.geometry()["C"])
        so.set_fields(U0, np.zeros(om.nc))
        sg.set_fields(U0, np.zeros(om.nc))
    so.start()
    sg.start()
    return so, sg


def compare_step(a, b):
    assert a.status == b.status and a.correctors == b.correctors and a.passes == b.passes
    assert list(a.mom_iters) == list(b.mom_iters) and a.p_iters == b.p_iters
    assert (a.accepted, a.converged, a.linear_ok) == (b.accepted, b.converged, b.linear_ok)
    for key in ("rU", "dU", "dP", "r_coupling_first", "r_coupling", "co_max", "co_mean"):
        assert abs(getattr(a, key) - getattr(b, key)) <= 1e-8 * abs(getattr(b, key)) + 1e-14, key
    # mass is a ratio of round-off-sized imbalances to the flux scale
    assert abs(a.mass - b.mass) <= 1e-3 * abs(b.mass) + 1e-13


CASES = {
    "cavity_bdf2": lambda: (cases.cavity2d_mesh(24), pc.cavity_piso_config(), cases.cavity_bcs(True), [0.01] * 5),
    "cavity_euler_variable": lambda: (cases.cavity2d_mesh(24), pc.cavity_piso_config(time_method=EULER),
                                      cases.cavity_bcs(True), [0.01, 0.02, 0.005, 0.01]),
    "cavity_bdf2_variable_3corr": lambda: (cases.cavity2d_mesh(20), pc.cavity_piso_config(n_correctors=3),
                                           cases.cavity_bcs(True), [0.02, 0.01, 0.015, 0.03, 0.01]),
    "polyhedral_c2": lambda: (cases.cavity3d_mesh(7, split=0.15, seed=3), pc.c2_piso_config(),
                              cases.cavity_bcs(False), [0.02, 0.01, 0.015, 0.015]),
    "polyhedral_c2_two_nonorth": lambda: (cases.cavity3d_mesh(6, split=0.2, seed=11),
                                          pc.c2_piso_config(non_orth_corr=2, n_correctors=3),
                                          cases.cavity_bcs(False), [0.05, 0.05, 0.05]),
    "channel_inlet_outlet": lambda: (pc.channel_mesh(24, 8), pc.cavity_piso_config(), pc.channel_bcs(),
                                     [0.05] * 5),
    "outer_passes_relaxed": lambda: (cases.cavity2d_mesh(16),
                                     pc.cavity_piso_config(max_iterations=3, alpha_u=0.8, tol_u=1e-3,
                                                           tol_mom=1e-3, tol_p=1e-2),
                                     cases.cavity_bcs(True), [0.01] * 3),
    # The cylinder_re100 case's pair: momentum bcgs + bjacobi, pressure cg + hypre
    # (cases/cylinder_re100/numerics/solution.bs:10-17).
    "bjacobi_momentum": lambda: (cases.cavity3d_mesh(7, split=0.15, seed=3),
                                 pc.c2_piso_config(mom=solver_config(solver=1, pc=3, rtol=1e-8, atol=1e-14)),
                                 cases.cavity_bcs(False), [0.02, 0.01, 0.015]),
    "warm_start": lambda: (cases.cavity2d_mesh(20),
                           pc.cavity_piso_config(mom=solver_config(solver=1, pc=1, rtol=1e-8, atol=1e-14,
                                                                   warm_start=1),
                                                 pre=solver_config(solver=0, pc=2, rtol=1e-8, atol=1e-14,
                                                                   warm_start=1)),
                           cases.cavity_bcs(True), [0.01] * 4),
}


@pytest.mark.parametrize("kind", list(CASES))
def test_piso_parity(kind):
    mesh, cfg, bcs, dts = CASES[kind]()
    so, sg = pair(mesh, cfg, bcs)
    for dt in dts:
        ro, rg = so.step(dt), sg.step(dt)
        compare_step(rg, ro)
        np.testing.assert_allclose(sg.temporal_flux(), so.temporal_flux(), rtol=1e-7,
                                   atol=1e-9 * max(np.abs(so.temporal_flux()).max(), 1e-300))
        if ro.status != STEP_COMPLETED:
            break
    Ug, pg, phig = sg.fields()
    Uo, po, phio = so.fields()
    assert rel_l2(Ug, Uo) < 1e-9
    assert rel_l2(pg, po) < 1e-9
    assert rel_l2(phig, phio) < 1e-9
    for which in range(4):
        hg, ho = sg.history(which), so.history(which)
        assert len(hg) == len(ho)
        np.testing.assert_allclose(hg, ho, rtol=1e-8, atol=1e-300)


def test_piso_adaptive_correctors_parity():
    """couplingTolerance below the first corrector's residual: the device takes
    the same number of extra correctors as the oracle."""
    mesh = cases.cavity2d_mesh(24)
    base = pc.cavity_piso_config(n_correctors=1)
    so, _ = pair(mesh, base, cases.cavity_bcs(True))
    tol = 0.3 * so.step(0.01).r_coupling
    cfg = pc.cavity_piso_config(n_correctors=1, coupling_tol=tol, max_correctors=20)
    so, sg = pair(mesh, cfg, cases.cavity_bcs(True))
    for k in range(3):
        ro, rg = so.step(0.01), sg.step(0.01)
        compare_step(rg, ro)
        assert rg.r_coupling <= tol
        if k == 0:
            assert rg.correctors > 1


def test_piso_coupling_rejection_on_device():
    """tests/piso_coupling_test.py:43-71 on the device: the cases/cavity 64 x 64
    step is rejected with one corrector at half its coupling residual and
    accepted with up to 30 correctors."""
    mesh = cases.cavity2d_mesh(64)
    arrays, _, _ = mesh.prepare()
    arrays = arrays.arrays()
    ub, pb = cases.cavity_bcs(True)
    ctx = Context(arrays, renumber=False)
    s = Piso(ctx, pc.cavity_piso_config(n_correctors=1), ub, pb)
    s.start()
    tol = 0.5 * s.step(0.01).r_coupling
    assert tol > 1e-10
    for adaptive in (False, True):
        s = Piso(ctx, pc.cavity_piso_config(n_correctors=1, coupling_tol=tol, max_correctors=30 if adaptive else 1),
                 ub, pb)
        s.start()
        r = s.step(0.01)
        if adaptive:
            assert r.status == STEP_COMPLETED and r.accepted and r.correctors > 1
            assert r.r_coupling <= tol and r.co_max > 0
        else:
            assert r.status == STEP_NOT_CONVERGED and not r.accepted and r.mass < 1e-8 and r.linear_ok


@pytest.mark.parametrize("method", [EULER, BDF2])
def test_piso_shear_decay_textbook_on_device(method):
    ny, nu, amp = 32, 0.1, 1e-6
    dts = [0.05, 0.025, 0.05, 0.04, 0.06]
    mesh = pc.shear_mesh(ny)
    ctx = Context(mesh, renumber=True)
    s = Piso(ctx, pc.shear_config(nu, method), *pc.shear_bcs())
    y = ctx.geometry()["C"][:, 1]
    U0 = np.zeros((ctx.nc, 3))
    U0[:, 0] = amp * np.sin(np.pi * y)
    s.set_fields(U0, np.zeros(ctx.nc))
    s.start()
    for dt in dts:
        assert s.step(dt).status == STEP_COMPLETED
    U, p, _ = s.fields()
    order = np.argsort(y)
    expect = pc.textbook_steps(pc.shear_operator(ny, nu), U0[order, 0], dts, method)
    assert np.abs(U[order, 0] - expect).max() < 1e-6 * amp
    assert np.abs(p).max() == 0.0


def test_piso_run_follows_time_stepper():
    """Piso.run takes the TimeStepper's steps (src/io/case.cpp:391-403): the
    last step is shortened to land on endTime."""
    mesh = cases.cavity2d_mesh(12)
    ctx = Context(mesh)
    s = Piso(ctx, pc.cavity_piso_config(), *cases.cavity_bcs(True))
    s.start()
    reps = s.run(0.0, 0.025, 0.01)
    assert [r.dt for r in reps] == pytest.approx([0.01, 0.01, 0.005], rel=1e-12)
    assert all(r.status == STEP_COMPLETED for r in reps)
    t = TimeStepper(0.0, 0.03, 0.01)
    while not t.finished():
        t.advance()
    assert t.value == 0.03 and t.step == 3


def test_piso_mixed_precision_momentum():
    """FP32 lockstep BiCGSTAB inside FP64 refinement for the momentum solves
    and FP32 AMG for pressure: the fields after five BDF2 steps agree with FP64
    to 1e-6 relative L2."""
    mesh = cases.cavity3d_mesh(8, split=0.1, seed=7)
    out = []
    for prec in (0, 1):
        cfg = pc.c2_piso_config(mom=solver_config(solver=1, pc=1, rtol=1e-8, atol=1e-14, precision=prec),
                                pre=solver_config(solver=0, pc=2, rtol=1e-8, atol=1e-14, precision=prec))
        ctx = Context(mesh)
        s = Piso(ctx, cfg, *cases.cavity_bcs(False))
        s.start()
        for _ in range(5):
            assert s.step(0.02).status == STEP_COMPLETED
        out.append(s.fields())
    for a, b in zip(out[1], out[0]):
        assert rel_l2(a, b) < 1e-6


# initializePotentialFlow (piso/main.cpp:127-150) on device and oracle.
POTENTIAL = {
    "channel_from_rest": lambda: (pc.channel_mesh(24, 8), pc.cavity_piso_config(initialize_potential_flow=1),
                                  pc.channel_bcs(), None, [0.05] * 4),
    "polyhedral_corrected": lambda: (cases.cavity3d_mesh(6, split=0.2, seed=11),
                                     pc.c2_piso_config(non_orth_corr=2, initialize_potential_flow=1,
                                                       initial_non_orth_corr=1),
                                     cases.cavity_bcs(False),
                                     lambda C: np.stack([np.sin(3 * C[:, 1]), np.cos(2 * C[:, 0]), 0.3 * C[:, 2]], 1),
                                     [0.02] * 3),
}


@pytest.mark.parametrize("kind", list(POTENTIAL))
def test_potential_flow_parity(kind):
    """The initial projection's record and fields, then the steps after it."""
    mesh, cfg, bcs, U0, dts = POTENTIAL[kind]()
    so, sg = pair(mesh, cfg, bcs, U0)
    io, ig = so.initial(), sg.initial()
    print(f"{kind}: {io.initial_pressure_solves} potential solves, {io.p_iters} CG iterations", flush=True)
    assert (ig.done, ig.status, ig.initial_pressure_solves, ig.p_iters) == \
        (io.done, io.status, io.initial_pressure_solves, io.p_iters) == (1, STEP_COMPLETED, io.initial_pressure_solves,
                                                                          io.p_iters)
    assert io.p_iters > 0
    assert ig.initial_u_max == pytest.approx(io.initial_u_max, rel=1e-10)
    assert abs(ig.initial_mass - io.initial_mass) <= 1e-3 * io.initial_mass + 1e-13
    for a, b in zip(sg.fields(), so.fields()):
        assert rel_l2(a, b) < 1e-10
    for dt in dts:
        ro, rg = so.step(dt), sg.step(dt)
        compare_step(rg, ro)
        assert ro.status == STEP_COMPLETED
    for a, b in zip(sg.fields(), so.fields()):
        assert rel_l2(a, b) < 1e-9


def test_potential_flow_failure_ends_the_run():
    """A potential solve that does not converge: the record says numerical
    failure and every step reports it, as runPiso returns numericalFailure."""
    cfg = pc.c2_piso_config(initialize_potential_flow=1,
                            pre=solver_config(solver=0, pc=0, rtol=1e-14, atol=1e-30, max_iter=1))
    so, sg = pair(cases.cavity3d_mesh(5, split=0.1, seed=2), cfg, cases.cavity_bcs(False),
                  lambda C: np.tile([0.3, -0.2, 0.1], (len(C), 1)))
    assert sg.initial().status == so.initial().status == 2
    assert sg.step(0.01).status == so.step(0.01).status == 2


# Transient k-omega inside PISO (piso/main.cpp:98-121, 165, 176-182, 251-266).
def test_transient_komega_decay_on_device():
    """The closed box at rest of the oracle's known answer: uniform k and
    omega follow the implicit step of dk/dt = -beta* k omega, domega/dt =
    -beta omega^2 (positive root for omega, Euler then variable-step BDF2)."""
    from babelsim_amd import komega_config
    arrays, _, _ = MeshData.hex((4, 3, 2)).prepare()
    ctx = Context(arrays.arrays(), renumber=True)
    ub, pb = cases.cavity_bcs(False)
    ub = (ub[0], np.zeros_like(ub[1]))
    zg = (np.full(6, pc.ZG, np.int32), np.zeros(6))
    kw = komega_config(tolerance=1e-13, k_solver=solver_config(solver=1, pc=1, rtol=1e-14, atol=1e-30),
                       w_solver=solver_config(solver=1, pc=1, rtol=1e-14, atol=1e-30))
    dts = [0.05, 0.1, 0.07]
    s = Piso(ctx, pc.cavity_piso_config(time_method=BDF2), ub, pb)
    s.enable_komega(kw, zg, zg)
    s.set_turbulence(np.full(ctx.nc, 0.3), np.full(ctx.nc, 2.0))
    s.start()
    kp, wp, ko, wo, dtp = 0.3, 2.0, 0.3, 2.0, 0.0
    for n, dt in enumerate(dts):
        r = s.step(dt)
        assert r.status == STEP_COMPLETED and r.turb_converged == 1
        if n > 0:
            q = dt / dtp
            a0, a1, a2 = (1 + 2 * q) / ((1 + q) * dt), (1 + q) / dt, -q * q / ((1 + q) * dt)
        else:
            a0, a1, a2 = 1 / dt, 1 / dt, 0.0
        rw, rk = a1 * wp + a2 * wo, a1 * kp + a2 * ko
        wn = (-a0 + np.sqrt(a0 * a0 + 4 * kw.beta * rw)) / (2 * kw.beta)
        kn = rk / (a0 + kw.beta_star * wn)
        k, w, _ = s.turbulence()
        np.testing.assert_allclose(w, wn, rtol=1e-10)
        np.testing.assert_allclose(k, kn, rtol=1e-10)
        ko, wo, kp, wp, dtp = kp, wp, kn, wn, dt


TRANSIENT_KOMEGA = {
    # the URANS cylinder case's coupling: iterated, eight transport corrections
    "iterated_8": dict(coupling=0, correctors=8, relaxation=0.7, max_iterations=1, tol=3e-3, tol_u=1e-4),
    # iterated until the transport settles (up to turbulenceMaxIterations)
    "iterated_settle": dict(coupling=0, correctors=0, relaxation=0.7, max_iterations=1, tol=1e-6, tol_u=1e-4),
    "segregated": dict(coupling=1, correctors=0, relaxation=1.0, max_iterations=1, tol=1e-6, tol_u=1e-4),
    "outer_passes": dict(coupling=0, correctors=0, relaxation=0.7, max_iterations=4, tol=5e-2, tol_u=2e-2),
}


@pytest.mark.parametrize("kind", list(TRANSIENT_KOMEGA))
def test_transient_komega_parity(kind):
    """A walled channel with inlet turbulence (the steady k-omega tests'
    boundary values): device and oracle steps agree in every count and to
    1e-9 in the fields."""
    from babelsim_amd import komega_config
    o = TRANSIENT_KOMEGA[kind]
    mesh = MeshData.hex((16, 10, 1), (0, 0, 0), (2.0, 1.0, 0.1))
    ub = (np.array([pc.FV, pc.ZG, pc.FV, pc.FV, pc.SYM, pc.SYM], np.int32), np.array([[1.0, 0, 0]] + [[0, 0, 0]] * 5))
    pb = (np.array([pc.ZG, pc.FV, pc.ZG, pc.ZG, pc.SYM, pc.SYM], np.int32), np.zeros(6))
    kb = (np.array([pc.FV, pc.ZG, pc.FV, pc.FV, pc.ZG, pc.ZG], np.int32), np.array([1e-3, 0, 1e-10, 1e-10, 0, 0]))
    wb = (np.array([pc.FV, pc.ZG, pc.FV, pc.FV, pc.ZG, pc.ZG], np.int32), np.array([1.0, 0, 200.0, 200.0, 0, 0]))
    cfg = pc.cavity_piso_config(mu=1e-3, max_iterations=o["max_iterations"], tol_u=o["tol_u"], tol_mom=o["tol_u"],
                                tol_p=10 * o["tol_u"])
    kw = komega_config(relaxation=o["relaxation"], tolerance=o["tol"])
    arrays, _, _ = mesh.prepare()
    arrays = arrays.arrays()
    om = orc.oracle_mesh(arrays)
    so = orc.OraclePiso(om, cfg, ub, pb)
    sg = Piso(Context(arrays, renumber=False), cfg, ub, pb)
    for s in (so, sg):
        s.enable_komega(kw, kb, wb, coupling=o["coupling"], correctors=o["correctors"])
        s.set_turbulence(np.full(om.nc, 1e-3), np.full(om.nc, 1.0))
        s.start()
    for dt in [0.02] * 4:
        ro, rg = so.step(dt), sg.step(dt)
        compare_step(rg, ro)
        assert (rg.k_iters, rg.w_iters, rg.turb_corrections, rg.turb_linear, rg.turb_converged) == \
            (ro.k_iters, ro.w_iters, ro.turb_corrections, ro.turb_linear, ro.turb_converged)
        for key in ("d_turb", "r_turb"):
            assert abs(getattr(rg, key) - getattr(ro, key)) <= 1e-8 * abs(getattr(ro, key)) + 1e-14, key
        assert ro.status == STEP_COMPLETED
    print(f"{kind}: last step {ro.passes} passes, {ro.turb_corrections} corrections, k {ro.k_iters} omega "
          f"{ro.w_iters} iterations", flush=True)
    for a, b in zip(sg.fields(), so.fields()):
        assert rel_l2(a, b) < 1e-9
    for a, b in zip(sg.turbulence(), so.turbulence()):
        assert rel_l2(a, b) < 1e-9
