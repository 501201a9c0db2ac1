"""Mixed precision on the device: FP32 AMG levels and smoothers inside FP64
CG, and FP32 lockstep BiCGSTAB inside FP64 iterative refinement.  The
refinement and the outer Krylov work on the FP64 true residual, so solutions
meet the same FP64 target as the FP64 path; the field errors against FP64 are
printed for the record."""
from __future__ import annotations

import numpy as np
import pytest

import cases
import oracle_lib as orc
from babelsim_amd import solver_config
from test_gpu_parity import laplace_system, rel_l2

pytestmark = pytest.mark.gpu


def test_mixed_vector_and_scalar_bicgstab():
    from babelsim_amd.device import Context, System
    arrays, om, diag, up, lo, b = laplace_system(anchor=True)
    ctx = Context(arrays, renumber=True)
    cfg = solver_config(solver=1, pc=1, rtol=1e-10, atol=1e-30, precision=1, inner_rtol=1e-3)
    B = np.stack([b, 2 * b + 0.3, -b], 1)
    sys = System(ctx)
    x, res = sys.solve(diag, up, lo, B, np.zeros_like(B), cfg)
    exact = solver_config(solver=1, pc=1, rtol=1e-12, atol=1e-30)
    csr = orc.csr_from_ldu(arrays, diag, up, lo)
    for comp in range(3):
        assert res[comp].status == 0 and res[comp].relative <= 1e-10
        xo, _, _ = orc.csr_solve(*csr, B[:, comp], np.zeros(len(b)), exact)
        assert rel_l2(x[:, comp], xo) < 1e-8
    x1, r1 = System(ctx).solve(diag, up, lo, B[:, 1], np.zeros(len(b)), cfg)
    assert r1[0].status == 0 and rel_l2(x1, x[:, 1]) < 1e-8


def test_mixed_cg_with_fp32_amg():
    from babelsim_amd.device import csr_solve
    arrays, om, diag, up, lo, b = laplace_system(anchor=True)
    csr = orc.csr_from_ldu(arrays, diag, up, lo)
    cfg = solver_config(solver=0, pc=2, rtol=1e-10, atol=1e-30, precision=1)
    x, r, h = csr_solve(*csr, b, np.zeros(len(b)), cfg)
    xo, ro, _ = orc.csr_solve(*csr, b, np.zeros(len(b)), solver_config(solver=0, pc=2, rtol=1e-12, atol=1e-30))
    assert r.status == 0 and r.relative <= 1e-10
    assert rel_l2(x, xo) < 1e-8


def test_mixed_simple_cavity_matches_fp64():
    """Config 1 on 48 x 48 run to convergence (SIMPLE tolerances 1e-6) in FP64
    and in mixed precision: the converged fields agree to 1e-6 relative L2."""
    from babelsim_amd.device import Context, Simple
    n = 48
    mesh = cases.cavity2d_mesh(n)
    ub, pb = cases.cavity_bcs(two_d=True)
    out = {}
    for prec in (0, 1):
        mom = solver_config(solver=1, pc=1, rtol=1e-8, atol=1e-14, precision=prec)
        pre = solver_config(solver=0, pc=2, rtol=1e-8, atol=1e-14, precision=prec)
        ctx = Context(mesh, renumber=True)
        sim = Simple(ctx, cases.c1_config(mom=mom, pre=pre), ub, pb)
        sim.start()
        its = 0
        while its < 20000:
            reps = sim.iterate(200)
            its += 200
            if any(r.converged for r in reps):
                break
        assert any(r.converged for r in reps)
        out[prec] = sim.fields()
        print(f"[mixed] precision={prec} converged within {its} iterations", flush=True)
    errs = {name: rel_l2(out[1][k], out[0][k]) for k, name in enumerate(("U", "p", "phi"))}
    print(f"[mixed] relative L2 vs FP64: {errs}", flush=True)
    assert all(e < 1e-6 for e in errs.values())
