"""Config 1 end to end on the device: the Re = 100 lid-driven cavity on the
128 x 128 lattice run to convergence (all SIMPLE tolerances 1e-6) reproduces
Ghia et al.'s centrelines within the reference's RMS limit of 0.01
(cases/cavity/validation/validate_cavity.py:133-153)."""
from __future__ import annotations

import numpy as np
import pytest

import cases
import ghia

pytestmark = pytest.mark.gpu


def test_c1_cavity_converges_to_ghia():
    from babelsim_amd.device import Context, Simple
    n = 128
    mesh = cases.cavity2d_mesh(n)
    ctx = Context(mesh, renumber=True)
    ub, pb = cases.cavity_bcs(two_d=True)
    sim = Simple(ctx, cases.c1_config(), ub, pb)
    sim.start()
    its, done = 0, False
    while its < 20000 and not done:
        reps = sim.iterate(500)
        its += 500
        done = any(r.converged for r in reps)
        print(f"[c1] {its} its rU={reps[-1].rU:.3e} mass={reps[-1].mass:.3e}", flush=True)
    assert done
    U, _, _ = sim.fields()
    u_l2, v_l2 = ghia.centreline_errors(ctx.geometry()["C"], U, n)
    print(f"[c1] converged after <= {its} iterations: u_L2={u_l2:.4f} v_L2={v_l2:.4f}", flush=True)
    assert u_l2 < 0.01 and v_l2 < 0.01
    assert np.abs(U[:, 2]).max() < 1e-10
