"""The AMG V-cycle replayed from a captured HIP graph against the same cycle
launched kernel by kernel (the probe switches graph replay off): every
replay, with the solver's own eager kernels and host waits between them,
must give the same bits."""
from __future__ import annotations

import numpy as np
import pytest

import cases
import piso_cases as pc
from babelsim_amd import solver_config
from babelsim_amd.device import Context, Piso, Simple

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_simple_graph_replay_matches_stream_launches(precision):
    mesh = cases.cavity3d_mesh(9, split=0.1, seed=5)
    cfg = cases.c2_config(mom=solver_config(solver=1, pc=1, rtol=1e-8, atol=1e-14, precision=precision),
                          pre=solver_config(solver=0, pc=2, rtol=1e-8, atol=1e-14, precision=precision))
    out = []
    for probe in (False, True):
        ctx = Context(mesh)
        s = Simple(ctx, cfg, *cases.cavity_bcs(False))
        s.start()
        if probe:
            s.probe_start(capacity=100000)
        reps = s.iterate(6)
        if probe:
            launches, _, _ = s.probe_stop()
            assert launches > 0
        out.append((s.fields(), s.history(3), [r.p_iters for r in reps]))
    (fa, ha, ia), (fb, hb, ib) = out
    assert ia == ib
    np.testing.assert_array_equal(ha, hb)
    for a, b in zip(fa, fb):
        np.testing.assert_array_equal(a, b)


def test_piso_graph_replay_matches_stream_launches():
    mesh = cases.cavity2d_mesh(32)
    out = []
    for probe in (False, True):
        ctx = Context(mesh)
        s = Piso(ctx, pc.cavity_piso_config(), *cases.cavity_bcs(True))
        s.start()
        if probe:
            s.probe_start(capacity=100000)
        for _ in range(4):
            s.step(0.01)
        if probe:
            s.probe_stop()
        out.append((s.fields(), s.history(3)))
    (fa, ha), (fb, hb) = out
    np.testing.assert_array_equal(ha, hb)
    for a, b in zip(fa, fb):
        np.testing.assert_array_equal(a, b)


def test_stream_launch_switch_gives_same_bits():
    """bsim_set_graph_replay(0) launches the cycles from the stream."""
    from babelsim_amd import set_graph_replay
    mesh = cases.cavity2d_mesh(24)
    out = []
    try:
        for replay in (True, False):
            set_graph_replay(replay)
            ctx = Context(mesh)
            s = Simple(ctx, cases.c1_config(), *cases.cavity_bcs(True))
            s.start()
            s.iterate(5)
            out.append((s.fields(), s.history(3)))
    finally:
        set_graph_replay(True)
    (fa, ha), (fb, hb) = out
    np.testing.assert_array_equal(ha, hb)
    for a, b in zip(fa, fb):
        np.testing.assert_array_equal(a, b)


def test_spmv_timing_modes_leave_the_solve_unchanged():
    """bsim_simple_time_spmv from HBM (Infinity Cache evicted before every
    launch) and back to back: both return a positive mean time and the matrix
    size, and the iterations after them give the same bits as a run without."""
    mesh = cases.cavity3d_mesh(9, split=0.1, seed=5)
    cfg = cases.c2_config()
    out = []
    for timed in (False, True):
        ctx = Context(mesh)
        s = Simple(ctx, cfg, *cases.cavity_bcs(False))
        s.start()
        s.iterate(2)
        if timed:
            for cold in (True, False):
                sec, n, nnz = s.time_spmv(1, reps=3, cold=cold)
                assert n == ctx.nc and nnz > 0 and 0.0 < sec < 1.0
        s.iterate(2)
        out.append((s.fields(), s.history(3)))
    (fa, ha), (fb, hb) = out
    np.testing.assert_array_equal(ha, hb)
    for a, b in zip(fa, fb):
        np.testing.assert_array_equal(a, b)
