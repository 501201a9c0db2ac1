"""The serving runtime (bsim_server_*): every served result equals the solo
run of the same request on its own context, bit for bit, whether it ran alone
or in a mix of concurrent requests, on a resident (reused) context or a fresh
one, under a budget that lets only one request run at a time, and in mixed
precision.  One served request is also checked against the oracle's SIMPLE.
Admission: arrival times hold, requests beyond the budget are rejected,
contexts are evicted when memory is needed."""
from __future__ import annotations

import numpy as np
import pytest

import cases
import oracle_lib as orc
import piso_cases as pc
import serve_cases as sc
from babelsim_amd import SERVE_CONVERGED, SERVE_REJECTED, SERVE_NOT_CONVERGED
from babelsim_amd.device import Context, Piso, Simple
from babelsim_amd.serve import REQ_STEADY, Server, estimate

pytestmark = pytest.mark.gpu

CAP = 600  # outer-iteration cap of the steady requests


def solo_steady(mesh, cfg, bcs, cap=CAP):
    ub, pb = bcs
    s = Simple(Context(mesh, renumber=True), cfg, ub, pb)
    s.start()
    its, status = 0, SERVE_NOT_CONVERGED
    for k in range(1, cap + 1):
        r = s.iterate(1)[-1]
        its = k
        if r.converged:
            status = SERVE_CONVERGED
            break
    return status, its, s.fields()


def solo_transient(mesh, cfg, bcs, steps, dt):
    ub, pb = bcs
    s = Piso(Context(mesh, renumber=True), cfg, ub, pb)
    s.start()
    for _ in range(steps):
        s.step(dt)
    return s.fields()


def workload():
    """Small cases of each kind the stream holds, plus a transient one."""
    return [
        dict(name="cavity2d", mesh=cases.cavity2d_mesh(16), bcs=cases.cavity_bcs(two_d=True),
             cfg=sc.serve_config(100.0)),
        dict(name="cavity3d", mesh=sc.cavity_mesh(8), bcs=sc.cavity_bcs(), cfg=sc.serve_config(400.0)),
        dict(name="channel3d", mesh=sc.channel_mesh(4), bcs=sc.channel_bcs(), cfg=sc.serve_config(200.0)),
        dict(name="cavity3d-re900", mesh=sc.cavity_mesh(8), bcs=sc.cavity_bcs(), cfg=sc.serve_config(900.0)),
        dict(name="piso-cavity", mesh=cases.cavity2d_mesh(12), bcs=cases.cavity_bcs(two_d=True),
             cfg=pc.cavity_piso_config(), steps=4, dt=0.01),
    ]


@pytest.fixture(scope="module")
def solo():
    out = []
    for w in workload():
        if "steps" in w:
            out.append((SERVE_CONVERGED, w["steps"], solo_transient(w["mesh"], w["cfg"], w["bcs"], w["steps"], w["dt"])))
        else:
            out.append(solo_steady(w["mesh"], w["cfg"], w["bcs"]))
    return out


def serve_all(server, ws, arrival=0.0):
    ids = []
    case_of = {}
    for w in ws:
        key = id(w["mesh"])
        if key not in case_of:
            case_of[key] = server.add_case(w["mesh"])
        cid = case_of[key]
        if "steps" in w:
            rid = server.submit_transient(cid, w["cfg"], *w["bcs"], steps=w["steps"], dt=w["dt"], arrival=arrival)
        else:
            rid = server.submit_steady(cid, w["cfg"], *w["bcs"], max_iterations=CAP, arrival=arrival)
        ids.append((rid, cid))
    return ids


def check_equal(server, ids, solo):
    for (rid, cid), (status, its, (U, p, phi)) in zip(ids, solo):
        r = server.wait(rid)
        assert r.status == status, r.message
        assert r.iterations == its
        # The admission reservation covers what the solver really allocated.
        assert 0 < r.measured_bytes <= r.reserved_bytes
        Us, ps, phis = server.fields(rid, cid)
        np.testing.assert_array_equal(Us, U)
        np.testing.assert_array_equal(ps, p)
        np.testing.assert_array_equal(phis, phi)


def test_solo_steady_requests_converge(solo):
    for w, (status, its, _) in zip(workload(), solo):
        if "steps" not in w:
            assert status == SERVE_CONVERGED, (w["name"], its)


def test_mix_equals_solo(solo):
    ws = workload()
    with Server(slots=4, memory_budget=8e9) as srv:
        ids = serve_all(srv, ws)
        check_equal(srv, ids, solo)
        st = srv.stats()
        assert st["completed"] == len(ws)
        assert st["peak_running"] >= 2  # the requests overlapped
        assert st["running"] == 0


def test_resident_contexts_are_reused_with_the_same_results(solo):
    ws = workload()
    with Server(slots=2, memory_budget=8e9) as srv:
        first = serve_all(srv, ws)
        check_equal(srv, first, solo)
        # Second wave on the same registered cases: resident contexts.
        case_ids = [cid for _, cid in first]
        ids = []
        for w, cid in zip(ws, case_ids):
            if "steps" in w:
                rid = srv.submit_transient(cid, w["cfg"], *w["bcs"], steps=w["steps"], dt=w["dt"])
            else:
                rid = srv.submit_steady(cid, w["cfg"], *w["bcs"], max_iterations=CAP)
            ids.append((rid, cid))
        check_equal(srv, ids, solo)
        assert any(srv.wait(rid).reused_context for rid, _ in ids)
        assert srv.stats()["context_reuses"] >= 1


def test_requests_on_one_mesh_reuse_resident_memory(solo):
    """One slot, the same steady request three times: after the first, the
    solver, fields and work vectors are rebuilt from the blocks the previous
    request returned to the device cache (no runtime allocation), and the
    results stay equal to the solo run."""
    from babelsim_amd.serve import pool_stats, pool_trim
    w = next(w for w in workload() if "steps" not in w)
    k = [i for i, x in enumerate(workload()) if "steps" not in x][0]
    pool_trim(0)
    with Server(slots=1, memory_budget=8e9) as srv:
        cid = srv.add_case(w["mesh"])
        ids, misses = [], []
        for _ in range(3):
            rid = srv.submit_steady(cid, w["cfg"], *w["bcs"], max_iterations=CAP)
            srv.wait(rid)
            ids.append((rid, cid))
            misses.append(pool_stats(0)["misses"])
        check_equal(srv, ids, [solo[k]] * 3)
        assert misses[2] == misses[1], misses  # the third request allocated nothing new
        assert pool_stats(0)["hits"] > 0
    pool_trim(0)
    assert pool_stats(0)["cached_bytes"] == 0


def test_tight_budget_runs_one_at_a_time(solo):
    ws = workload()
    footprint = 0.0
    for w in ws:
        _, nf, _, _, nc = w["mesh"].sizes()
        ctx_b, sol_b = estimate(REQ_STEADY if "steps" not in w else 1, nc, nf)
        footprint = max(footprint, ctx_b + sol_b)
    with Server(slots=4, memory_budget=1.2 * footprint) as srv:
        ids = serve_all(srv, ws)
        check_equal(srv, ids, solo)
        st = srv.stats()
        assert st["peak_running"] == 1
        assert st["evictions"] >= 1
        assert st["peak_reserved"] <= 1.2 * footprint


def test_request_beyond_budget_is_rejected():
    with Server(slots=1, memory_budget=1 << 20) as srv:
        w = workload()[0]
        cid = srv.add_case(w["mesh"])
        r = srv.wait(srv.submit_steady(cid, w["cfg"], *w["bcs"]))
        assert r.status == SERVE_REJECTED
        assert "budget" in r.message
        assert srv.stats()["rejected"] == 1


def test_arrival_time_is_respected():
    with Server(slots=2, memory_budget=8e9) as srv:
        w = workload()[2]
        cid = srv.add_case(w["mesh"])
        t = srv.clock()
        r = srv.wait(srv.submit_steady(cid, w["cfg"], *w["bcs"], max_iterations=5, arrival=t + 0.4))
        assert r.released >= t + 0.4
        assert r.admitted >= r.released
        assert r.finished >= r.started >= r.admitted
        assert r.status == SERVE_NOT_CONVERGED and r.iterations == 5


def test_invalid_request_is_refused_at_submit():
    with Server(slots=1, memory_budget=8e9) as srv:
        w = workload()[0]
        cid = srv.add_case(w["mesh"])
        with pytest.raises(ValueError):
            srv.submit_steady(cid, sc.serve_config(100.0), *w["bcs"], max_iterations=0)
        with pytest.raises(ValueError):
            srv.submit_steady(cid + 5, w["cfg"], *w["bcs"])


def test_mixed_precision_served_equals_solo():
    w = dict(mesh=sc.cavity_mesh(8), bcs=sc.cavity_bcs(), cfg=sc.serve_config(400.0, precision=1))
    ref = solo_steady(w["mesh"], w["cfg"], w["bcs"])
    ref64 = solo_steady(w["mesh"], sc.serve_config(400.0), w["bcs"])
    with Server(slots=2, memory_budget=8e9) as srv:
        ids = serve_all(srv, [w, dict(w)])
        check_equal(srv, ids, [ref, ref])
    # FP32 AMG / inner Krylov against the FP64 request: converged fields agree.
    for a, b in zip(ref[2][:2], ref64[2][:2]):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-4


def test_served_request_matches_the_oracle():
    # The server renumbers the mesh and the oracle keeps the caller's order,
    # so their AMG aggregates differ; tight linear tolerances make the fields
    # independent of the preconditioner to well below the 1e-9 bar.
    mesh = sc.channel_mesh(4)
    ub, pb = sc.channel_bcs()
    cfg = sc.serve_config(200.0, rtol=1e-12)
    with Server(slots=1, memory_budget=8e9) as srv:
        cid = srv.add_case(mesh)
        rid = srv.submit_steady(cid, cfg, ub, pb, max_iterations=12)
        r = srv.wait(rid)
        U, p, phi = srv.fields(rid, cid)
    so = orc.OracleSimple(orc.oracle_mesh(mesh.arrays()), cfg, ub, pb)
    so.start()
    so.iterate(r.iterations)
    Uo, po, phio = so.fields()
    for a, b in ((U, Uo), (p, po), (phi, phio)):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-9


def test_failed_request_reports_an_error_and_the_server_goes_on(solo):
    """A request whose solver rejects its configuration inside the worker
    (negative viscosity: SimpleSolver's own check) ends with SERVE_ERROR and
    the message; its context is dropped and the next request is served as a
    solo run would be."""
    from babelsim_amd import SERVE_ERROR
    w = workload()[1]
    with Server(slots=1, memory_budget=8e9) as srv:
        cid = srv.add_case(w["mesh"])
        bad = sc.serve_config(400.0)
        bad.mu = -1.0
        r = srv.wait(srv.submit_steady(cid, bad, *w["bcs"], max_iterations=5))
        assert r.status == SERVE_ERROR
        assert "viscosity" in r.message
        st = srv.stats()
        assert st["resident_idle"] == 0 and st["reserved"] == 0
        ids = [(srv.submit_steady(cid, w["cfg"], *w["bcs"], max_iterations=CAP), cid)]
        check_equal(srv, ids, [solo[1]])


def test_numerical_failure_ends_a_served_request(solo):
    """runSimple returns numericalFailure at the first non-finite measure or
    unhealthy solve (simple/main.cpp:61-67, 88, 115-116); a served request
    ends there with SERVE_NUMERICAL_FAILURE after one iteration instead of
    running to its cap.  Green-Gauss gradients carry the non-finite value into
    the momentum residual (least-squares gradients would raise, as the
    reference's operators.cpp:358-360 does, which is an error, not a failure)."""
    from babelsim_amd import SERVE_NUMERICAL_FAILURE
    w = dict(workload()[0], cfg=cases.c1_config())
    mesh = w["mesh"]
    nc = mesh.sizes()[4]
    U0 = np.zeros((nc, 3))
    U0[nc // 2, 0] = np.nan
    with Server(slots=1, memory_budget=8e9) as srv:
        cid = srv.add_case(mesh)
        r = srv.wait(srv.submit_steady(cid, w["cfg"], *w["bcs"], max_iterations=50, U0=U0))
        assert r.status == SERVE_NUMERICAL_FAILURE and r.iterations == 1, r.message
        ok = workload()[0]
        ids = [(srv.submit_steady(cid, ok["cfg"], *ok["bcs"], max_iterations=CAP), cid)]
        check_equal(srv, ids, [solo[0]])
    ub, pb = w["bcs"]
    s = Simple(Context(mesh, renumber=True), w["cfg"], ub, pb)
    s.set_fields(U=U0)
    s.start()
    reps = s.iterate(3)
    assert [r.numerical_failure for r in reps] == [1, 1, 1]
    assert reps[1].p_iters == 0 and reps[2].mom_iters[0] == 0
