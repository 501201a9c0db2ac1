"""Parity of the HIP library against the oracle on the same inputs.

Tolerances: geometry and explicit operators agree to 1e-11 of the array scale
(FP64, different summation order); linear-solver residual histories to 1e-10
relative; SIMPLE fields after a fixed number of outer iterations to a relative
L2 error of 1e-9 (the required 1e-6 with margin)."""
import numpy as np
import pytest

import cases
import meshes
from babelsim_amd import solver_config
from babelsim_amd.device import Context, Simple, System, amg_hierarchy, csr_solve
import oracle_lib as orc
from op_fields import fields, scalar_bcs, vector_bcs

pytestmark = pytest.mark.gpu

RTOL = 1e-11


def close(a, b, rtol=RTOL):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    err = float(np.max(np.abs(a - b))) if a.size else 0.0
    assert err <= rtol * scale, f"max abs err {err:.3e} (scale {scale:.3e})"


def rel_l2(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module", params=list(meshes.ALL))
def pair(request):
    arrays = meshes.ALL[request.param]()
    return request.param, orc.oracle_mesh(arrays), Context(arrays)


def test_geometry(pair):
    name, om, gm = pair
    go, gg = om.geometry(), gm.geometry()
    for key in go:
        close(gg[key], go[key])
    assert gm.orthogonal == (name in ("box", "cavity2d"))


@pytest.mark.parametrize("method", [0, 1])
def test_gradients(pair, method):
    name, om, gm = pair
    s, U, phi = fields(om)
    t, v = scalar_bcs(om.npatch)
    close(gm.gradient(s, t, v, method, phi), om.gradient(s, t, v, method, phi))
    t, v = vector_bcs(om.npatch)
    close(gm.gradient(U, t, v, method, phi), om.gradient(U, t, v, method, phi))


@pytest.mark.parametrize("method,gmethod", [(0, 0), (1, 0), (1, 1)])
def test_interpolation_and_flux(pair, method, gmethod):
    name, om, gm = pair
    s, U, phi = fields(om)
    t, v = scalar_bcs(om.npatch)
    close(gm.interpolate(s, t, v, method, gmethod, phi), om.interpolate(s, t, v, method, gmethod, phi))
    t, v = vector_bcs(om.npatch)
    close(gm.interpolate(U, t, v, method, gmethod, phi), om.interpolate(U, t, v, method, gmethod, phi))
    close(gm.flux(U, t, v, method, gmethod, phi), om.flux(U, t, v, method, gmethod, phi))
    close(gm.flux(U, t, v, method, gmethod, None), om.flux(U, t, v, method, gmethod, None))


@pytest.mark.parametrize("dmethod", [0, 1, 2])
@pytest.mark.parametrize("gmethod", [0, 1])
def test_normal_gradient_and_laplacian(pair, dmethod, gmethod):
    name, om, gm = pair
    s, U, phi = fields(om)
    t, v = scalar_bcs(om.npatch)
    t = np.where(t == 3, 2, t).astype(np.int32)
    close(gm.normal_gradient(s, t, v, gmethod, dmethod), om.normal_gradient(s, t, v, gmethod, dmethod))
    kf = 0.5 + np.abs(phi)
    close(gm.explicit_laplacian(s, t, v, kf, 0, gmethod, dmethod), om.explicit_laplacian(s, t, v, kf, 0, gmethod, dmethod))


def test_divergence(pair):
    name, om, gm = pair
    s, U, phi = fields(om)
    close(gm.divergence_face(phi), om.divergence_face(phi))


@pytest.mark.parametrize("conv,interp", [(0, 0), (1, 0), (2, 0), (2, 1)])
@pytest.mark.parametrize("gmethod", [0, 1])
def test_convection(pair, conv, interp, gmethod):
    name, om, gm = pair
    s, U, phi = fields(om)
    for x, (t, v) in ((s, scalar_bcs(om.npatch)), (U, vector_bcs(om.npatch))):
        close(gm.explicit_convection(x, t, v, phi, conv, interp, gmethod),
              om.explicit_convection(x, t, v, phi, conv, interp, gmethod))
        a = gm.assemble("convection", x, t, v, phi=phi, conv=conv, interp=interp, gmethod=gmethod, scale=1.3)
        b = om.assemble("convection", x, t, v, phi=phi, conv=conv, interp=interp, gmethod=gmethod, scale=1.3)
        for key in a:
            close(a[key], b[key])


@pytest.mark.parametrize("dmethod", [0, 1, 2])
@pytest.mark.parametrize("gmethod", [0, 1])
def test_diffusion_assembly(pair, dmethod, gmethod):
    name, om, gm = pair
    s, U, phi = fields(om)
    kf = 0.5 + np.abs(phi)
    for x, (t, v) in ((s, scalar_bcs(om.npatch)), (U, vector_bcs(om.npatch))):
        a = gm.assemble("diffusion", x, t, v, kf=kf, gmethod=gmethod, dmethod=dmethod, ctx=phi)
        b = om.assemble("diffusion", x, t, v, kf=kf, gmethod=gmethod, dmethod=dmethod, ctx=phi)
        for key in a:
            close(a[key], b[key])


# ------------------------------------------------------------ linear solves
def laplace_system(n=(12, 11, 10), jitter=0.2, split=0.1, anchor=True, seed=5):
    """A pressure-like SPD system on a jittered polyhedral mesh, assembled by
    the oracle: diffusion with positive face coefficients, one anchored cell."""
    from babelsim_amd import MeshData
    arrays, _, _ = MeshData.hex(n, jitter=jitter, split=split, seed=seed).prepare()
    arrays = arrays.arrays()
    om = orc.oracle_mesh(arrays)
    rng = np.random.default_rng(seed)
    kf = 0.5 + rng.random(om.nf)
    x = np.zeros(om.nc)
    t = np.full(6, 2, np.int32)
    q = om.assemble("diffusion", x, t, np.zeros(6), kf=kf)
    diag = q["diag"].copy()
    if anchor:
        diag[0] *= 2.0
    return arrays, om, diag, q["upper"], q["lower"], rng.standard_normal(om.nc)


def test_amg_hierarchy_is_bit_exact():
    arrays, om, diag, up, lo, b = laplace_system()
    P, Cc, D, Vv = orc.csr_from_ldu(arrays, diag, up, lo)
    cfg = solver_config(solver=0, pc=2)
    s1, a1, d1 = orc.amg_hierarchy(P, Cc, D, Vv, cfg)
    s2, a2, d2 = amg_hierarchy(P, Cc, D, Vv, cfg)
    assert list(s1) == list(s2)
    assert np.array_equal(a1, a2)
    assert np.array_equal(d1, d2)
    assert len(s1) >= 3


@pytest.mark.parametrize("solver,pc", [(0, 0), (0, 1), (0, 2), (0, 3), (1, 0), (1, 1), (1, 2), (1, 3)])
def test_csr_solver_histories(solver, pc):
    arrays, om, diag, up, lo, b = laplace_system()
    P, Cc, D, Vv = orc.csr_from_ldu(arrays, diag, up, lo)
    cfg = solver_config(solver=solver, pc=pc, rtol=1e-10, atol=1e-30, max_iter=2000)
    x1, r1, h1 = orc.csr_solve(P, Cc, D, Vv, b, np.zeros(len(b)), cfg)
    x2, r2, h2 = csr_solve(P, Cc, D, Vv, b, np.zeros(len(b)), cfg)
    assert r1.status == r2.status == 0
    assert r1.iterations == r2.iterations
    assert len(h1) == len(h2)
    np.testing.assert_allclose(h2, h1, rtol=1e-10, atol=0)
    assert rel_l2(x2, x1) < 1e-9


def test_one_level_amg_on_the_mesh_layout():
    """A hierarchy with a single level (coarse size above the row count) puts
    the dense factorisation on the fine matrix in its sliced mesh layout."""
    arrays, om, diag, up, lo, b = laplace_system(n=(6, 5, 4), anchor=True)
    assert len(diag) < 200
    ctx = Context(arrays, renumber=True)
    sys = System(ctx)
    cfg = solver_config(solver=0, pc=2, rtol=1e-10, atol=1e-30, amg_coarse_size=len(diag) + 1)
    x, res = sys.solve(diag, up, lo, b, np.zeros_like(b), cfg)
    assert res[0].status == 0 and res[0].iterations <= 2
    xo, ro, _ = orc.csr_solve(*orc.csr_from_ldu(arrays, diag, up, lo), b, np.zeros(len(b)), cfg)
    assert rel_l2(x, xo) < 1e-9


def test_block_jacobi_factors_on_the_mesh_layout():
    """The device's block Jacobi on a mesh system (sliced rows, caller's cell
    order, so the 64-row blocks are the oracle's): each preconditioned solve
    equals the oracle's CSR solve with the same preconditioner to 1e-12, a
    mixed-precision request is refused, and with renumbered cells (other
    blocks) both still converge to the same solution."""
    arrays, om, diag, up, lo, b = laplace_system(anchor=True)
    csr = orc.csr_from_ldu(arrays, diag, up, lo)
    for solver in (0, 1):
        cfg = solver_config(solver=solver, pc=3, rtol=1e-10, atol=1e-30)
        xo, ro, ho = orc.csr_solve(*csr, b, np.zeros(len(b)), cfg)
        sys = System(Context(arrays, renumber=False))
        x, res = sys.solve(diag, up, lo, b, np.zeros_like(b), cfg)
        assert res[0].status == 0 and res[0].iterations == ro.iterations
        assert rel_l2(x, xo) < 1e-12
        sysr = System(Context(arrays, renumber=True))
        xr, rr = sysr.solve(diag, up, lo, b, np.zeros_like(b), cfg)
        assert rr[0].status == 0 and rel_l2(xr, xo) < 1e-8
    with pytest.raises(ValueError):
        System(Context(arrays, renumber=False)).solve(diag, up, lo, b, np.zeros_like(b),
                                                      solver_config(solver=1, pc=3, precision=1))


def test_manufactured_solution():
    # Manufactured SPD system with x = 1 + 0.01 * gid (petsc_backend_test.cpp:23-53), to 2e-11.
    arrays, om, diag, up, lo, _ = laplace_system()
    P, Cc, D, Vv = orc.csr_from_ldu(arrays, diag, up, lo)
    xs = 1.0 + 0.01 * np.arange(len(D))
    from babelsim_amd.device import csr_spmv
    b = csr_spmv(P, Cc, D, Vv, xs)
    cfg = solver_config(solver=0, pc=2, rtol=1e-14, atol=1e-30, max_iter=2000)
    x, r, _ = csr_solve(P, Cc, D, Vv, b, np.zeros(len(b)), cfg)
    assert np.max(np.abs(x - xs)) <= 2e-11 * np.max(np.abs(xs))


def test_system_contract_and_counters():
    # LDU seam: vector system with three right-hand sides, exact-compare reuse.
    arrays, om, diag, up, lo, b = laplace_system(anchor=True)
    ctx = Context(arrays, renumber=False)
    sys = System(ctx)
    cfg = solver_config(solver=1, pc=1, rtol=1e-10, atol=1e-30)
    B = np.stack([b, 2 * b, -b], 1)
    x, res = sys.solve(diag, up, lo, B, np.zeros_like(B), cfg)
    assert all(r.status == 0 for r in res)
    for comp in range(3):
        xo, ro, _ = orc.csr_solve(*orc.csr_from_ldu(arrays, diag, up, lo), B[:, comp], np.zeros(len(b)), cfg)
        assert rel_l2(x[:, comp], xo) < 1e-9
    x, res = sys.solve(diag, up, lo, B, np.zeros_like(B), cfg)  # unchanged matrix: rhs-only
    c = sys.counters()
    assert c["matrix_updates"] == 1 and c["rhs_only"] == 1 and c["solves"] == 6


# ------------------------------------------------------------------ SIMPLE
def run_pair(mesh_data, cfg, two_d, iters):
    arrays, _, _ = mesh_data.prepare()
    arrays = arrays.arrays()
    ub, pb = cases.cavity_bcs(two_d)
    om = orc.oracle_mesh(arrays)
    so = orc.OracleSimple(om, cfg, ub, pb)
    ctx = Context(arrays, renumber=False)
    sg = Simple(ctx, cfg, ub, pb)
    so.start()
    sg.start()
    ro = so.iterate(iters)
    rg = sg.iterate(iters)
    return so, sg, ro, rg


@pytest.mark.parametrize("kind", ["c1", "c2", "c1_retry", "c1_warm", "c1_bjacobi", "c2_bjacobi", "c1_icc", "c2_asm"])
def test_simple_parity(kind):
    """Device SIMPLE against the oracle.  c1_retry caps the momentum solves at 3
    BiCGSTAB iterations so that lagged-preconditioner failures trigger the
    rebuild-and-retry path of the three-component solve; c1_warm warm-starts
    every solve; the bjacobi kinds solve momentum with BiCGSTAB and block
    Jacobi, the reference cavity case's pair (cases/cavity/numerics/solution.bs:7-8);
    c1_icc solves pressure with CG and icc, as the reference's SIMPLE tests do
    (tests/cavity_regression_test.cpp:52-53); c2_asm solves momentum with
    BiCGSTAB and asm."""
    bj = solver_config(solver=1, pc=3, rtol=1e-8, atol=1e-14)
    if kind == "c1_icc":
        pre = solver_config(solver=0, pc=4, rtol=1e-8, atol=1e-14)
        mesh, cfg, two_d, iters = cases.cavity2d_mesh(24), cases.c1_config(pre=pre), True, 12
    elif kind == "c2_asm":
        mom = solver_config(solver=1, pc=5, rtol=1e-8, atol=1e-14)
        mesh, cfg, two_d, iters = cases.cavity3d_mesh(7, split=0.15, seed=3), cases.c2_config(mom=mom), False, 8
    elif kind == "c1_bjacobi":
        mesh, cfg, two_d, iters = cases.cavity2d_mesh(24), cases.c1_config(mom=bj), True, 12
    elif kind == "c2_bjacobi":
        mesh, cfg, two_d, iters = cases.cavity3d_mesh(7, split=0.15, seed=3), cases.c2_config(mom=bj), False, 8
    elif kind == "c1":
        mesh, cfg, two_d, iters = cases.cavity2d_mesh(24), cases.c1_config(), True, 12
    elif kind == "c1_retry":
        mom = solver_config(solver=1, pc=1, rtol=1e-8, atol=1e-14, max_iter=3)
        mesh, cfg, two_d, iters = cases.cavity2d_mesh(24), cases.c1_config(mom=mom), True, 12
    elif kind == "c1_warm":
        mom = solver_config(solver=1, pc=1, rtol=1e-8, atol=1e-14, warm_start=1)
        pre = solver_config(solver=0, pc=2, rtol=1e-8, atol=1e-14, warm_start=1)
        mesh, cfg, two_d, iters = cases.cavity2d_mesh(24), cases.c1_config(mom=mom, pre=pre), True, 12
    else:
        mesh, cfg, two_d, iters = cases.cavity3d_mesh(7, split=0.15, seed=3), cases.c2_config(), False, 8
    so, sg, ro, rg = run_pair(mesh, cfg, two_d, iters)
    for a, b in zip(rg, ro):
        assert list(a.mom_iters) == list(b.mom_iters) and a.p_iters == b.p_iters
        for key in ("rU", "dU", "dP"):
            assert abs(getattr(a, key) - getattr(b, key)) <= 1e-8 * abs(getattr(b, key)) + 1e-14
        assert a.referenced == b.referenced
    Ug, pg, phig = sg.fields()
    Uo, po, phio = so.fields()
    assert rel_l2(Ug, Uo) < 1e-9
    assert rel_l2(pg, po) < 1e-9
    assert rel_l2(phig, phio) < 1e-9
    for which in range(4):
        hg, ho = sg.history(which), so.history(which)
        assert len(hg) == len(ho)
        np.testing.assert_allclose(hg, ho, rtol=1e-10, atol=1e-300)
