"""The device equation seam (bsim_field_*, bsim_equ_*, bsim_math_*): the
reference's own expectations of tests/equ_cases.py (procedural_equation_test,
public_equation_test, physics_contract_test), which the oracle passes in
test_oracle_equ.py, and parity with the oracle's equation layer on a
jittered, split mesh with every term (convection with an inletOutlet
context, corrected diffusion with an interpolated coefficient, BDF2 time
term, sources, reaction, relaxation, scale/copy/add) to 1e-11 of each
array's scale, and the solved unknowns to 1e-9."""
import types

import numpy as np
import pytest

import equ as oracle_equ
import equ_cases as ec
import meshes
import oracle_lib as orc
from babelsim_amd import equ as dequ
from babelsim_amd import solver_config
from babelsim_amd.device import Context
from op_fields import fields, scalar_bcs, vector_bcs

pytestmark = pytest.mark.gpu

B = types.SimpleNamespace(**{k: getattr(dequ, k) for k in dir(dequ) if not k.startswith("_")})
B.mesh = lambda arrays: Context(arrays, renumber=True)
O = types.SimpleNamespace(**{k: getattr(oracle_equ, k) for k in dir(oracle_equ) if not k.startswith("_")})
O.mesh = orc.oracle_mesh


def test_procedural_equation_known_answers():
    ec.procedural(B)


def test_public_equation_known_answers():
    ec.public(B)


def test_frozen_diffusion_flux_belongs_to_the_matrix():
    ec.frozen_diffusion(B)


def test_courant_contract():
    ec.courant_contract(B)


def test_reductions():
    ec.reductions(B)


def close(a, b, rtol=1e-11):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(1.0, float(np.max(np.abs(b))))
    err = float(np.max(np.abs(a - b)))
    assert err <= rtol * scale, f"max abs err {err:.3e} (scale {scale:.3e})"


def build(S, mesh, ncomp, arrays, om):
    s, U, phi = fields(om)
    g = om.geometry()
    C = g["C"]
    x0 = s if ncomp == 1 else U
    bc = scalar_bcs(om.npatch) if ncomp == 1 else vector_bcs(om.npatch)
    flux = S.Field(mesh, 1, 1, g["Sf"] @ np.array([1.0, 0.4, -0.3]) + 0.05 * phi)
    x = S.Field(mesh, 0, ncomp, x0, bc)
    prev = S.Field(mesh, 0, ncomp, 0.9 * x0)
    older = S.Field(mesh, 0, ncomp, 0.8 * x0)
    k = S.Field(mesh, 0, 1, 1.0 + 0.2 * C[:, 0] ** 2, (np.full(om.npatch, 2, np.int32), np.zeros(om.npatch)))
    f = S.Field(mesh, 0, ncomp, np.cos(x0))
    c = S.Field(mesh, 0, 1, 0.5 + C[:, 1] ** 2)
    opts = S.methods(interp=1, grad=1, conv=1, diff=1, coef_interp=1, coef_grad=0)
    a = S.Equation(x, solver_config(solver=1, pc=1, rtol=1e-11, atol=1e-30, max_iter=2000), opts)
    a.div(flux, 1.3)
    a.laplacian(k, -0.7)
    a.laplacian(0.25, -1.0)
    a.ddt(2.0, prev, 0.1, S.BDF2, older, 0.2)
    a.source(f, 0.5)
    a.reaction(c, 1.5)
    a.relax(prev, 0.7)
    a.scale(0.9)
    b = a.copy()
    b.add(a, 0.25)
    return b, x, flux, (flux, x, prev, older, k, f, c, a)


@pytest.mark.parametrize("ncomp", [1, 3])
def test_equation_parity_with_the_oracle(ncomp):
    arrays = meshes.jittered()
    om = orc.oracle_mesh(arrays)
    db, dx, dflux, dkeep = build(B, B.mesh(arrays), ncomp, arrays, om)
    ob, ox, oflux, okeep = build(O, O.mesh(arrays), ncomp, arrays, om)
    close(db.diagonal(), ob.diagonal())
    close(db.rhs(), ob.rhs())
    close(db.apply(dx), ob.apply(ox))
    close(db.residual(dx), ob.residual(ox))
    assert abs(db.relative_residual(dx) - ob.relative_residual(ox)) <= 1e-11
    if ncomp == 1:
        close(db.face_flux(dx), ob.face_flux(ox))
    rd, ro = db.solve(), ob.solve()
    assert all(r.status == 0 for r in rd) and all(r.status == 0 for r in ro)
    assert [r.iterations for r in rd] == [r.iterations for r in ro]
    np.testing.assert_allclose(dx.values(), ox.values(), rtol=0, atol=1e-9 * np.max(np.abs(ox.values())))


def test_reference_if_unanchored_parity():
    arrays = meshes.jittered()
    om = orc.oracle_mesh(arrays)
    for S, mesh in ((B, B.mesh(arrays)), (O, O.mesh(arrays))):
        p = S.Field(mesh, 0, 1)
        e = S.Equation(p, solver_config(solver=0, pc=2, rtol=1e-12, atol=1e-30))
        e.laplacian(1.0, -1.0)
        assert e.reference_if_unanchored(3.0)
        S.result = (e.diagonal(), e.rhs())
    close(B.result[0], O.result[0])
    close(B.result[1], O.result[1])
