"""The reference's own expectations for the procedural equation layer, written
once for any backend with the Field / Equation / math interface of
babelsim_amd.equ (the HIP library, tests/test_gpu_equ.py) and oracle/equ.py
(the oracle, tests/test_oracle_equ.py):

- tests/procedural_equation_test.cpp:9-86 (time diagonal 12, source 24, the
  frozen solve 2, V/aP 1/6, relaxed 1/12, variable-step BDF2 2.25, diffusion
  1 and 3, the Laplacian sign, homogeneous corrections, the pressure
  reference 3, the vector reaction/source solve (1, 2, 3));
- tests/public_equation_test.cpp:13-111 (vector ddt + source 0.2 f with V/aP
  0.1, reference-constrained Poisson 3, the rejections, scalar relaxation
  0.1);
- tests/physics_contract_test.cpp:43-91 (frozen diffusion: V div(faceFlux)
  = A x - b for every diffusion method, constant and variable coefficient,
  after copy/scale/add and away from the solution; Courant 0.5 and 3.0).

`B` provides mesh(arrays), Field(mesh, location, ncomp, values, bc),
Equation(unknown, linear, options), methods(...), EULER, BDF2, and the
reductions norms, relative_change, relative_magnitude, flux_balance, courant,
face_add."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as orc
from babelsim_amd import MeshData, solver_config

CELL, FACE = 0, 1
FV, FG, ZG, IO, SYM = range(5)
MINUS_X, PLUS_X = 0, 1


def linear(**kw):
    base = dict(solver=0, pc=1, rtol=1e-12, atol=1e-14, max_iter=500)
    base.update(kw)
    return solver_config(**base)


def near(a, b, tol=1e-9):
    return np.allclose(a, b, rtol=0, atol=tol)


def hexbox(n, hi):
    return MeshData.hex(n, (0, 0, 0), hi).arrays()


def procedural(B):
    """procedural_equation_test.cpp:9-86 on makeHexBox({2,1,1},{0,0,0},{4,1,1})."""
    arrays = hexbox((2, 1, 1), (4, 1, 1))
    m = B.mesh(arrays)
    V = orc.oracle_mesh(arrays).geometry()["V"]
    x = B.Field(m, CELL, 1, np.ones(2))
    old = B.Field(m, CELL, 1, np.ones(2))
    capacity = B.Field(m, CELL, 1, np.full(2, 3.0))
    a = B.Equation(x, linear())
    a.ddt(capacity, old, 0.5)
    a.source(6.0)
    capacity.set(np.full(2, 99.0))  # assembled coefficients are frozen
    old.set(np.full(2, 99.0))
    assert near(a.diagonal()[0], 12), "integrated time diagonal"
    assert near(a.rhs()[0], 24), "source sign/volume or frozen history"
    r = a.solve()
    assert r[0].status == 0 and near(x.values()[0], 2), "frozen system changed during solve"
    assert near(a.residual(x)[0], 0), "b-Ax residual"
    saved = a.copy()
    a.reset()
    a.add(saved, 2)
    a.scale(0.5)
    assert near(V[0] / a.diagonal()[0], 1.0 / 6), "V/aP response"
    a.relax(x, 0.5)
    assert near(V[0] / a.diagonal()[0], 1.0 / 12), "relaxed response"
    assert a.solve()[0].status == 0 and near(x.values()[0], 2), "relax fixed point"
    # Variable-step BDF2 differentiates a quadratic exactly.
    old.set(np.ones(2))
    older = B.Field(m, CELL, 1, np.zeros(2))
    a.reset()
    a.ddt(1.0, old, 0.5, B.BDF2, older, 1.0)
    a.source(3.0)
    assert a.solve()[0].status == 0 and near(x.values()[0], 2.25), "variable-step BDF2"
    # Known linear diffusion solution: boundary RHS and Laplacian sign.
    types = np.full(6, ZG, np.int32)
    types[MINUS_X] = types[PLUS_X] = FV
    vals = np.zeros(6)
    vals[PLUS_X] = 4.0
    x.set(np.zeros(2))
    x.set_boundary(types, vals)
    a.reset()
    a.laplacian(1.0, -1)
    assert a.solve()[0].status == 0
    assert near(x.values(), [1, 3]), "diffusion boundary/sign"
    negative = a.copy()
    a.reset()
    a.laplacian(1.0, +1)
    assert near(a.diagonal() + negative.diagonal(), 0), "laplacian LHS sign"
    assert near(a.rhs() + negative.rhs(), 0), "laplacian boundary RHS sign"
    before, d0, b0 = x.values(), a.diagonal(), a.rhs()
    assert a.relative_residual(x) < 1e-12, "assembled residual"
    assert near(x.values(), before) and near(a.diagonal(), d0) and near(a.rhs(), b0), "diagnostic changed state"
    # Correction boundaries are homogeneous (field.h:382-409); an anchored
    # matrix gets no extra reference.
    corr = B.Field(m, CELL, 1, None, (types, np.zeros(6)))
    pressure = B.Equation(corr, linear())
    pressure.laplacian(1.0, -1)
    anchored = pressure.diagonal()
    assert not pressure.reference_if_unanchored(7.0)
    assert near(pressure.diagonal(), anchored), "anchored pressure received a redundant reference"
    corr.set(np.full(2, 2.0))
    assert pressure.solve()[0].status == 0
    assert near(B.norms(corr)[0], 0.0), "correction boundary is not homogeneous"
    free = B.Field(m, CELL, 1)
    fs = B.Equation(free, linear())
    fs.laplacian(1.0, -1)
    assert fs.reference_if_unanchored(3.0)
    assert fs.solve()[0].status == 0 and near(free.values(), [3, 3]), "pressure reference value"
    u = B.Field(m, CELL, 3)
    v = B.Equation(u, linear(solver=1))
    v.reaction(2.0)
    v.source([2.0, 4.0, 6.0])
    assert all(r.status == 0 for r in v.solve())
    assert near(u.values()[0], [1, 2, 3]), "vector source/diagonal"


def public(B):
    """public_equation_test.cpp:13-111 on makeHexBox({2,1,1},{0,0,0},{2,1,1}), dt 0.1, Euler."""
    arrays = hexbox((2, 1, 1), (2, 1, 1))
    m = B.mesh(arrays)
    V = orc.oracle_mesh(arrays).geometry()["V"]
    dt = 0.1
    U = B.Field(m, CELL, 3)
    U_old = B.Field(m, CELL, 3)
    force_v = np.array([[1.0, 2, 3], [3, 2, 1]])
    force = B.Field(m, CELL, 3, force_v)
    momentum = B.Equation(U, linear(solver=1))
    momentum.ddt(1.0, U_old, dt, B.EULER)
    momentum.source(force, 2.0)
    assert all(r.status == 0 for r in momentum.solve()), "vector Field source did not converge"
    assert near(U.values(), 0.2 * force_v, 1e-12), "vector source components are incorrect"
    assert near(V / momentum.diagonal(), 0.1, 1e-12), "diagonal response is incorrect"
    p = B.Field(m, CELL, 1)
    poisson = B.Equation(p, linear())
    poisson.laplacian(1.0, -1.0)
    poisson.reference(0, 3.0)
    assert poisson.solve()[0].status == 0
    assert near(p.values(), 3.0, 1e-10), "reference value was ignored"
    with pytest.raises(ValueError):
        momentum.reference(0, 0.0)  # a vector equation has no scalar reference
    with pytest.raises(ValueError):
        poisson.relax(p, 0.0)
    flux = B.Field(m, FACE, 1, np.ones(len(arrays["owner"])))
    other = B.Field(m, FACE, 1, np.ones(len(arrays["owner"])))
    transported = B.Field(m, CELL, 1)
    eq = B.Equation(transported, linear(solver=1))
    eq.div(flux)
    with pytest.raises(ValueError):
        eq.div(other)  # one equation, one boundary flux context
    T = B.Field(m, CELL, 1)
    T_old = B.Field(m, CELL, 1)
    energy = B.Equation(T, linear())
    energy.ddt(1.0, T_old, dt, B.EULER)
    energy.source(2.0)
    energy.relax(T, 0.5)
    assert energy.solve()[0].status == 0
    assert near(T.values(), 0.1, 1e-12), "scalar relaxation formula changed"


def sheared():
    """makeHexFromVertices({4,3,1}) with vertices (i + 0.7 j, j, k) (physics_contract_test.cpp:121-124)."""
    vx = np.array([[i + 0.7 * j, j, k] for k in range(2) for j in range(4) for i in range(5)], np.float64)
    return MeshData.hex((4, 3, 1), vertices=vx).arrays()


def outward_sums(arrays, phi):
    own, nei = arrays["owner"], arrays["neighbour"]
    inner = nei >= 0
    s = np.zeros(int(max(own.max(), nei.max()) + 1))
    np.add.at(s, own, phi)
    np.add.at(s, nei[inner], -phi[inner])
    return s


def frozen_diffusion(B):
    """frozenDiffusion (physics_contract_test.cpp:43-77): the recovered face
    flux belongs to the assembled matrix, V div(faceFlux) = A x - b, for every
    diffusion method with a constant and a variable coefficient, through
    copy/scale/add, after the coefficient is changed, at the solution and
    away from it."""
    arrays = sheared()
    m = B.mesh(arrays)
    C = orc.oracle_mesh(arrays).geometry()["C"]
    types = np.full(6, FV, np.int32)
    for method in (0, 1, 2):
        for variable in (False, True):
            p = B.Field(m, CELL, 1, 1 + C[:, 0] * C[:, 1], (types, np.zeros(6)))
            k = B.Field(m, CELL, 1, 1 + 0.1 * C[:, 0])
            a = B.Equation(p, linear(), B.methods(grad=0, diff=method))
            a.laplacian(k if variable else 1.0, -1)
            combined = a.copy()
            combined.scale(0.4)
            combined.add(a, 0.6)
            k.set(np.full(len(C), 20.0))
            assert combined.solve()[0].status == 0, "diffusion solve failed"

            def check():
                matrix_flux = combined.apply(p) - combined.rhs()
                recovered = outward_sums(arrays, combined.face_flux(p))
                assert np.linalg.norm(recovered - matrix_flux) < 1e-10, (method, variable)

            check()
            p.set(0.1 + 0.2 * C[:, 1] ** 2)
            check()


def courant_contract(B):
    """courantContract (physics_contract_test.cpp:79-91): sheared unit-volume
    cells with U = (2, 0, 0), only the x-normal faces carry flux."""
    arrays = sheared()
    m = B.mesh(arrays)
    g = orc.oracle_mesh(arrays).geometry()
    phi = g["Sf"] @ np.array([2.0, 0.0, 0.0])
    f = B.Field(m, FACE, 1, phi)
    co = B.courant(f, 0.25)
    assert near(co["maximum"], 0.5, 1e-12) and near(co["mean"], 0.5, 1e-12)
    f3 = B.Field(m, FACE, 1, -3.0 * phi)
    co = B.courant(f3, 0.5)
    assert near(co["maximum"], 3.0, 1e-12) and near(co["mean"], 3.0, 1e-12)
    with pytest.raises(ValueError):
        B.courant(f, 0.0)


def reductions(B):
    """Flux balance of a divergence-free and of a one-face flux, face
    increments by region, norms and relative measures against their
    definitions (fvm_execution.cpp:91-175, field_math.cpp:30-82)."""
    arrays = sheared()
    m = B.mesh(arrays)
    g = orc.oracle_mesh(arrays).geometry()
    nei = arrays["neighbour"]
    phi = g["Sf"] @ np.array([2.0, -1.0, 0.5])  # uniform velocity: every cell balances
    fb = B.flux_balance(B.Field(m, FACE, 1, phi))
    assert fb["l1"] < 1e-12 and fb["maximum"] < 1e-12 and fb["relative"] < 1e-12
    one = np.zeros(len(nei))
    f0 = int(np.flatnonzero(nei >= 0)[0])
    one[f0] = 1.0
    fb = B.flux_balance(B.Field(m, FACE, 1, one))
    assert near(fb["l1"], 2.0) and near(fb["l2"], np.sqrt(2.0)) and near(fb["maximum"], 1.0) and near(fb["relative"], 1.0)
    target = B.Field(m, FACE, 1, np.ones(len(nei)))
    inc = B.Field(m, FACE, 1, np.full(len(nei), 2.0))
    B.face_add(inc, target, -1.0, 1)  # interior faces only
    assert near(target.values(), np.where(nei >= 0, -1.0, 1.0))
    B.face_add(inc, target, 0.5, 0)
    assert near(target.values(), np.where(nei >= 0, 0.0, 2.0))
    rng = np.random.default_rng(5)
    a, b = rng.standard_normal((12, 3)), rng.standard_normal((12, 3))
    fa, fb3 = B.Field(m, CELL, 3, a), B.Field(m, CELL, 3, b)
    s = B.Field(m, CELL, 1, a[:, 0])
    n = B.norms(fa, fb3, s)
    assert near(n, [np.linalg.norm(a), np.linalg.norm(b), np.linalg.norm(a[:, 0])], 1e-12)
    assert near(B.relative_change(fa, fb3), np.linalg.norm(a - b) / np.linalg.norm(a), 1e-12)
    assert near(B.relative_magnitude(s, B.Field(m, CELL, 1, b[:, 1])), np.linalg.norm(a[:, 0]) / np.linalg.norm(b[:, 1]), 1e-12)
