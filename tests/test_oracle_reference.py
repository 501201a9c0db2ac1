"""Pins the oracle against the reference's own compiled discretisation layer
(oracle/_ref: src/core/mesh.cpp + src/discretization/operators.cpp built from
the reference sources).  Skipped where that build is absent (the GPU box)."""
import numpy as np
import pytest

import meshes
from oracle_lib import oracle_mesh, reference, reference_mesh
from op_fields import fields, scalar_bcs, vector_bcs

pytestmark = pytest.mark.skipif(reference() is None, reason="reference discretisation layer not built here")

RTOL = 1e-11


def close(a, b, rtol=RTOL):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
    err = float(np.max(np.abs(a - b))) if a.size else 0.0
    assert err <= rtol * scale, f"max abs err {err:.3e} (scale {scale:.3e})"


@pytest.fixture(params=list(meshes.ALL))
def pair(request):
    arrays = meshes.ALL[request.param]()
    return request.param, oracle_mesh(arrays), reference_mesh(arrays)


def test_geometry(pair):
    name, om, rm = pair
    go, gr = om.geometry(), rm.geometry()
    for key in go:
        close(go[key], gr[key])


@pytest.mark.parametrize("method", [0, 1])
def test_gradients(pair, method):
    name, om, rm = pair
    s, U, phi = fields(om)
    t, v = scalar_bcs(om.npatch)
    close(om.gradient(s, t, v, method, phi), rm.gradient(s, t, v, method, phi))
    t, v = vector_bcs(om.npatch)
    close(om.gradient(U, t, v, method, phi), rm.gradient(U, t, v, method, phi))


@pytest.mark.parametrize("method,gmethod", [(0, 0), (1, 0), (1, 1)])
def test_interpolation_and_flux(pair, method, gmethod):
    name, om, rm = pair
    s, U, phi = fields(om)
    t, v = scalar_bcs(om.npatch)
    close(om.interpolate(s, t, v, method, gmethod, phi), rm.interpolate(s, t, v, method, gmethod, phi))
    t, v = vector_bcs(om.npatch)
    close(om.interpolate(U, t, v, method, gmethod, phi), rm.interpolate(U, t, v, method, gmethod, phi))
    close(om.flux(U, t, v, method, gmethod, phi), rm.flux(U, t, v, method, gmethod, phi))
    close(om.flux(U, t, v, method, gmethod, None), rm.flux(U, t, v, method, gmethod, None))


@pytest.mark.parametrize("dmethod", [0, 1, 2])
@pytest.mark.parametrize("gmethod", [0, 1])
def test_normal_gradient_and_laplacian(pair, dmethod, gmethod):
    name, om, rm = pair
    s, U, phi = fields(om)
    t, v = scalar_bcs(om.npatch)
    t = np.where(t == 3, 2, t).astype(np.int32)  # inletOutlet needs a context these operators do not bind
    close(om.normal_gradient(s, t, v, gmethod, dmethod), rm.normal_gradient(s, t, v, gmethod, dmethod))
    kf = 0.5 + np.abs(phi)
    close(om.explicit_laplacian(s, t, v, kf, 0, gmethod, dmethod),
          rm.explicit_laplacian(s, t, v, kf, 0, gmethod, dmethod))
    close(om.explicit_laplacian(s, t, v, None, 0.3, gmethod, dmethod),
          rm.explicit_laplacian(s, t, v, None, 0.3, gmethod, dmethod))


def test_divergence(pair):
    name, om, rm = pair
    s, U, phi = fields(om)
    close(om.divergence_face(phi), rm.divergence_face(phi))


@pytest.mark.parametrize("conv,interp", [(0, 0), (1, 0), (2, 0), (2, 1)])
@pytest.mark.parametrize("gmethod", [0, 1])
def test_convection(pair, conv, interp, gmethod):
    name, om, rm = pair
    s, U, phi = fields(om)
    t, v = scalar_bcs(om.npatch)
    close(om.explicit_convection(s, t, v, phi, conv, interp, gmethod),
          rm.explicit_convection(s, t, v, phi, conv, interp, gmethod))
    a = om.assemble("convection", s, t, v, phi=phi, conv=conv, interp=interp, gmethod=gmethod, scale=1.3)
    b = rm.assemble("convection", s, t, v, phi=phi, conv=conv, interp=interp, gmethod=gmethod, scale=1.3)
    for key in a:
        close(a[key], b[key])
    t, v = vector_bcs(om.npatch)
    close(om.explicit_convection(U, t, v, phi, conv, interp, gmethod),
          rm.explicit_convection(U, t, v, phi, conv, interp, gmethod))
    a = om.assemble("convection", U, t, v, phi=phi, conv=conv, interp=interp, gmethod=gmethod, scale=0.8)
    b = rm.assemble("convection", U, t, v, phi=phi, conv=conv, interp=interp, gmethod=gmethod, scale=0.8)
    for key in a:
        close(a[key], b[key])


@pytest.mark.parametrize("dmethod", [0, 1, 2])
@pytest.mark.parametrize("gmethod", [0, 1])
def test_diffusion_assembly(pair, dmethod, gmethod):
    name, om, rm = pair
    s, U, phi = fields(om)
    kf = 0.5 + np.abs(phi)
    t, v = scalar_bcs(om.npatch)
    a = om.assemble("diffusion", s, t, v, kf=kf, gmethod=gmethod, dmethod=dmethod, ctx=phi)
    b = rm.assemble("diffusion", s, t, v, kf=kf, gmethod=gmethod, dmethod=dmethod, ctx=phi)
    for key in a:
        close(a[key], b[key])
    t, v = vector_bcs(om.npatch)
    a = om.assemble("diffusion", U, t, v, kf=kf, gmethod=gmethod, dmethod=dmethod, ctx=phi)
    b = rm.assemble("diffusion", U, t, v, kf=kf, gmethod=gmethod, dmethod=dmethod, ctx=phi)
    for key in a:
        close(a[key], b[key])
