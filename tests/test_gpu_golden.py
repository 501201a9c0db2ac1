"""The HIP operators (through the C ABI) against the golden fixtures written
from the reference's own compiled mesh and operator sources: geometry,
gradients, interpolation, fluxes, normal gradients, Laplacians, divergence
and the convection/diffusion assembly on five small meshes."""
from __future__ import annotations

import pytest

import golden_cases as gc
from test_golden import compare, load

pytestmark = pytest.mark.gpu

RTOL = 1e-10  # device summation orders differ from the reference's face loops


@pytest.mark.parametrize("renumber", [True, False])
@pytest.mark.parametrize("name", gc.MESHES)
def test_device_matches_reference_fixtures(name, renumber):
    from babelsim_amd.device import Context
    arrays, x, want = load(name)
    print(f"[golden] mesh={name} renumber={renumber}", flush=True)
    ctx = Context(arrays, renumber=renumber)
    compare(gc.evaluate(ctx, arrays["npatch"], x), want, rtol=RTOL)
