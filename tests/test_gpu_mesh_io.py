"""A mesh file through the device path: the reference's channel fixture read
with bsim_mesh_read, renumbered on the device, stepped by SIMPLE, and its
fields written and read back by the caller's original cell ids."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import cases
from babelsim_amd import MeshData, read_field_values, write_field_csv, write_result_metadata
from babelsim_amd.device import Context, Simple

pytestmark = pytest.mark.gpu
MESHES = Path(__file__).resolve().parent / "golden" / "meshes"


def test_channel_mesh_file_on_device(tmp_path):
    mesh = MeshData.read(MESHES / "babelsim_channel.mesh")
    ctx = Context(mesh, renumber=True)
    g = ctx.geometry()
    np.testing.assert_allclose(g["V"], 1.0, rtol=1e-12)
    np.testing.assert_allclose(g["C"][0], (0.5, 0.5, 0.5), rtol=1e-12)
    names = [n for n, _ in mesh.patches()]
    # inlet (1,0,0), outlet zeroGradient U and p = 0, no-slip walls, symmetry front/back
    ut = np.array([cases.FIXED_VALUE, cases.ZERO_GRADIENT, cases.FIXED_VALUE, cases.FIXED_VALUE, cases.SYMMETRY,
                   cases.SYMMETRY], np.int32)
    uv = np.zeros((6, 3))
    uv[names.index("inlet")] = (1.0, 0.0, 0.0)
    pt = np.array([cases.ZERO_GRADIENT, cases.FIXED_VALUE, cases.ZERO_GRADIENT, cases.ZERO_GRADIENT, cases.SYMMETRY,
                   cases.SYMMETRY], np.int32)
    s = Simple(ctx, cases.c1_config(), (ut, uv), (pt, np.zeros(6)))
    s.start()
    reps = s.iterate(5)
    assert all(np.isfinite(r.rU) for r in reps)
    U, p, _ = s.fields()
    write_field_csv(tmp_path / "final", "U", U, g["C"])
    write_field_csv(tmp_path / "final", "p", p, g["C"])
    write_result_metadata(tmp_path / "final", mesh, "final", [("U", "vector", "cell"), ("p", "scalar", "cell")])
    np.testing.assert_array_equal(read_field_values(tmp_path / "final", ctx.nc, 3, name="U"), U)
    np.testing.assert_array_equal(read_field_values(tmp_path / "final", ctx.nc, 1, name="p"), p)
    # a restart from the written values gives the same next iteration
    s2 = Simple(ctx, cases.c1_config(), (ut, uv), (pt, np.zeros(6)))
    s2.set_fields(read_field_values(tmp_path / "final", ctx.nc, 3, name="U"),
                  read_field_values(tmp_path / "final", ctx.nc, 1, name="p"))
    s2.start()
    assert np.isfinite(s2.iterate(1)[0].rU)
