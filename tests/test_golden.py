"""The oracle against the golden fixtures written from the reference's own
compiled mesh and operator sources (tests/golden/make_golden.py).  Runs
wherever the fixtures are, with or without the reference tree."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import golden_cases as gc
import oracle_lib as orc

GOLDEN = Path(__file__).resolve().parent / "golden"
RTOL = 1e-11


def load(name):
    z = np.load(GOLDEN / f"ops_{name}.npz")
    arrays = {k[5:]: z[k] for k in z.files if k.startswith("mesh_")}
    arrays["npatch"] = int(arrays["npatch"])
    x = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    out = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    return arrays, x, out


def compare(got: dict, want: dict, rtol=RTOL):
    assert set(got) == set(want)
    bad = []
    for key, ref in want.items():
        a = np.asarray(got[key]).reshape(ref.shape)
        scale = max(1.0, float(np.max(np.abs(ref))) if ref.size else 1.0)
        err = float(np.max(np.abs(a - ref))) if ref.size else 0.0
        if not err <= rtol * scale:
            bad.append(f"{key}: {err:.3e} (scale {scale:.2e})")
    assert not bad, "; ".join(bad)


@pytest.mark.parametrize("name", gc.MESHES)
def test_oracle_matches_reference_fixtures(name):
    arrays, x, want = load(name)
    om = orc.oracle_mesh(arrays)
    compare(gc.evaluate(om, arrays["npatch"], x), want)
