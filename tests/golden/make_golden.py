#!/usr/bin/env python
"""Writes the operator golden fixtures tests/golden/ops_<mesh>.npz from the
reference's own discretisation layer (src/core/mesh.cpp and
src/discretization/operators.cpp, compiled by oracle/Makefile into
oracle/_ref/libbsimref.so).  Inputs (mesh arrays, fields, boundary
conditions) and outputs are stored together; nothing of the reference's
source goes into the fixture.  Run from the repository root after
`make -C oracle`:  python tests/golden/make_golden.py"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

import golden_cases as gc  # noqa: E402
import oracle_lib as orc  # noqa: E402


def main() -> int:
    if orc.reference() is None:
        print("oracle/_ref/libbsimref.so is not built (make -C oracle with the reference tree present)")
        return 1
    for name in gc.MESHES:
        arrays = gc.mesh_arrays(name)
        rm = orc.reference_mesh(arrays)
        x = gc.inputs(arrays, rm.geometry()["C"])
        out = gc.evaluate(rm, int(arrays["npatch"]), x)
        blob = {f"mesh_{k}": np.asarray(v) for k, v in arrays.items()}
        blob.update({f"in_{k}": v for k, v in x.items()})
        blob.update({f"out_{k}": np.asarray(v) for k, v in out.items()})
        path = HERE / f"ops_{name}.npz"
        np.savez_compressed(path, **blob)
        print(f"{path.name}: {len(out)} outputs, {path.stat().st_size} bytes")
    return 0


if __name__ == "__main__":
    sys.exit(main())
