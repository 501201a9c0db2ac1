"""The ctypes mirrors of the C ABI's structs have the C compiler's layout:
a small C program built with gcc against include/bsim.h prints sizeof and
offsetof of every field, and each must equal the ctypes value."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

from babelsim_amd import _capi

ROOT = Path(__file__).resolve().parents[1]

STRUCTS = {
    "bsim_solver_config": _capi.SolverConfig,
    "bsim_solve_result": _capi.SolveResult,
    "bsim_simple_config": _capi.SimpleConfig,
    "bsim_iter_report": _capi.IterReport,
    "bsim_komega_config": _capi.KOmegaConfig,
    "bsim_piso_config": _capi.PisoConfig,
    "bsim_step_report": _capi.StepReport,
    "bsim_piso_initial": _capi.PisoInitial,
    "bsim_piso_turbulence": _capi.PisoTurbulence,
    "bsim_request": _capi.Request,
    "bsim_request_result": _capi.RequestResult,
    "bsim_equation_counters": _capi.EquationCounters,
    "bsim_performance": _capi.Performance,
}


def test_struct_layouts_match_the_header(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc is not available")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "bsim.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'    printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'    printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    seen = 0
    for line in filter(None, out):
        cname, field, value = line.split()
        cls = STRUCTS[cname]
        expect = C_sizeof(cls) if field == "sizeof" else getattr(cls, field).offset
        assert int(value) == expect, (cname, field, value, expect)
        seen += 1
    assert seen == sum(len(c._fields_) + 1 for c in STRUCTS.values())


def C_sizeof(cls):
    import ctypes
    return ctypes.sizeof(cls)
