"""The reference-side binding printed in INTEGRATION.md (a ComputeBackend,
src/internal/compute_backend.h, over include/bsim.h) compiles against the
reference's own headers: the C++ block is extracted and run through
`g++ -std=c++17 -fsyntax-only`.  Skipped where the reference sources or an
mpi.h (include/babelsim/parallel.h includes it) are absent."""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REFERENCE = Path("/root/reference")
MPI_DIRS = ["/opt/conda/include", "/usr/include/mpi", "/usr/lib/x86_64-linux-gnu/openmpi/include",
            "/usr/include/x86_64-linux-gnu/mpich"]


def binding_source() -> str:
    text = (ROOT / "INTEGRATION.md").read_text()
    section = text[text.index("## Reference-side binding (C++)"):]
    return section.split("```cpp\n", 1)[1].split("```", 1)[0]


def mpi_include():
    for d in MPI_DIRS:
        if (Path(d) / "mpi.h").is_file():
            return d
    return None


def test_binding_names_every_solver_and_preconditioner_type():
    src = binding_source()
    for name in ("ConjugateGradient", "BiCGSTAB", "PreconditionerType::None", "PreconditionerType::Jacobi",
                 "PreconditionerType::Hypre", "PreconditionerType::GAMG", "PreconditionerType::BlockJacobi"):
        assert name in src, name
    # Types the library does not implement are refused, not substituted.
    assert src.count("throw std::invalid_argument") >= 3  # Krylov type, PC type, AMG controls
    assert "bsim_system_performance" in src


@pytest.mark.skipif(not (REFERENCE / "include" / "babelsim").is_dir(), reason="reference sources absent")
@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_binding_compiles_against_the_reference_headers(tmp_path):
    mpi = mpi_include()
    if mpi is None:
        pytest.skip("no mpi.h for include/babelsim/parallel.h")
    path = tmp_path / "hip_backend.cpp"
    path.write_text(binding_source())
    cmd = ["g++", "-std=c++17", "-fsyntax-only", f"-I{REFERENCE / 'include'}", f"-I{REFERENCE / 'src'}",
           f"-I{mpi}", f"-I{ROOT / 'include'}", str(path)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300, env=dict(os.environ))
    assert r.returncode == 0, r.stderr[-4000:]
