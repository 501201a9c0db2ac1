"""MI355X-native BabelSim hot path: HIP kernels for gfx950 behind a C ABI.

The Python layer only loads ``libbsim.so`` and forwards plain arrays to it.
"""
from ._capi import (IterReport, MeshData, SimpleConfig, SolveResult, SolverConfig, check, lib,  # noqa: F401
                    simple_config, solver_config)

__all__ = ["MeshData", "SolverConfig", "SolveResult", "SimpleConfig", "IterReport", "lib", "check",
           "simple_config", "solver_config"]
