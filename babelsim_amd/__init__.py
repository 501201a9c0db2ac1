"""MI355X-native BabelSim hot path: HIP kernels for gfx950 behind a C ABI.

The Python layer only loads ``libbsim.so`` and forwards plain arrays to it.
"""
from ._capi import (BDF2, EULER, SERVE_CANCELLED, SERVE_CONVERGED, SERVE_ERROR,  # noqa: F401
                    SERVE_NOT_CONVERGED, SERVE_NUMERICAL_FAILURE, SERVE_REJECTED, STEP_COMPLETED,
                    STEP_NOT_CONVERGED, STEP_NUMERICAL_FAILURE,
                    IterReport, KOmegaConfig, MeshData, PisoConfig, SimpleConfig, SolveResult, SolverConfig, StepReport, check,
                    komega_config, lib, piso_config, read_field_values, set_graph_replay, simple_config, solver_config,
                    write_field_csv, write_result_metadata)

__all__ = ["MeshData", "SolverConfig", "SolveResult", "SimpleConfig", "IterReport", "PisoConfig", "StepReport",
           "lib", "check", "simple_config", "solver_config", "piso_config", "EULER", "BDF2", "STEP_COMPLETED",
           "STEP_NOT_CONVERGED", "STEP_NUMERICAL_FAILURE", "set_graph_replay", "read_field_values",
           "write_field_csv", "write_result_metadata", "KOmegaConfig", "komega_config", "SERVE_CONVERGED",
           "SERVE_NOT_CONVERGED", "SERVE_NUMERICAL_FAILURE", "SERVE_REJECTED", "SERVE_ERROR", "SERVE_CANCELLED"]
