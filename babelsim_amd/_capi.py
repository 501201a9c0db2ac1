"""ctypes binding of libbsim.so (the C ABI declared in include/bsim.h).

The HIP library is the product: there is no Python or CPU fallback.  Loading
fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("BSIM_LIBRARY", _PKG / "libbsim.so"))

_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


class SolverConfig(C.Structure):
    """LinearSolverConfig (solver_control.h:165-187) + AMG/precision controls."""

    _fields_ = [("solver", C.c_int), ("pc", C.c_int), ("atol", C.c_double), ("rtol", C.c_double),
                ("max_iter", C.c_int), ("warm_start", C.c_int), ("rebuild_interval", C.c_int),
                ("reuse_growth", C.c_double), ("amg_max_levels", C.c_int), ("amg_coarse_size", C.c_int),
                ("amg_pre_sweeps", C.c_int), ("amg_post_sweeps", C.c_int), ("amg_omega", C.c_double),
                ("amg_strength", C.c_double), ("amg_correction", C.c_double), ("precision", C.c_int),
                ("inner_rtol", C.c_double)]


def solver_config(**kw) -> SolverConfig:
    base = dict(solver=1, pc=2, atol=1e-12, rtol=1e-8, max_iter=1000, warm_start=0, rebuild_interval=8,
                reuse_growth=2.0, amg_max_levels=12, amg_coarse_size=48, amg_pre_sweeps=2, amg_post_sweeps=2,
                amg_omega=0.8, amg_strength=0.5, amg_correction=1.9, precision=0, inner_rtol=1e-3)
    base.update(kw)
    return SolverConfig(**base)


class SolveResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("initial", C.c_double),
                ("final_residual", C.c_double), ("relative", C.c_double)]


class SimpleConfig(C.Structure):
    """Steady SIMPLE controls (src/physics/simple/main.cpp:21-31) and per-equation methods."""

    _fields_ = [("rho", C.c_double), ("mu", C.c_double),
                ("m_interp", C.c_int), ("m_grad", C.c_int), ("m_conv", C.c_int), ("m_diff", C.c_int),
                ("m_coef_interp", C.c_int), ("m_coef_grad", C.c_int),
                ("p_interp", C.c_int), ("p_grad", C.c_int), ("p_diff", C.c_int),
                ("p_coef_interp", C.c_int), ("p_coef_grad", C.c_int),
                ("alpha_u", C.c_double), ("alpha_p", C.c_double), ("non_orth_corr", C.c_int),
                ("tol_mass", C.c_double), ("tol_u", C.c_double), ("tol_mom", C.c_double), ("tol_p", C.c_double),
                ("mom", SolverConfig), ("pre", SolverConfig)]


class IterReport(C.Structure):
    _fields_ = [("rU", C.c_double), ("dU", C.c_double), ("dP", C.c_double), ("mass", C.c_double),
                ("linP", C.c_double), ("mom_iters", C.c_int * 3), ("p_iters", C.c_int),
                ("mom_status", C.c_int), ("p_status", C.c_int), ("converged", C.c_int), ("referenced", C.c_int),
                ("d_turb", C.c_double), ("r_turb", C.c_double), ("bounded_omega", C.c_double),
                ("k_iters", C.c_int), ("w_iters", C.c_int), ("turb_linear", C.c_int), ("turb_converged", C.c_int),
                ("numerical_failure", C.c_int)]


class KOmegaConfig(C.Structure):
    """Wilcox k-omega constants, bounds and controls (src/physics/RANS/k_omega.cpp:26-48) and the
    kTransport / omegaTransport equation methods and solvers."""

    _fields_ = [("beta_star", C.c_double), ("beta", C.c_double), ("gamma", C.c_double), ("sigma_k", C.c_double),
                ("sigma_omega", C.c_double), ("k_min", C.c_double), ("omega_min", C.c_double),
                ("viscosity_ratio_max", C.c_double), ("relaxation", C.c_double), ("tolerance", C.c_double),
                ("k_interp", C.c_int), ("k_grad", C.c_int), ("k_conv", C.c_int), ("k_diff", C.c_int),
                ("k_coef_interp", C.c_int), ("k_coef_grad", C.c_int),
                ("w_interp", C.c_int), ("w_grad", C.c_int), ("w_conv", C.c_int), ("w_diff", C.c_int),
                ("w_coef_interp", C.c_int), ("w_coef_grad", C.c_int),
                ("k_solver", SolverConfig), ("w_solver", SolverConfig)]


def komega_config(**kw) -> KOmegaConfig:
    """Defaults of the reference's KOmega constructor (k_omega.cpp:33-48)."""
    base = dict(beta_star=0.09, beta=3.0 / 40.0, gamma=5.0 / 9.0, sigma_k=0.5, sigma_omega=0.5, k_min=1e-12,
                omega_min=1e-12, viscosity_ratio_max=1e5, relaxation=0.7, tolerance=1e-6,
                k_interp=0, k_grad=0, k_conv=1, k_diff=0, k_coef_interp=0, k_coef_grad=0,
                w_interp=0, w_grad=0, w_conv=1, w_diff=0, w_coef_interp=0, w_coef_grad=0)
    ks = kw.pop("k_solver", None) or solver_config(solver=1, pc=1, rtol=1e-8, atol=1e-14)
    ws = kw.pop("w_solver", None) or solver_config(solver=1, pc=1, rtol=1e-8, atol=1e-14)
    base.update(kw)
    return KOmegaConfig(k_solver=ks, w_solver=ws, **base)


def simple_config(**kw) -> SimpleConfig:
    base = dict(rho=1.0, mu=0.01, m_interp=0, m_grad=0, m_conv=1, m_diff=0, m_coef_interp=0, m_coef_grad=0,
                p_interp=0, p_grad=0, p_diff=0, p_coef_interp=0, p_coef_grad=0, alpha_u=0.7, alpha_p=0.3,
                non_orth_corr=1, tol_mass=1e-6, tol_u=1e-6, tol_mom=1e-6, tol_p=1e-6)
    mom = kw.pop("mom", None) or solver_config(solver=1, pc=1, rtol=1e-8, atol=1e-14)
    pre = kw.pop("pre", None) or solver_config(solver=0, pc=2, rtol=1e-8, atol=1e-14)
    base.update(kw)
    return SimpleConfig(mom=mom, pre=pre, **base)


class PisoConfig(C.Structure):
    """Transient PISO controls (src/physics/piso/main.cpp:83-105) and per-equation methods."""

    _fields_ = [("rho", C.c_double), ("mu", C.c_double),
                ("m_interp", C.c_int), ("m_grad", C.c_int), ("m_conv", C.c_int), ("m_diff", C.c_int),
                ("m_coef_interp", C.c_int), ("m_coef_grad", C.c_int),
                ("p_interp", C.c_int), ("p_grad", C.c_int), ("p_diff", C.c_int),
                ("p_coef_interp", C.c_int), ("p_coef_grad", C.c_int),
                ("time_method", C.c_int), ("max_iterations", C.c_int), ("n_correctors", C.c_int),
                ("max_correctors", C.c_int), ("non_orth_corr", C.c_int),
                ("alpha_u", C.c_double), ("coupling_tol", C.c_double), ("tol_mass", C.c_double),
                ("tol_u", C.c_double), ("tol_mom", C.c_double), ("tol_p", C.c_double),
                ("mom", SolverConfig), ("pre", SolverConfig),
                ("initialize_potential_flow", C.c_int), ("initial_non_orth_corr", C.c_int)]


class PisoInitial(C.Structure):
    """The initial potential-flow projection's record (main.cpp:147-149) and outcome."""

    _fields_ = [("initial_mass", C.c_double), ("initial_u_max", C.c_double),
                ("initial_pressure_solves", C.c_int), ("p_iters", C.c_int), ("done", C.c_int), ("status", C.c_int)]


class StepReport(C.Structure):
    """One PISO time step: the values runPiso reports (main.cpp:282-289) and the outcome."""

    _fields_ = [("dt", C.c_double), ("co_max", C.c_double), ("co_mean", C.c_double), ("mass", C.c_double),
                ("dU", C.c_double), ("rU", C.c_double), ("dP", C.c_double), ("r_coupling_first", C.c_double),
                ("r_coupling", C.c_double), ("lin_u", C.c_double), ("lin_p", C.c_double),
                ("passes", C.c_int), ("correctors", C.c_int), ("mom_iters", C.c_int * 3), ("p_iters", C.c_int),
                ("accepted", C.c_int), ("converged", C.c_int), ("linear_ok", C.c_int), ("status", C.c_int),
                ("d_turb", C.c_double), ("r_turb", C.c_double), ("k_iters", C.c_int), ("w_iters", C.c_int),
                ("turb_linear", C.c_int), ("turb_converged", C.c_int), ("turb_corrections", C.c_int)]


class PisoTurbulence(C.Structure):
    """turbulenceCoupling (0 iterated, 1 segregated), turbulenceCorrectors (<= 0: the reference's
    default) and turbulenceMaxIterations (src/physics/piso/main.cpp:98-120)."""

    _fields_ = [("coupling", C.c_int), ("correctors", C.c_int), ("max_iterations", C.c_int)]


_U64, _F64 = C.c_uint64, C.c_double


class EquationCounters(C.Structure):
    """EquationPerformanceCounters (solver_control.h:36-60) of one persistent system."""

    _fields_ = ([("identity", C.c_char * 64), ("ksp_type", C.c_char * 16), ("pc_type", C.c_char * 16),
                 ("last_status", C.c_char * 24)] +
                [(n, _U64) for n in ("matrix_pattern_builds", "matrix_value_updates", "rhs_only_solves",
                                     "linear_solves", "krylov_iterations", "preconditioner_setups",
                                     "preconditioner_reuses", "preconditioner_retries", "true_residual_checks",
                                     "sparse_matvecs", "preconditioner_applications")] +
                [(n, _F64) for n in ("last_initial_residual", "last_final_residual", "last_relative_residual",
                                     "pattern_build_seconds", "matrix_update_seconds", "preconditioner_seconds",
                                     "rhs_seconds", "solve_seconds", "residual_check_seconds")])

    def as_dict(self) -> dict:
        return {n: (getattr(self, n).decode() if isinstance(getattr(self, n), bytes) else getattr(self, n))
                for n, _ in self._fields_}


class EquMethods(C.Structure):
    """Operator options of an equation (include/babelsim/methods.h)."""

    _fields_ = [("interp", C.c_int), ("grad", C.c_int), ("conv", C.c_int), ("diff", C.c_int),
                ("coef_interp", C.c_int), ("coef_grad", C.c_int)]


class Performance(C.Structure):
    """PerformanceCounters: the 31 values of Case::performance (src/io/case.cpp:72-104), in order."""

    _fields_ = ([(n, _U64) for n in ("linear_solves", "krylov_iterations", "sparse_matvecs", "halo_exchanges",
                                     "halo_bytes", "global_reductions", "equation_assemblies",
                                     "preconditioner_setups", "preconditioner_applications", "output_writes")] +
                [(n, _F64) for n in ("elapsed_seconds", "assembly_seconds", "preconditioner_seconds",
                                     "preconditioner_apply_seconds", "linear_solve_seconds",
                                     "sparse_matvec_seconds", "halo_seconds", "global_reduction_seconds",
                                     "output_seconds")] +
                [(n, _U64) for n in ("matrix_pattern_builds", "matrix_value_updates", "rhs_only_solves",
                                     "true_residual_checks")] +
                [(n, _F64) for n in ("matrix_update_seconds", "rhs_seconds", "residual_check_seconds",
                                     "pattern_build_seconds")] +
                [(n, _U64) for n in ("preconditioner_reuses", "preconditioner_retries",
                                     "reported_iteration_intervals")] +
                [("iteration_compute_seconds", _F64)])

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


EULER, BDF2 = 1, 2
STEP_COMPLETED, STEP_NOT_CONVERGED, STEP_NUMERICAL_FAILURE = 0, 1, 2


def piso_config(**kw) -> PisoConfig:
    """Defaults of runPiso's solution dictionary (main.cpp:83-105)."""
    base = dict(rho=1.0, mu=0.01, m_interp=0, m_grad=0, m_conv=1, m_diff=0, m_coef_interp=0, m_coef_grad=0,
                p_interp=0, p_grad=0, p_diff=0, p_coef_interp=0, p_coef_grad=0, time_method=BDF2,
                max_iterations=1, n_correctors=2, max_correctors=0, non_orth_corr=1, alpha_u=1.0,
                coupling_tol=np.finfo(np.float64).max, tol_mass=1e-8, tol_u=1e-7, tol_mom=1e-6, tol_p=1e-6,
                initialize_potential_flow=0, initial_non_orth_corr=-1)
    mom = kw.pop("mom", None) or solver_config(solver=1, pc=1, rtol=1e-8, atol=1e-14)
    pre = kw.pop("pre", None) or solver_config(solver=0, pc=2, rtol=1e-8, atol=1e-14)
    base.update(kw)
    return PisoConfig(mom=mom, pre=pre, **base)


REQ_STEADY, REQ_TRANSIENT = 0, 1
(SERVE_CONVERGED, SERVE_NOT_CONVERGED, SERVE_NUMERICAL_FAILURE, SERVE_REJECTED, SERVE_ERROR,
 SERVE_CANCELLED) = range(6)


class Request(C.Structure):
    """One served solve: a steady SIMPLE run or a transient PISO run on a registered case."""

    _fields_ = [("kind", C.c_int), ("case_id", C.c_int), ("simple", C.POINTER(SimpleConfig)),
                ("piso", C.POINTER(PisoConfig)), ("u_bc_type", ip), ("u_bc_value", dp), ("p_bc_type", ip),
                ("p_bc_value", dp), ("U0", dp), ("p0", dp), ("max_iterations", C.c_int), ("dt", C.c_double),
                ("arrival", C.c_double)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("slot", C.c_int), ("reused_context", C.c_int),
                ("released", C.c_double), ("admitted", C.c_double), ("started", C.c_double),
                ("finished", C.c_double), ("rU", C.c_double), ("dU", C.c_double), ("dP", C.c_double),
                ("mass", C.c_double), ("reserved_bytes", C.c_double), ("measured_bytes", C.c_double),
                ("message", C.c_char * 256)]


def set_graph_replay(enable: bool) -> None:
    """AMG V-cycles as captured-graph replays (default) or stream launches."""
    check(lib().bsim_set_graph_replay(int(bool(enable))))


def ptr(a: np.ndarray | None, kind=dp):
    if a is None:
        return None
    return a.ctypes.data_as(kind)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"libbsim.so is not built ({LIB_PATH}); run babelsim_amd.build")
        _lib = C.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


def check(code: int) -> None:
    if code != 0:
        msg = lib().bsim_last_error().decode()
        raise (ValueError if code == 2 else RuntimeError)(msg)


def _declare(L: C.CDLL) -> None:
    L.bsim_last_error.restype = C.c_char_p
    L.bsim_mesh_generate_hex.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, dp, C.c_double, C.c_double,
                                         C.c_uint64, C.POINTER(C.c_void_p)]
    L.bsim_mesh_from_arrays.argtypes = [C.c_int, dp, C.c_int, ip, ip, ip, ip, ip, C.c_int, C.POINTER(C.c_void_p)]
    L.bsim_mesh_prepare.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), ip, ip]
    L.bsim_mesh_sizes.argtypes = [C.c_void_p, ip, ip, ip, ip, ip]
    L.bsim_mesh_sizes.restype = None
    L.bsim_mesh_arrays.argtypes = [C.c_void_p, dp, ip, ip, ip, ip, ip]
    L.bsim_mesh_arrays.restype = None
    L.bsim_mesh_free.argtypes = [C.c_void_p]
    L.bsim_mesh_free.restype = None

    for name, args in _DEVICE_SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _declare_void(L)


vp = C.c_void_p
u64p = C.POINTER(C.c_uint64)
_DEVICE_SIGNATURES: dict = {
    "bsim_set_graph_replay": [C.c_int],
    "bsim_mesh_read": [C.c_char_p, C.POINTER(vp)],
    "bsim_mesh_write": [vp, C.c_char_p],
    "bsim_mesh_patch": [vp, C.c_int, C.c_char_p, C.c_int, ip],
    "bsim_write_field_csv": [C.c_char_p, C.c_int, C.c_char_p, C.c_int, ip, dp, C.c_int, dp],
    "bsim_write_result_metadata": [C.c_char_p, C.c_int, C.c_int, vp, ip, C.c_longlong, C.c_char_p, C.c_int,
                                   C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), ip],
    "bsim_read_field_values": [C.c_char_p, C.c_char_p, C.c_int, C.c_int, dp],
    "bsim_ctx_create": [vp, C.c_int, C.c_int, C.POINTER(vp)],
    "bsim_ctx_sizes": [vp, ip, ip, ip, ip, ip],
    "bsim_ctx_geometry": [vp] + [dp] * 10,
    "bsim_op_gradient": [vp, C.c_int, dp, ip, dp, dp, C.c_int, dp],
    "bsim_op_interpolate": [vp, C.c_int, dp, ip, dp, dp, C.c_int, C.c_int, dp],
    "bsim_op_flux": [vp, dp, ip, dp, dp, C.c_int, C.c_int, dp],
    "bsim_op_normal_gradient": [vp, dp, ip, dp, C.c_int, C.c_int, dp],
    "bsim_op_divergence_face": [vp, dp, dp],
    "bsim_op_explicit_laplacian": [vp, dp, ip, dp, dp, C.c_double, C.c_int, C.c_int, dp],
    "bsim_op_explicit_convection": [vp, C.c_int, dp, ip, dp, dp, C.c_int, C.c_int, C.c_int, dp],
    "bsim_op_assemble_convection": [vp, C.c_int, dp, ip, dp, dp, C.c_int, C.c_int, C.c_int, C.c_double,
                                    dp, dp, dp, dp],
    "bsim_op_assemble_diffusion": [vp, C.c_int, dp, ip, dp, dp, dp, C.c_int, C.c_int, dp, dp, dp, dp],
    "bsim_system_create": [vp, C.POINTER(vp)],
    "bsim_system_solve_ldu": [vp, dp, dp, dp, C.c_int, dp, dp, C.POINTER(SolverConfig), C.POINTER(SolveResult)],
    "bsim_system_history": [vp, dp, C.c_int, ip],
    "bsim_system_counters": [vp, u64p],
    "bsim_system_performance": [vp, C.POINTER(EquationCounters)],
    "bsim_field_create": [vp, C.c_int, C.c_int, dp, C.POINTER(vp)],
    "bsim_field_set": [vp, dp],
    "bsim_field_get": [vp, dp],
    "bsim_field_set_boundary": [vp, ip, dp],
    "bsim_field_set_flux": [vp, vp],
    "bsim_equ_create": [vp, C.POINTER(EquMethods), C.POINTER(SolverConfig), C.POINTER(vp)],
    "bsim_equ_reset": [vp],
    "bsim_equ_copy": [vp, C.POINTER(vp)],
    "bsim_equ_add": [vp, vp, C.c_double],
    "bsim_equ_scale": [vp, C.c_double],
    "bsim_equ_div": [vp, vp, C.c_double],
    "bsim_equ_laplacian": [vp, vp, C.c_double, C.c_double],
    "bsim_equ_reaction": [vp, vp, C.c_double, C.c_double],
    "bsim_equ_source": [vp, vp, dp, C.c_double],
    "bsim_equ_ddt": [vp, vp, C.c_double, vp, C.c_double, C.c_int, vp, C.c_double],
    "bsim_equ_relax": [vp, vp, C.c_double],
    "bsim_equ_reference": [vp, C.c_int, C.c_double],
    "bsim_equ_reference_if_unanchored": [vp, C.c_double, ip],
    "bsim_equ_diagonal": [vp, vp],
    "bsim_equ_rhs": [vp, vp],
    "bsim_equ_apply": [vp, vp, vp],
    "bsim_equ_residual": [vp, vp, vp],
    "bsim_equ_face_flux": [vp, vp, vp],
    "bsim_equ_relative_residual": [vp, vp, dp],
    "bsim_equ_solve": [vp, C.POINTER(SolverConfig), C.POINTER(SolveResult)],
    "bsim_equ_performance": [vp, C.POINTER(EquationCounters)],
    "bsim_math_norms": [C.POINTER(vp), C.c_int, dp],
    "bsim_math_relative_change": [vp, vp, dp],
    "bsim_math_relative_magnitude": [vp, vp, dp],
    "bsim_math_flux_balance": [vp, dp],
    "bsim_math_courant": [vp, C.c_double, dp],
    "bsim_math_face_add": [vp, vp, C.c_double, C.c_int],
    "bsim_csr_solve": [C.c_int, C.c_int, ip, ip, dp, dp, dp, dp, C.POINTER(SolverConfig), C.POINTER(SolveResult),
                       dp, C.c_int, ip],
    "bsim_csr_spmv": [C.c_int, C.c_int, ip, ip, dp, dp, dp, dp],
    "bsim_csr_solve_noalloc": [C.c_int, vp, C.c_int, vp, vp, vp, vp, vp, vp, C.POINTER(SolverConfig),
                               C.POINTER(SolveResult), dp, C.c_int, ip],
    "bsim_amg_hierarchy": [C.c_int, C.c_int, ip, ip, dp, dp, C.POINTER(SolverConfig), ip, ip, ip, dp],
    "bsim_simple_create": [vp, C.POINTER(SimpleConfig), ip, dp, ip, dp, C.POINTER(vp)],
    "bsim_simple_set_fields": [vp, dp, dp],
    "bsim_simple_enable_komega": [vp, C.POINTER(KOmegaConfig), ip, dp, ip, dp],
    "bsim_simple_set_turbulence": [vp, dp, dp],
    "bsim_simple_get_turbulence": [vp, dp, dp, dp],
    "bsim_simple_start": [vp],
    "bsim_simple_iterate": [vp, C.c_int, C.POINTER(IterReport)],
    "bsim_simple_get_fields": [vp, dp, dp, dp],
    "bsim_simple_history": [vp, C.c_int, dp, C.c_int, ip],
    "bsim_simple_timing": [vp, dp],
    "bsim_simple_performance": [vp, C.POINTER(Performance), C.POINTER(EquationCounters), C.c_int, ip],
    "bsim_simple_time_spmv": [vp, C.c_int, C.c_int, C.c_int, dp, ip, ip],
    "bsim_simple_probe_start": [vp, C.c_int],
    "bsim_simple_probe_stop": [vp, ip, dp, dp],
    "bsim_piso_create": [vp, C.POINTER(PisoConfig), ip, dp, ip, dp, C.POINTER(vp)],
    "bsim_piso_set_fields": [vp, dp, dp],
    "bsim_piso_enable_komega": [vp, C.POINTER(KOmegaConfig), ip, dp, ip, dp, C.POINTER(PisoTurbulence)],
    "bsim_piso_set_turbulence": [vp, dp, dp],
    "bsim_piso_get_turbulence": [vp, dp, dp, dp],
    "bsim_piso_start": [vp],
    "bsim_piso_initial_report": [vp, C.POINTER(PisoInitial)],
    "bsim_piso_step": [vp, C.c_double, C.POINTER(StepReport)],
    "bsim_piso_get_fields": [vp, dp, dp, dp],
    "bsim_piso_temporal_flux": [vp, dp],
    "bsim_piso_history": [vp, C.c_int, dp, C.c_int, ip],
    "bsim_piso_timing": [vp, dp],
    "bsim_piso_performance": [vp, C.POINTER(Performance), C.POINTER(EquationCounters), C.c_int, ip],
    "bsim_piso_probe_start": [vp, C.c_int],
    "bsim_piso_probe_stop": [vp, ip, dp, dp],
    "bsim_server_create": [C.c_int, C.c_int, C.c_double, C.POINTER(vp)],
    "bsim_server_add_case": [vp, vp, ip],
    "bsim_server_submit": [vp, C.POINTER(Request), ip],
    "bsim_server_wait": [vp, C.c_int, C.POINTER(RequestResult)],
    "bsim_server_done": [vp, C.c_int, ip],
    "bsim_server_fields": [vp, C.c_int, dp, dp, dp],
    "bsim_server_release": [vp, C.c_int],
    "bsim_server_stats": [vp, dp],
    "bsim_server_estimate": [C.c_int, C.c_int, C.c_int, dp, dp],
    "bsim_device_pool_stats": [C.c_int, dp],
    "bsim_device_pool_trim": [C.c_int],
}
_VOID = {"bsim_ctx_destroy": [vp], "bsim_system_destroy": [vp], "bsim_simple_destroy": [vp],
         "bsim_piso_destroy": [vp], "bsim_server_destroy": [vp], "bsim_field_destroy": [vp],
         "bsim_equ_destroy": [vp]}


def _declare_void(L):
    for name, args in _VOID.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = None
    L.bsim_server_clock.argtypes = [vp]
    L.bsim_server_clock.restype = C.c_double


class MeshData:
    """A polyhedral face list (Mesh::polyhedral inputs) held by the library."""

    def __init__(self, handle: int):
        self.handle = C.c_void_p(handle)

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.bsim_mesh_free(self.handle)
            self.handle = None

    @classmethod
    def hex(cls, n, lo=(0, 0, 0), hi=(1, 1, 1), vertices=None, jitter=0.0, split=0.0, seed=1) -> "MeshData":
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        v = None if vertices is None else np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1)
        h = C.c_void_p()
        check(lib().bsim_mesh_generate_hex(int(n[0]), int(n[1]), int(n[2]), ptr(lo), ptr(hi), ptr(v),
                                           float(jitter), float(split), int(seed), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_arrays(cls, a: dict) -> "MeshData":
        v = np.ascontiguousarray(a["vertices"], dtype=np.float64).reshape(-1)
        fpo = np.ascontiguousarray(a["fpo"], dtype=np.int32)
        fpi = np.ascontiguousarray(a["fpi"], dtype=np.int32)
        own = np.ascontiguousarray(a["owner"], dtype=np.int32)
        nei = np.ascontiguousarray(a["neighbour"], dtype=np.int32)
        pat = np.ascontiguousarray(a["patch"], dtype=np.int32)
        h = C.c_void_p()
        check(lib().bsim_mesh_from_arrays(len(v) // 3, ptr(v), len(own), ptr(fpo, ip), ptr(fpi, ip), ptr(own, ip),
                                          ptr(nei, ip), ptr(pat, ip), int(a["npatch"]), C.byref(h)))
        return cls(h.value)

    def sizes(self):
        vals = [C.c_int() for _ in range(5)]
        lib().bsim_mesh_sizes(self.handle, *[C.byref(x) for x in vals])
        return tuple(x.value for x in vals)  # nv, nf, nfp, npatch, ncells

    def arrays(self) -> dict:
        nv, nf, nfp, npatch, nc = self.sizes()
        out = dict(vertices=np.zeros(3 * nv), fpo=np.zeros(nf + 1, np.int32), fpi=np.zeros(nfp, np.int32),
                   owner=np.zeros(nf, np.int32), neighbour=np.zeros(nf, np.int32), patch=np.zeros(nf, np.int32))
        lib().bsim_mesh_arrays(self.handle, ptr(out["vertices"]), ptr(out["fpo"], ip), ptr(out["fpi"], ip),
                               ptr(out["owner"], ip), ptr(out["neighbour"], ip), ptr(out["patch"], ip))
        out["npatch"] = npatch
        out["ncells"] = nc
        return out

    def prepare(self, renumber: bool = True):
        nv, nf, nfp, npatch, nc = self.sizes()
        cperm = np.zeros(nc, np.int32)
        fperm = np.zeros(nf, np.int32)
        h = C.c_void_p()
        check(lib().bsim_mesh_prepare(self.handle, int(renumber), C.byref(h), ptr(cperm, ip), ptr(fperm, ip)))
        return MeshData(h.value), cperm, fperm

    @classmethod
    def read(cls, path) -> "MeshData":
        """A BABELSIM_MESH 3 file (src/io/mesh_reader.cpp:44-111)."""
        h = C.c_void_p()
        check(lib().bsim_mesh_read(str(path).encode(), C.byref(h)))
        return cls(h.value)

    def write(self, path) -> None:
        check(lib().bsim_mesh_write(self.handle, str(path).encode()))

    def patches(self) -> list:
        """[(name, kind)] with kind one of PATCH_KINDS."""
        out = []
        buf = C.create_string_buffer(256)
        for k in range(self.sizes()[3]):
            kind = C.c_int()
            check(lib().bsim_mesh_patch(self.handle, k, buf, len(buf), C.byref(kind)))
            out.append((buf.value.decode(), PATCH_KINDS[kind.value]))
        return out


PATCH_KINDS = ("generic", "wall", "inlet", "outlet", "symmetry", "processor")


def write_field_csv(time_dir, name, values, centres, global_ids=None, rank=0) -> None:
    """<time_dir>/rank-NNNN/<name>.csv (src/parallel/parallel_writer.cpp:42-75)."""
    v = np.ascontiguousarray(values, np.float64)
    n = v.shape[0]
    ncomp = 1 if v.ndim == 1 else int(np.prod(v.shape[1:]))
    c = np.ascontiguousarray(centres, np.float64).reshape(n, 3)
    g = None if global_ids is None else np.ascontiguousarray(global_ids, np.int32)
    check(lib().bsim_write_field_csv(str(time_dir).encode(), int(rank), name.encode(), n, ptr(g, ip), ptr(c), ncomp,
                                     ptr(v)))


def write_result_metadata(time_dir, mesh: "MeshData", time_name, fields, global_ids=None, rank=0, ranks=1,
                          global_cells=None) -> None:
    """metadata.bs and mesh.geometry (parallel_writer.cpp:116-169); fields are
    (name, type, location) with location "cell", "face" or "vertex"."""
    locs = {"cell": 0, "face": 1, "vertex": 2}
    names = (C.c_char_p * len(fields))(*[f[0].encode() for f in fields])
    types = (C.c_char_p * len(fields))(*[f[1].encode() for f in fields])
    where = np.array([locs[f[2]] for f in fields], np.int32)
    g = None if global_ids is None else np.ascontiguousarray(global_ids, np.int32)
    nc = mesh.sizes()[4]
    check(lib().bsim_write_result_metadata(str(time_dir).encode(), int(rank), int(ranks), mesh.handle, ptr(g, ip),
                                           int(nc if global_cells is None else global_cells),
                                           str(time_name).encode(), len(fields), names, types, ptr(where, ip)))


def read_field_values(path, ncells, ncomp=1, name="") -> np.ndarray:
    """Nonuniform cell values by global id (src/io/field_reader.cpp:143-241)."""
    out = np.zeros(ncells * ncomp)
    check(lib().bsim_read_field_values(str(path).encode(), name.encode(), int(ncells), int(ncomp), ptr(out)))
    return out if ncomp == 1 else out.reshape(ncells, ncomp)
