"""Python handles over the device library: a case context (mesh in HBM), its
face operators, persistent linear systems and the SIMPLE and PISO drivers.

The method signatures mirror the operator set of BabelSim's
``include/babelsim/operators.h`` (one field at a time, boundary conditions per
patch), so a test can run the same call against the CPU oracle and the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._capi import (EquationCounters, IterReport, KOmegaConfig, MeshData, Performance, PisoConfig, PisoInitial,
                    PisoTurbulence, SimpleConfig, SolveResult, SolverConfig, StepReport, check, dp, ip, lib, ptr)


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _run_performance(fn, handle) -> tuple[dict, list[dict]]:
    """The run's PerformanceCounters and its systems' counters."""
    total = Performance()
    systems = (EquationCounters * 8)()
    count = C.c_int()
    check(fn(handle, C.byref(total), systems, 8, C.byref(count)))
    return total.as_dict(), [systems[k].as_dict() for k in range(count.value)]


class Context:
    """One device, one stream, the mesh resident in HBM."""

    def __init__(self, mesh: MeshData | dict, renumber: bool = True, device: int = 0):
        if isinstance(mesh, dict):
            mesh = MeshData.from_arrays(mesh)
        self._mesh = mesh
        h = C.c_void_p()
        check(lib().bsim_ctx_create(mesh.handle, int(renumber), int(device), C.byref(h)))
        self.h = h
        vals = [C.c_int() for _ in range(5)]
        check(lib().bsim_ctx_sizes(self.h, *[C.byref(v) for v in vals]))
        self.nc, self.nf, self.nfi, self.nnz_off, orth = (v.value for v in vals)
        self.orthogonal = bool(orth)
        self.npatch = mesh.sizes()[3]

    def close(self):
        if getattr(self, "h", None):
            lib().bsim_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def geometry(self) -> dict:
        nc, nf = self.nc, self.nf
        g = dict(C=np.zeros(3 * nc), V=np.zeros(nc), Cf=np.zeros(3 * nf), Sf=np.zeros(3 * nf), Af=np.zeros(nf),
                 n=np.zeros(3 * nf), w=np.zeros(nf), orth=np.zeros(nf), k=np.zeros(3 * nf), skew=np.zeros(3 * nf))
        check(lib().bsim_ctx_geometry(self.h, ptr(g["C"]), ptr(g["V"]), ptr(g["Cf"]), ptr(g["Sf"]), ptr(g["Af"]),
                                      ptr(g["n"]), ptr(g["w"]), ptr(g["orth"]), ptr(g["k"]), ptr(g["skew"])))
        for key in ("C", "Cf", "Sf", "n", "k", "skew"):
            g[key] = g[key].reshape(-1, 3)
        return g

    # ---- operators -----------------------------------------------------
    def gradient(self, x, bct, bcv, method, ctx=None):
        x = _f64(x)
        ncomp = 1 if x.ndim == 1 else 3
        out = np.zeros(self.nc * 3 * ncomp)
        c = None if ctx is None else _f64(ctx)
        check(lib().bsim_op_gradient(self.h, ncomp, ptr(x), ptr(_i32(bct), ip), ptr(_f64(bcv).reshape(-1)), ptr(c),
                                     method, ptr(out)))
        return out.reshape(self.nc, 3) if ncomp == 1 else out.reshape(self.nc, 3, 3)

    def interpolate(self, x, bct, bcv, method, gmethod, ctx=None):
        x = _f64(x)
        ncomp = 1 if x.ndim == 1 else 3
        out = np.zeros(self.nf * ncomp)
        c = None if ctx is None else _f64(ctx)
        check(lib().bsim_op_interpolate(self.h, ncomp, ptr(x), ptr(_i32(bct), ip), ptr(_f64(bcv).reshape(-1)),
                                        ptr(c), method, gmethod, ptr(out)))
        return out if ncomp == 1 else out.reshape(self.nf, 3)

    def flux(self, U, bct, bcv, method, gmethod, ctx=None):
        out = np.zeros(self.nf)
        c = None if ctx is None else _f64(ctx)
        check(lib().bsim_op_flux(self.h, ptr(_f64(U)), ptr(_i32(bct), ip), ptr(_f64(bcv).reshape(-1)), ptr(c), method,
                                 gmethod, ptr(out)))
        return out

    def normal_gradient(self, x, bct, bcv, gmethod, dmethod):
        out = np.zeros(self.nf)
        check(lib().bsim_op_normal_gradient(self.h, ptr(_f64(x)), ptr(_i32(bct), ip), ptr(_f64(bcv).reshape(-1)),
                                            gmethod, dmethod, ptr(out)))
        return out

    def divergence_face(self, phi):
        out = np.zeros(self.nc)
        check(lib().bsim_op_divergence_face(self.h, ptr(_f64(phi)), ptr(out)))
        return out

    def explicit_laplacian(self, x, bct, bcv, kf, k, gmethod, dmethod):
        out = np.zeros(self.nc)
        kfa = None if kf is None else _f64(kf)
        check(lib().bsim_op_explicit_laplacian(self.h, ptr(_f64(x)), ptr(_i32(bct), ip), ptr(_f64(bcv).reshape(-1)),
                                               ptr(kfa), float(k), gmethod, dmethod, ptr(out)))
        return out

    def explicit_convection(self, x, bct, bcv, phi, conv, interp, gmethod):
        x = _f64(x)
        ncomp = 1 if x.ndim == 1 else 3
        out = np.zeros(self.nc * ncomp)
        check(lib().bsim_op_explicit_convection(self.h, ncomp, ptr(x), ptr(_i32(bct), ip),
                                                ptr(_f64(bcv).reshape(-1)), ptr(_f64(phi)), conv, interp, gmethod,
                                                ptr(out)))
        return out if ncomp == 1 else out.reshape(self.nc, 3)

    def assemble(self, kind, x, bct, bcv, phi=None, kf=None, conv=1, interp=0, gmethod=0, dmethod=0, scale=1.0,
                 ctx=None):
        x = _f64(x)
        ncomp = 1 if x.ndim == 1 else 3
        diag, up, lo, src = np.zeros(self.nc), np.zeros(self.nf), np.zeros(self.nf), np.zeros(self.nc * ncomp)
        if kind == "convection":
            check(lib().bsim_op_assemble_convection(self.h, ncomp, ptr(x), ptr(_i32(bct), ip),
                                                    ptr(_f64(bcv).reshape(-1)), ptr(_f64(phi)), conv, interp, gmethod,
                                                    float(scale), ptr(diag), ptr(up), ptr(lo), ptr(src)))
        else:
            c = None if ctx is None else _f64(ctx)
            check(lib().bsim_op_assemble_diffusion(self.h, ncomp, ptr(x), ptr(_i32(bct), ip),
                                                   ptr(_f64(bcv).reshape(-1)), ptr(c), ptr(_f64(kf)), gmethod,
                                                   dmethod, ptr(diag), ptr(up), ptr(lo), ptr(src)))
        return dict(diag=diag, upper=up, lower=lo, source=src if ncomp == 1 else src.reshape(self.nc, 3))


class System:
    """A persistent linear system (ComputeBackend::solve for one equation)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().bsim_system_create(ctx.h, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().bsim_system_destroy(self.h)
            self.h = None

    def solve(self, diag, upper, lower, source, x, cfg: SolverConfig):
        source = _f64(source)
        nrhs = 1 if source.ndim == 1 else 3
        xx = _f64(x).copy()
        res = (SolveResult * nrhs)()
        check(lib().bsim_system_solve_ldu(self.h, ptr(_f64(diag)), ptr(_f64(upper)), ptr(_f64(lower)), nrhs,
                                          ptr(source), ptr(xx), C.byref(cfg), res))
        return xx, list(res)

    def history(self):
        n = C.c_int()
        check(lib().bsim_system_history(self.h, None, 0, C.byref(n)))
        out = np.zeros(n.value)
        check(lib().bsim_system_history(self.h, ptr(out), n.value, C.byref(n)))
        return out

    def counters(self) -> dict:
        out = (C.c_uint64 * 10)()
        check(lib().bsim_system_counters(self.h, out))
        keys = ["matrix_updates", "rhs_only", "pc_setups", "pc_reuses", "retries", "solves", "krylov_iterations",
                "true_residual_checks", "spmvs", "pc_applications"]
        return dict(zip(keys, list(out)))

    def performance(self) -> dict:
        """EquationPerformanceCounters of this system (bsim_system_performance)."""
        out = EquationCounters()
        check(lib().bsim_system_performance(self.h, C.byref(out)))
        return out.as_dict()


def csr_solve(ptr_, col, diag, val, b, x0, cfg: SolverConfig, device=0, hist_cap=100000):
    n = len(diag)
    x = _f64(x0).copy()
    res = SolveResult()
    hist = np.zeros(hist_cap)
    ln = C.c_int()
    check(lib().bsim_csr_solve(device, n, ptr(_i32(ptr_), ip), ptr(_i32(col), ip), ptr(_f64(diag)), ptr(_f64(val)),
                               ptr(_f64(b)), ptr(x), C.byref(cfg), C.byref(res), ptr(hist), hist_cap, C.byref(ln)))
    return x, res, hist[:ln.value].copy()


def csr_solve_noalloc(ptr_, col, diag, val, b, x, cfg: SolverConfig, device=0, stream=0, hist_cap=100000):
    """bsim_csr_solve_noalloc on device arrays: ptr_ and col int32, diag, val, b
    and x float64, each an object with data_ptr() (a torch tensor on the
    device); x holds the initial guess and is overwritten.  stream is a
    hipStream_t handle (0: the null stream)."""
    res = SolveResult()
    hist = np.zeros(hist_cap)
    ln = C.c_int()
    check(lib().bsim_csr_solve_noalloc(device, stream or None, len(diag), ptr_.data_ptr(), col.data_ptr(),
                                       diag.data_ptr(), val.data_ptr(), b.data_ptr(), x.data_ptr(), C.byref(cfg),
                                       C.byref(res), ptr(hist), hist_cap, C.byref(ln)))
    return res, hist[:ln.value].copy()


def csr_spmv(ptr_, col, diag, val, x, device=0):
    y = np.zeros(len(diag))
    check(lib().bsim_csr_spmv(device, len(diag), ptr(_i32(ptr_), ip), ptr(_i32(col), ip), ptr(_f64(diag)),
                              ptr(_f64(val)), ptr(_f64(x)), ptr(y)))
    return y


def amg_hierarchy(ptr_, col, diag, val, cfg: SolverConfig, device=0):
    n = len(diag)
    nl = C.c_int()
    sizes = np.zeros(64, np.int32)
    agg0 = np.full(n, -1, np.int32)
    cd1 = np.zeros(n)
    check(lib().bsim_amg_hierarchy(device, n, ptr(_i32(ptr_), ip), ptr(_i32(col), ip), ptr(_f64(diag)),
                                   ptr(_f64(val)), C.byref(cfg), C.byref(nl), ptr(sizes, ip), ptr(agg0, ip), ptr(cd1)))
    sizes = sizes[:nl.value]
    return sizes, agg0, cd1[:sizes[1]] if len(sizes) > 1 else None


class Simple:
    """Steady SIMPLE resident on the GPU (src/physics/simple/main.cpp)."""

    def __init__(self, ctx: Context, cfg: SimpleConfig, u_bc, p_bc):
        self.ctx = ctx
        self.cfg = cfg
        ut, uv = u_bc
        pt, pv = p_bc
        h = C.c_void_p()
        check(lib().bsim_simple_create(ctx.h, C.byref(cfg), ptr(_i32(ut), ip), ptr(_f64(uv).reshape(-1)),
                                       ptr(_i32(pt), ip), ptr(_f64(pv).reshape(-1)), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().bsim_simple_destroy(self.h)
            self.h = None

    def set_fields(self, U=None, p=None):
        check(lib().bsim_simple_set_fields(self.h, ptr(None if U is None else _f64(U)),
                                           ptr(None if p is None else _f64(p))))

    def enable_komega(self, kw, k_bc, w_bc):
        """turbulenceModel kOmega (src/physics/RANS/k_omega.cpp); before start()."""
        kt, kv = k_bc
        wt, wv = w_bc
        check(lib().bsim_simple_enable_komega(self.h, C.byref(kw), ptr(_i32(kt), ip), ptr(_f64(kv).reshape(-1)),
                                              ptr(_i32(wt), ip), ptr(_f64(wv).reshape(-1))))

    def set_turbulence(self, k=None, omega=None):
        check(lib().bsim_simple_set_turbulence(self.h, ptr(None if k is None else _f64(k)),
                                               ptr(None if omega is None else _f64(omega))))

    def turbulence(self):
        k, w, mut = np.zeros(self.ctx.nc), np.zeros(self.ctx.nc), np.zeros(self.ctx.nc)
        check(lib().bsim_simple_get_turbulence(self.h, ptr(k), ptr(w), ptr(mut)))
        return k, w, mut

    def start(self):
        check(lib().bsim_simple_start(self.h))

    def iterate(self, n=1):
        reps = (IterReport * n)()
        check(lib().bsim_simple_iterate(self.h, n, reps))
        return list(reps)

    def fields(self):
        U = np.zeros((self.ctx.nc, 3))
        p = np.zeros(self.ctx.nc)
        phi = np.zeros(self.ctx.nf)
        check(lib().bsim_simple_get_fields(self.h, ptr(U), ptr(p), ptr(phi)))
        return U, p, phi

    def history(self, which):
        n = C.c_int()
        check(lib().bsim_simple_history(self.h, which, None, 0, C.byref(n)))
        out = np.zeros(n.value)
        check(lib().bsim_simple_history(self.h, which, ptr(out), n.value, C.byref(n)))
        return out

    def timing(self) -> dict:
        out = np.zeros(9)
        check(lib().bsim_simple_timing(self.h, ptr(out)))
        keys = ["assembly", "momentum_solve", "rhie_chow", "pressure_assembly", "pressure_solve", "correction",
                "convergence", "iterations", "turbulence"]
        return dict(zip(keys, out.tolist()))

    def performance(self) -> tuple[dict, list[dict]]:
        """PerformanceCounters of the run (Case::performance's 31 values) and per system."""
        return _run_performance(lib().bsim_simple_performance, self.h)

    def time_spmv(self, which=1, reps=20, cold=False):
        """Mean seconds of one SpMV on the momentum (0) or pressure (1) matrix;
        cold: each launch follows a 512 MiB pass that evicts the Infinity Cache."""
        s = C.c_double()
        n = C.c_int()
        nnz = C.c_int()
        check(lib().bsim_simple_time_spmv(self.h, which, reps, int(cold), C.byref(s), C.byref(n), C.byref(nnz)))
        return s.value, n.value, nnz.value

    def probe_start(self, capacity=8192):
        """Start HIP-event timing of the pressure AMG's fine-level smoother sweeps."""
        check(lib().bsim_simple_probe_start(self.h, capacity))

    def probe_stop(self):
        """Stop the probe: (launches, summed kernel seconds, algorithmic bytes per launch)."""
        n = C.c_int()
        s = C.c_double()
        b = C.c_double()
        check(lib().bsim_simple_probe_stop(self.h, C.byref(n), C.byref(s), C.byref(b)))
        return n.value, s.value, b.value


class Piso:
    """Transient PISO resident on the GPU (src/physics/piso/main.cpp).

    ``step(dt)`` is one ``time.advance()`` of length dt and the body of
    runPiso's time loop; ``run`` adds the TimeStepper of src/io/case.cpp:391-403.
    """

    def __init__(self, ctx: Context, cfg: PisoConfig, u_bc, p_bc):
        self.ctx = ctx
        self.cfg = cfg
        ut, uv = u_bc
        pt, pv = p_bc
        h = C.c_void_p()
        check(lib().bsim_piso_create(ctx.h, C.byref(cfg), ptr(_i32(ut), ip), ptr(_f64(uv).reshape(-1)),
                                     ptr(_i32(pt), ip), ptr(_f64(pv).reshape(-1)), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().bsim_piso_destroy(self.h)
            self.h = None

    def set_fields(self, U=None, p=None):
        check(lib().bsim_piso_set_fields(self.h, ptr(None if U is None else _f64(U)),
                                         ptr(None if p is None else _f64(p))))

    def enable_komega(self, kw: KOmegaConfig, k_bc, w_bc, coupling=0, correctors=0, max_iterations=1000):
        """Transient Wilcox k-omega (piso/main.cpp:98-121), before start()."""
        kt, kv = k_bc
        wt, wv = w_bc
        t = PisoTurbulence(coupling=coupling, correctors=correctors, max_iterations=max_iterations)
        check(lib().bsim_piso_enable_komega(self.h, C.byref(kw), ptr(_i32(kt), ip), ptr(_f64(kv)),
                                            ptr(_i32(wt), ip), ptr(_f64(wv)), C.byref(t)))

    def set_turbulence(self, k=None, omega=None):
        check(lib().bsim_piso_set_turbulence(self.h, ptr(None if k is None else _f64(k)),
                                             ptr(None if omega is None else _f64(omega))))

    def turbulence(self):
        k, w, mut = (np.zeros(self.ctx.nc) for _ in range(3))
        check(lib().bsim_piso_get_turbulence(self.h, ptr(k), ptr(w), ptr(mut)))
        return k, w, mut

    def start(self):
        check(lib().bsim_piso_start(self.h))

    def initial(self) -> PisoInitial:
        """The initial potential-flow projection's record (after start)."""
        rec = PisoInitial()
        check(lib().bsim_piso_initial_report(self.h, C.byref(rec)))
        return rec

    def step(self, dt) -> StepReport:
        rep = StepReport()
        check(lib().bsim_piso_step(self.h, float(dt), C.byref(rep)))
        return rep

    def run(self, start_time, end_time, delta_t, max_steps=None):
        """Steps from start_time to end_time with TimeStepper's rule (src/io/case.cpp:391-403);
        stops at the first step that is not completed, like runPiso.  Returns the reports."""
        stepper = TimeStepper(start_time, end_time, delta_t)
        reports = []
        while not stepper.finished() and (max_steps is None or len(reports) < max_steps):
            dt = stepper.advance()
            rep = self.step(dt)
            reports.append(rep)
            if rep.status != 0:
                break
        return reports

    def fields(self):
        U = np.zeros((self.ctx.nc, 3))
        p = np.zeros(self.ctx.nc)
        phi = np.zeros(self.ctx.nf)
        check(lib().bsim_piso_get_fields(self.h, ptr(U), ptr(p), ptr(phi)))
        return U, p, phi

    def temporal_flux(self):
        out = np.zeros(self.ctx.nf)
        check(lib().bsim_piso_temporal_flux(self.h, ptr(out)))
        return out

    def history(self, which):
        n = C.c_int()
        check(lib().bsim_piso_history(self.h, which, None, 0, C.byref(n)))
        out = np.zeros(n.value)
        check(lib().bsim_piso_history(self.h, which, ptr(out), n.value, C.byref(n)))
        return out

    def timing(self) -> dict:
        out = np.zeros(8)
        check(lib().bsim_piso_timing(self.h, ptr(out)))
        keys = ["history", "assembly", "momentum_solve", "corrector", "pressure_assembly", "pressure_solve",
                "diagnostics", "steps"]
        return dict(zip(keys, out.tolist()))

    def performance(self) -> tuple[dict, list[dict]]:
        """PerformanceCounters of the run and per system (momentum, pressure)."""
        return _run_performance(lib().bsim_piso_performance, self.h)

    def probe_start(self, capacity=8192):
        check(lib().bsim_piso_probe_start(self.h, capacity))

    def probe_stop(self):
        n = C.c_int()
        s = C.c_double()
        b = C.c_double()
        check(lib().bsim_piso_probe_stop(self.h, C.byref(n), C.byref(s), C.byref(b)))
        return n.value, s.value, b.value


class TimeStepper:
    """src/io/case.cpp:388-403: t_{n+1} = start + (n+1) delta_t in long double,
    snapped to end_time within 32 eps of the time scale; dt = t_{n+1} - t_n."""

    def __init__(self, start_time, end_time, delta_t):
        if not (np.isfinite(start_time) and np.isfinite(end_time) and np.isfinite(delta_t)) or \
                end_time < start_time or delta_t <= 0:
            raise ValueError("time control is invalid")
        self.start, self.end, self.delta = float(start_time), float(end_time), float(delta_t)
        self.value = self.start
        self.step = 0

    def finished(self) -> bool:
        return self.value >= self.end

    def advance(self) -> float:
        if self.finished():
            raise RuntimeError("cannot advance beyond endTime")
        eps = np.finfo(np.float64).eps
        tolerance = 32 * eps * max(abs(self.start), abs(self.end), self.delta)
        nxt = float(np.longdouble(self.start) + np.longdouble(self.step + 1) * np.longdouble(self.delta))
        if nxt >= self.end - tolerance:
            nxt = self.end
        dt = nxt - self.value
        if not dt > 0:
            raise RuntimeError("time step is below representable time precision")
        self.value = nxt
        self.step += 1
        return dt
