"""Python handle over the serving runtime of libbsim.so (bsim_server_*).

A :class:`Server` owns worker threads on one device.  Meshes are registered
once as cases (their contexts stay resident in HBM between requests); each
request is a steady SIMPLE run to convergence or a transient PISO run of a
fixed number of steps, admitted by a device-memory budget as it arrives and
run on its own stream next to the others.  The reference has no counterpart:
its RunTime is single-active per thread (src/runtime/runtime.cpp:15, 48-51).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._capi import (REQ_STEADY, REQ_TRANSIENT, MeshData, PisoConfig, Request, RequestResult,  # noqa: F401
                    SimpleConfig, check, dp, ip, lib, ptr)

STATS_KEYS = ("budget", "peak_reserved", "reserved", "peak_running", "running", "completed", "rejected",
              "contexts_created", "context_reuses", "evictions", "resident_idle")


@dataclass
class Result:
    id: int
    status: int
    iterations: int
    slot: int
    reused_context: bool
    released: float
    admitted: float
    started: float
    finished: float
    rU: float
    dU: float
    dP: float
    mass: float
    reserved_bytes: float
    measured_bytes: float
    message: str

    @property
    def latency(self) -> float:
        """Seconds from eligibility (arrival) to the finished result."""
        return self.finished - self.released


def estimate(kind: int, ncells: int, nfaces: int) -> tuple[float, float]:
    """Admission estimates in bytes: (one context, one request's solver)."""
    a, b = C.c_double(), C.c_double()
    check(lib().bsim_server_estimate(int(kind), int(ncells), int(nfaces), C.byref(a), C.byref(b)))
    return a.value, b.value


class Server:
    def __init__(self, slots: int = 4, memory_budget: float = 64e9, device: int = 0):
        h = C.c_void_p()
        check(lib().bsim_server_create(int(device), int(slots), float(memory_budget), C.byref(h)))
        self.h = h
        self._cases: dict[int, tuple[MeshData, int, int]] = {}

    def close(self):
        if getattr(self, "h", None):
            lib().bsim_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self) -> float:
        return lib().bsim_server_clock(self.h)

    def add_case(self, mesh: MeshData | dict) -> int:
        if isinstance(mesh, dict):
            mesh = MeshData.from_arrays(mesh)
        cid = C.c_int()
        check(lib().bsim_server_add_case(self.h, mesh.handle, C.byref(cid)))
        _, nf, _, _, nc = mesh.sizes()
        self._cases[cid.value] = (mesh, nc, nf)
        return cid.value

    def _submit(self, kind, case, simple, piso, u_bc, p_bc, max_iterations, dt, arrival, U0, p0) -> int:
        ut, uv = np.ascontiguousarray(u_bc[0], np.int32), np.ascontiguousarray(u_bc[1], np.float64).reshape(-1)
        pt, pv = np.ascontiguousarray(p_bc[0], np.int32), np.ascontiguousarray(p_bc[1], np.float64).reshape(-1)
        u0 = None if U0 is None else np.ascontiguousarray(U0, np.float64).reshape(-1)
        p0_ = None if p0 is None else np.ascontiguousarray(p0, np.float64).reshape(-1)
        req = Request(kind=kind, case_id=case, simple=C.pointer(simple) if simple is not None else None,
                      piso=C.pointer(piso) if piso is not None else None, u_bc_type=ptr(ut, ip), u_bc_value=ptr(uv),
                      p_bc_type=ptr(pt, ip), p_bc_value=ptr(pv), U0=ptr(u0), p0=ptr(p0_),
                      max_iterations=int(max_iterations), dt=float(dt), arrival=float(arrival))
        rid = C.c_int()
        check(lib().bsim_server_submit(self.h, C.byref(req), C.byref(rid)))
        return rid.value

    def submit_steady(self, case: int, cfg: SimpleConfig, u_bc, p_bc, max_iterations: int = 1000,
                      arrival: float = 0.0, U0=None, p0=None) -> int:
        """SIMPLE to convergence (outer tolerances of cfg), at most max_iterations outer iterations."""
        return self._submit(REQ_STEADY, case, cfg, None, u_bc, p_bc, max_iterations, 0.0, arrival, U0, p0)

    def submit_transient(self, case: int, cfg: PisoConfig, u_bc, p_bc, steps: int, dt: float,
                         arrival: float = 0.0, U0=None, p0=None) -> int:
        """`steps` PISO time steps of length dt from the initial fields."""
        return self._submit(REQ_TRANSIENT, case, None, cfg, u_bc, p_bc, steps, dt, arrival, U0, p0)

    def wait(self, rid: int) -> Result:
        r = RequestResult()
        check(lib().bsim_server_wait(self.h, int(rid), C.byref(r)))
        return Result(rid, r.status, r.iterations, r.slot, bool(r.reused_context), r.released, r.admitted,
                      r.started, r.finished, r.rU, r.dU, r.dP, r.mass, r.reserved_bytes, r.measured_bytes,
                      r.message.decode(errors="replace"))

    def done(self, rid: int) -> bool:
        d = C.c_int()
        check(lib().bsim_server_done(self.h, int(rid), C.byref(d)))
        return bool(d.value)

    def fields(self, rid: int, case: int):
        _, nc, nf = self._cases[case]
        U, p, phi = np.zeros(3 * nc), np.zeros(nc), np.zeros(nf)
        check(lib().bsim_server_fields(self.h, int(rid), ptr(U), ptr(p), ptr(phi)))
        return U.reshape(nc, 3), p, phi

    def release(self, rid: int) -> None:
        check(lib().bsim_server_release(self.h, int(rid)))

    def stats(self) -> dict:
        out = np.zeros(len(STATS_KEYS))
        check(lib().bsim_server_stats(self.h, out.ctypes.data_as(dp)))
        return dict(zip(STATS_KEYS, out.tolist()))


POOL_KEYS = ("cached_bytes", "live_bytes", "hits", "misses", "released")


def pool_stats(device: int = 0) -> dict:
    """The device block cache: cached and live bytes, cache hits, runtime
    allocations and blocks returned to the runtime."""
    out = np.zeros(len(POOL_KEYS))
    check(lib().bsim_device_pool_stats(int(device), out.ctypes.data_as(dp)))
    return dict(zip(POOL_KEYS, out.tolist()))


def pool_trim(device: int = 0) -> None:
    check(lib().bsim_device_pool_trim(int(device)))
