"""Host mirror of the procedural equation seam on the device (bsim_field_*,
bsim_equ_*, bsim_math_*): Field (include/babelsim/field.h), Equation
(include/babelsim/equ.h) and the math / diagnostics reductions, with the
reference's names, argument meaning and exceptions (ValueError for
std::invalid_argument, RuntimeError otherwise).  Fields and matrices stay in
HBM; only the values a caller asks for cross PCIe."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._capi import EquationCounters, EquMethods, SolveResult, SolverConfig, check, dp, ip, lib, ptr
from .device import Context

CELL, FACE = 0, 1
STEADY, EULER, BDF2 = 0, 1, 2
ALL_FACES, INTERIOR_FACES = 0, 1


class Field:
    """A cell field (ncomp 1 or 3, boundary conditions per patch, zero gradient
    until set) or a face field (one value per face) resident on a context."""

    def __init__(self, ctx: Context, location: int = CELL, ncomp: int = 1, values=None, bc=None):
        self.ctx = ctx
        self.location, self.ncomp = location, ncomp
        h = C.c_void_p()
        vals = None if values is None else np.ascontiguousarray(values, np.float64)
        check(lib().bsim_field_create(ctx.h, location, ncomp, ptr(vals), C.byref(h)))
        self.h = h
        self._flux = None
        if bc is not None:
            self.set_boundary(*bc)

    def __del__(self):
        if getattr(self, "h", None):
            lib().bsim_field_destroy(self.h)
            self.h = None

    @property
    def size(self) -> int:
        return self.ctx.nc * self.ncomp if self.location == CELL else self.ctx.nf

    def set(self, values) -> None:
        v = np.ascontiguousarray(values, np.float64)
        if v.size != self.size:
            raise ValueError(f"field holds {self.size} values, got {v.size}")
        check(lib().bsim_field_set(self.h, ptr(v)))

    def values(self) -> np.ndarray:
        out = np.zeros(self.size)
        check(lib().bsim_field_get(self.h, ptr(out)))
        return out.reshape(-1, 3) if self.ncomp == 3 else out

    def set_boundary(self, types, values) -> None:
        t = np.ascontiguousarray(types, np.int32)
        v = np.ascontiguousarray(np.asarray(values, np.float64).reshape(-1))
        if v.size != t.size * self.ncomp:
            raise ValueError("one boundary value per patch and component")
        check(lib().bsim_field_set_boundary(self.h, ptr(t, ip), ptr(v)))

    def set_boundary_flux(self, flux: "Field") -> None:
        check(lib().bsim_field_set_flux(self.h, flux.h))
        self._flux = flux


def methods(interp=0, grad=0, conv=1, diff=0, coef_interp=0, coef_grad=0) -> EquMethods:
    return EquMethods(interp, grad, conv, diff, coef_interp, coef_grad)


class Equation:
    """equ::Equation<T> bound to `unknown`: operators add to A, sources to b,
    assembly freezes its inputs; solve() writes the unknown in place."""

    def __init__(self, unknown: Field, linear: SolverConfig, options: EquMethods | None = None, _handle=None):
        self.x = unknown
        self.linear = linear
        self._keep = []
        if _handle is not None:
            self.h = _handle
            return
        h = C.c_void_p()
        check(lib().bsim_equ_create(unknown.h, C.byref(options or methods()), C.byref(linear), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().bsim_equ_destroy(self.h)
            self.h = None

    def reset(self):
        check(lib().bsim_equ_reset(self.h))
        self._keep = []

    def copy(self) -> "Equation":
        h = C.c_void_p()
        check(lib().bsim_equ_copy(self.h, C.byref(h)))
        e = Equation(self.x, self.linear, _handle=h)
        e._keep = list(self._keep)
        return e

    def add(self, other: "Equation", factor=1.0):
        check(lib().bsim_equ_add(self.h, other.h, factor))
        self._keep += other._keep

    def scale(self, factor):
        check(lib().bsim_equ_scale(self.h, factor))

    def div(self, flux: Field, factor=1.0):
        check(lib().bsim_equ_div(self.h, flux.h, factor))
        self._keep.append(flux)

    def laplacian(self, coefficient, multiplier):
        """coefficient: a number, or a cell or face Field."""
        if isinstance(coefficient, Field):
            check(lib().bsim_equ_laplacian(self.h, coefficient.h, 0.0, multiplier))
        else:
            check(lib().bsim_equ_laplacian(self.h, None, float(coefficient), multiplier))

    def reaction(self, coefficient, factor=1.0):
        if isinstance(coefficient, Field):
            check(lib().bsim_equ_reaction(self.h, coefficient.h, 0.0, factor))
        else:
            check(lib().bsim_equ_reaction(self.h, None, float(coefficient), 1.0))

    def source(self, value, factor=1.0):
        """value: a constant (scalar or 3-vector) or a cell Field."""
        if isinstance(value, Field):
            check(lib().bsim_equ_source(self.h, value.h, None, factor))
        else:
            c = np.zeros(3)
            c[:self.x.ncomp] = np.asarray(value, np.float64).reshape(-1)[:self.x.ncomp]
            check(lib().bsim_equ_source(self.h, None, ptr(c), 1.0))

    def ddt(self, capacity, previous: Field, dt, method=EULER, older: Field | None = None, previous_dt=0.0):
        cap = capacity.h if isinstance(capacity, Field) else None
        val = 0.0 if isinstance(capacity, Field) else float(capacity)
        check(lib().bsim_equ_ddt(self.h, cap, val, previous.h, dt, method, older.h if older is not None else None,
                                 previous_dt))

    def relax(self, previous: Field, alpha):
        check(lib().bsim_equ_relax(self.h, previous.h, alpha))

    def reference(self, cell, value):
        check(lib().bsim_equ_reference(self.h, cell, value))

    def reference_if_unanchored(self, value) -> bool:
        applied = C.c_int()
        check(lib().bsim_equ_reference_if_unanchored(self.h, value, C.byref(applied)))
        return bool(applied.value)

    def _cell(self, ncomp):
        return Field(self.x.ctx, CELL, ncomp)

    def diagonal(self) -> np.ndarray:
        out = self._cell(1)
        check(lib().bsim_equ_diagonal(self.h, out.h))
        return out.values()

    def rhs(self) -> np.ndarray:
        out = self._cell(self.x.ncomp)
        check(lib().bsim_equ_rhs(self.h, out.h))
        return out.values()

    def apply(self, x: Field) -> np.ndarray:
        out = self._cell(self.x.ncomp)
        check(lib().bsim_equ_apply(self.h, x.h, out.h))
        return out.values()

    def residual(self, x: Field) -> np.ndarray:
        out = self._cell(self.x.ncomp)
        check(lib().bsim_equ_residual(self.h, x.h, out.h))
        return out.values()

    def face_flux(self, solution: Field) -> np.ndarray:
        out = Field(self.x.ctx, FACE, 1)
        check(lib().bsim_equ_face_flux(self.h, solution.h, out.h))
        return out.values()

    def relative_residual(self, x: Field) -> float:
        out = C.c_double()
        check(lib().bsim_equ_relative_residual(self.h, x.h, C.byref(out)))
        return out.value

    def solve(self, cfg: SolverConfig | None = None):
        res = (SolveResult * 3)()
        check(lib().bsim_equ_solve(self.h, C.byref(cfg) if cfg is not None else None, res))
        return list(res)[:self.x.ncomp]

    def performance(self) -> dict:
        out = EquationCounters()
        check(lib().bsim_equ_performance(self.h, C.byref(out)))
        return out.as_dict()


def norms(*fields: Field) -> list[float]:
    arr = (C.c_void_p * len(fields))(*[f.h.value for f in fields])
    out = np.zeros(len(fields))
    check(lib().bsim_math_norms(arr, len(fields), ptr(out)))
    return out.tolist()


def relative_change(current: Field, previous: Field) -> float:
    out = C.c_double()
    check(lib().bsim_math_relative_change(current.h, previous.h, C.byref(out)))
    return out.value


def relative_magnitude(value: Field, reference: Field) -> float:
    out = C.c_double()
    check(lib().bsim_math_relative_magnitude(value.h, reference.h, C.byref(out)))
    return out.value


def flux_balance(flux: Field) -> dict:
    out = np.zeros(4)
    check(lib().bsim_math_flux_balance(flux.h, ptr(out)))
    return dict(zip(("l1", "l2", "maximum", "relative"), out.tolist()))


def courant(flux: Field, dt: float) -> dict:
    out = np.zeros(2)
    check(lib().bsim_math_courant(flux.h, dt, ptr(out)))
    return dict(zip(("maximum", "mean"), out.tolist()))


def face_add(increment: Field, target: Field, factor=1.0, region=ALL_FACES) -> None:
    check(lib().bsim_math_face_add(increment.h, target.h, factor, region))
