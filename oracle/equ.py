"""Oracle (test infrastructure): a NumPy restatement of the reference's
procedural equation layer -- src/discretization/procedural_equation.cpp,
the math reductions of src/discretization/field_math.cpp and the diagnostics
of src/discretization/fvm_execution.cpp:91-203 -- on top of the oracle's C
operators (liboracle.so through tests/oracle_lib.py).  The coefficients are
kept as the reference keeps them, in LDU form (include/babelsim/
discrete_equation.h:14-50): diagonal and source per cell, upper (row owner)
and lower (row neighbour) per face.  Only tests import this module."""
from __future__ import annotations

import numpy as np

import oracle_lib as orc

CELL, FACE = 0, 1
STEADY, EULER, BDF2 = 0, 1, 2
ORTHOGONAL, CORRECTED = 0, 1
UPWIND, LINEAR_UPWIND, CENTRAL = 0, 1, 2
ZERO_GRADIENT = 2


class Field:
    """Field<T> on an oracle mesh: cell values (ncomp 1 or 3) with a boundary
    condition per patch, or one value per face."""

    def __init__(self, mesh: orc.CpuMesh, location=CELL, ncomp=1, values=None, bc=None):
        self.m, self.location, self.ncomp = mesh, location, ncomp
        n = mesh.nc if location == CELL else mesh.nf
        shape = (n,) if ncomp == 1 else (n, 3)
        self.v = np.zeros(shape) if values is None else np.array(values, np.float64).reshape(shape)
        self.bct = np.full(mesh.npatch, ZERO_GRADIENT, np.int32)
        self.bcv = np.zeros(mesh.npatch * ncomp)
        self.flux = None
        if bc is not None:
            self.set_boundary(*bc)

    def set_boundary(self, types, values):
        self.bct = np.array(types, np.int32)
        self.bcv = np.array(values, np.float64).reshape(-1)

    def set_boundary_flux(self, flux: "Field"):
        """Field::setBoundaryFlux (field.h:277-282): a snapshot of the values."""
        if self.location != CELL or flux.location != FACE:
            raise ValueError("boundary flux must be a face field on the same mesh")
        self.flux = flux.v.copy()

    def set(self, values):
        self.v = np.array(values, np.float64).reshape(self.v.shape)

    def values(self):
        return self.v.copy()

    def ctx(self):
        return self.flux


def methods(interp=0, grad=0, conv=1, diff=0, coef_interp=0, coef_grad=0) -> dict:
    return dict(interp=interp, grad=grad, conv=conv, diff=diff, coef_interp=coef_interp, coef_grad=coef_grad)


def _finite(v):
    if not np.isfinite(v):
        raise ValueError("nonfinite equ coefficient")


class Equation:
    """equ::Equation<T> (procedural_equation.cpp:15-36, 102-390)."""

    def __init__(self, unknown: Field, linear, options: dict | None = None):
        if unknown.location != CELL:
            raise ValueError("equ requires compatible cell fields")
        self.x, self.linear = unknown, linear
        self.opt = options or methods()
        self.m = unknown.m
        self.system = None
        self.reset()

    def reset(self):  # :124-126
        m, k = self.m, self.x.ncomp
        self.diag = np.zeros(m.nc)
        self.upper = np.zeros(m.nf)
        self.lower = np.zeros(m.nf)
        self.src = np.zeros(m.nc) if k == 1 else np.zeros((m.nc, 3))
        self.diffusion = []  # (snapshot dict, factor)
        self.boundary_flux = None

    def copy(self) -> "Equation":  # :127-131
        b = Equation(self.x, self.linear, self.opt)
        b.diag, b.upper, b.lower, b.src = (a.copy() for a in (self.diag, self.upper, self.lower, self.src))
        b.diffusion = list(self.diffusion)
        b.boundary_flux = self.boundary_flux
        return b

    def _accumulate(self, c: dict, f):  # :62-70
        _finite(f)
        self.diag += f * c["diag"]
        self.src = self.src + f * c["source"]
        self.upper += f * c["upper"]
        self.lower += f * c["lower"]

    def div(self, flux: Field, factor=1.0):  # :132-145
        _finite(factor)
        if flux.location != FACE:
            raise ValueError("equ::div requires a face flux")
        if self.boundary_flux is not None and self.boundary_flux is not flux:
            raise ValueError("one equation requires one boundary flux context")
        self.boundary_flux = flux
        self.x.set_boundary_flux(flux)
        o = self.opt
        c = self.m.assemble("convection", self.x.v, self.x.bct, self.x.bcv, phi=flux.v, conv=o["conv"],
                            interp=o["interp"], gmethod=o["grad"], scale=factor)
        self.diag += c["diag"]
        self.upper += c["upper"]
        self.lower += c["lower"]
        self.src = self.src + c["source"]

    def _face_coefficient(self, coefficient) -> np.ndarray:
        m, o = self.m, self.opt
        if not isinstance(coefficient, Field):
            _finite(coefficient)
            return np.full(m.nf, float(coefficient))
        if coefficient.location == FACE:
            return coefficient.v.copy()
        return m.interpolate(coefficient.v, coefficient.bct, coefficient.bcv, o["coef_interp"], o["coef_grad"])

    def laplacian(self, coefficient, multiplier):  # :146-193
        _finite(multiplier)
        m, o, x = self.m, self.opt, self.x
        kf = self._face_coefficient(coefficient)
        c = m.assemble("diffusion", x.v, x.bct, x.bcv, kf=kf, gmethod=o["grad"], dmethod=o["diff"], ctx=x.ctx())
        if x.ncomp == 1:
            if o["diff"] != ORTHOGONAL:  # diffusionCorrection (:77-87)
                full = m.diffusion_flux(x.v, x.bct, x.bcv, kf, o["grad"], o["diff"], x.ctx())
                orth = m.diffusion_flux(x.v, x.bct, x.bcv, kf, o["grad"], ORTHOGONAL, x.ctx())
                corr = full + (-1.0) * orth
            else:
                corr = np.zeros(m.nf)
            snapshot = dict(kf=kf, bct=x.bct.copy(), bcv=x.bcv.copy(), corr=corr,
                            ctx=None if x.flux is None else x.flux.copy())
            self.diffusion.append((snapshot, multiplier))
        self._accumulate(c, -multiplier)

    def reaction(self, coefficient, factor=1.0):  # :194-204
        V = self.m.geometry()["V"]
        if isinstance(coefficient, Field):
            _finite(factor)
            if not np.all(np.isfinite(coefficient.v)):
                raise ValueError("nonfinite equ coefficient")
            self.diag += factor * coefficient.v * V
        else:
            _finite(coefficient)
            self.diag += coefficient * V

    def source(self, value, factor=1.0):  # :205-213
        V = self.m.geometry()["V"]
        Vb = V if self.x.ncomp == 1 else V[:, None]
        if isinstance(value, Field):
            _finite(factor)
            self.src = self.src + factor * Vb * value.v
        else:
            self.src = self.src + Vb * np.asarray(value, np.float64)

    def ddt(self, capacity, previous: Field, dt, method=EULER, older: Field | None = None, previous_dt=0.0):
        """timeTerm (:214-240)."""
        if method == STEADY:
            return
        if not (dt > 0) or not np.isfinite(dt):
            raise ValueError("equ::ddt requires positive finite dt")
        a0, a1, a2 = 1 / dt, 1 / dt, 0.0
        if method == BDF2:
            if older is None:
                raise ValueError("BDF2 requires explicit older field; use Euler to start")
            olddt = dt if previous_dt == 0 else previous_dt
            if not (olddt > 0) or not np.isfinite(olddt):
                raise ValueError("invalid previous dt")
            r = dt / olddt
            a0, a1, a2 = (1 + 2 * r) / ((1 + r) * dt), (1 + r) / dt, -r * r / ((1 + r) * dt)
        c = capacity.v if isinstance(capacity, Field) else np.full(self.m.nc, float(capacity))
        if not np.all((c > 0) & np.isfinite(c)):
            raise ValueError("time capacity must be positive and finite")
        w = c * self.m.geometry()["V"]
        wb = w if self.x.ncomp == 1 else w[:, None]
        self.diag += w * a0
        self.src = self.src + wb * a1 * previous.v
        if a2 != 0:
            self.src = self.src + wb * a2 * older.v

    def scale(self, f):  # :259-266
        _finite(f)
        self.diag *= f
        self.upper *= f
        self.lower *= f
        self.src = f * self.src
        self.diffusion = [(s, factor * f) for s, factor in self.diffusion]

    def add(self, other: "Equation", f=1.0):  # :267-280
        if other.x is not self.x:
            raise ValueError("matrix addition requires same unknown binding")
        if other.boundary_flux is not None:
            if self.boundary_flux is not None and self.boundary_flux is not other.boundary_flux:
                raise ValueError("one equation requires one boundary flux context")
            self.boundary_flux = other.boundary_flux
        contributions = [(s, factor * f) for s, factor in other.diffusion]
        self._accumulate(dict(diag=other.diag, upper=other.upper, lower=other.lower, source=other.src), f)
        self.diffusion += contributions

    def relax(self, previous: Field, alpha):  # :317-324
        if not (0 < alpha <= 1):
            raise ValueError("relaxation must be in (0,1]")
        extra = self.diag * (1 / alpha - 1)
        self.diag = self.diag + extra
        self.src = self.src + (extra if self.x.ncomp == 1 else extra[:, None]) * previous.v

    def reference(self, cell, value):  # :325-334
        _finite(value)
        if self.x.ncomp != 1:
            raise ValueError("reference requires a scalar equation")
        if not 0 <= cell < self.m.nc:
            raise ValueError("reference global cell does not exist")
        if not self.diag[cell] > 0:
            raise ValueError("reference requires a positive diagonal")
        self.src[cell] += self.diag[cell] * value
        self.diag[cell] *= 2

    def reference_if_unanchored(self, value) -> bool:  # :335-350
        _finite(value)
        if self.x.ncomp != 1:
            raise ValueError("reference requires a scalar equation")
        own, nei = self.m.owner, self.m.neighbour
        inner = nei >= 0
        row = self.diag.copy()
        mag = np.abs(self.diag)
        np.add.at(row, own[inner], self.upper[inner])
        np.add.at(mag, own[inner], np.abs(self.upper[inner]))
        np.add.at(row, nei[inner], self.lower[inner])
        np.add.at(mag, nei[inner], np.abs(self.lower[inner]))
        constant = np.all(np.isfinite(row) & (np.abs(row) <= 64 * np.finfo(float).eps * mag))
        if constant:
            self.reference(0, value)
        return bool(constant)

    def diagonal(self):
        return self.diag.copy()

    def rhs(self):
        return self.src.copy()

    def apply(self, x: Field):  # :299-310
        own, nei = self.m.owner, self.m.neighbour
        inner = nei >= 0
        d = self.diag if self.x.ncomp == 1 else self.diag[:, None]
        up = self.upper[inner] if self.x.ncomp == 1 else self.upper[inner][:, None]
        lo = self.lower[inner] if self.x.ncomp == 1 else self.lower[inner][:, None]
        y = d * x.v
        np.add.at(y, own[inner], up * x.v[nei[inner]])
        np.add.at(y, nei[inner], lo * x.v[own[inner]])
        return y

    def residual(self, x: Field):  # :311-315
        return self.src - self.apply(x)

    def relative_residual(self, x: Field) -> float:  # :447-452
        ax = self.apply(x)
        r = self.src - ax
        n = [float(np.sqrt(np.sum(a * a))) for a in (r, ax, self.src)]
        return n[0] / max(n[1] + n[2], 1e-30)

    def face_flux(self, solution: Field):  # :352-370
        if self.x.ncomp != 1:
            raise ValueError("faceFlux requires a scalar equation")
        m = self.m
        out = np.zeros(m.nf)
        for s, factor in self.diffusion:
            inc = m.diffusion_flux(solution.v, s["bct"], s["bcv"], s["kf"], 0, ORTHOGONAL, s["ctx"])
            inc = inc + s["corr"]
            out += factor * inc
        return out

    def solve(self, cfg=None):  # :371-390
        if self.system is None:
            self.system = orc.OracleSystem(self.m)
        x, res = self.system.solve(self.diag, self.upper, self.lower, self.src, self.x.v,
                                   cfg if cfg is not None else self.linear)
        self.x.v = x.reshape(self.x.v.shape)
        return res


def norms(*fields: Field):  # field_math.cpp:30-62
    return [float(np.sqrt(np.sum(f.v * f.v))) for f in fields]


def relative_change(current: Field, previous: Field) -> float:  # fvm_execution.cpp:91-128
    d = current.v - previous.v
    return float(np.sqrt(np.sum(d * d)) / max(np.sqrt(np.sum(current.v * current.v)), 1e-30))


def relative_magnitude(value: Field, reference: Field) -> float:  # fvm_execution.cpp:130-146
    return float(np.sqrt(np.sum(value.v * value.v)) / max(np.sqrt(np.sum(reference.v * reference.v)), 1e-30))


def _outward(mesh, phi):
    own, nei = mesh.owner, mesh.neighbour
    inner = nei >= 0
    imb = np.zeros(mesh.nc)
    scale = np.zeros(mesh.nc)
    np.add.at(imb, own, phi)
    np.add.at(scale, own, np.abs(phi))
    np.add.at(imb, nei[inner], -phi[inner])
    np.add.at(scale, nei[inner], np.abs(phi[inner]))
    return imb, scale


def flux_balance(flux: Field) -> dict:  # fvm_execution.cpp:148-175
    imb, scale = _outward(flux.m, flux.v)
    l1 = float(np.sum(np.abs(imb)))
    return dict(l1=l1, l2=float(np.sqrt(np.sum(imb * imb))), maximum=float(np.max(np.abs(imb))),
                relative=l1 / max(float(np.sum(scale)), 1e-30))


def courant(flux: Field, dt: float) -> dict:  # fvm_execution.cpp:177-199
    if not (dt > 0) or not np.isfinite(dt):
        raise ValueError("Courant time step must be finite and positive")
    _, total = _outward(flux.m, flux.v)
    V = flux.m.geometry()["V"]
    co = 0.5 * dt * total / V
    return dict(maximum=float(np.max(co)), mean=float(np.sum(co * V) / np.sum(V)))


def face_add(increment: Field, target: Field, factor=1.0, region=0):  # field_math.cpp:65-82
    if region not in (0, 1):
        raise ValueError("invalid face region")
    mask = np.ones(target.m.nf, bool) if region == 0 else target.m.neighbour >= 0
    target.v[mask] += factor * increment.v[mask]
