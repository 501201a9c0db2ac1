"""The serving stream: steady laminar lid-driven cavities and straight channels
on jittered hex lattices, sizes log-uniform between 1e4 and 1e6 cells,
Re uniform in [50, 1000], Poisson arrivals at a fixed rate.

Sizes are rounded to a geometric ladder of lattice sides (ratio ~1.2), so
requests of one size bucket share a mesh: the server registers each distinct
mesh once and keeps its contexts resident between requests.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from babelsim_amd import MeshData, simple_config, solver_config

FV, FG, ZG, IO, SYM = range(5)

# Cube sides of the cavity ladder: 22^3 = 10,648 ... 100^3 = 1,000,000 cells.
CAVITY_SIDES = (22, 26, 31, 37, 44, 53, 63, 75, 89, 100)
# Channel lattices 4m x m x m on [0, 4] x [0, 1]^2: 4 * 14^3 = 10,976 ...
# 4 * 63^3 = 1,000,188 cells.
CHANNEL_SIDES = (14, 17, 20, 24, 29, 34, 41, 49, 58, 63)


@dataclass
class ServeRequestSpec:
    index: int
    kind: str        # "cavity" or "channel"
    side: int        # lattice side (cube side, or channel cross-section)
    cells: int
    re: float
    arrival: float   # seconds after the stream starts


def cavity_mesh(n: int, seed: int = 20261205) -> MeshData:
    return MeshData.hex((n, n, n), jitter=0.2, split=0.0, seed=seed + n)


def channel_mesh(m: int, seed: int = 20261206) -> MeshData:
    return MeshData.hex((4 * m, m, m), (0, 0, 0), (4, 1, 1), jitter=0.2, split=0.0, seed=seed + m)


def cavity_bcs():
    """Walls at rest, lid (+y) moving with (1, 0, 0)."""
    ut = np.full(6, FV, np.int32)
    uv = np.zeros((6, 3))
    uv[3] = (1.0, 0.0, 0.0)
    return (ut, uv), (np.full(6, ZG, np.int32), np.zeros(6))


def channel_bcs():
    """-x inlet U = (1, 0, 0); +x outlet inletOutlet with p = 0; four walls."""
    ut = np.array([FV, IO, FV, FV, FV, FV], np.int32)
    uv = np.zeros((6, 3))
    uv[0] = (1.0, 0.0, 0.0)
    pt = np.array([ZG, FV, ZG, ZG, ZG, ZG], np.int32)
    return (ut, uv), (pt, np.zeros(6))


def serve_config(re: float, precision: int = 0, tol: float = 1e-5, rtol: float = 1e-6, warm_start: int = 0):
    """Laminar SIMPLE on a jittered mesh: least-squares gradients, corrected
    diffusion, linearUpwind, alpha_U 0.7 / alpha_p 0.3; BiCGSTAB+Jacobi
    momentum and CG+AMG pressure (strength threshold 0.65, over-correction
    2.0, Jacobi weight 0.9) at rtol (1e-6); the four outer tolerances
    `tol`.  precision 1 runs the AMG and the momentum Krylov in FP32 inside
    FP64 refinement.  warm_start starts both linear solves from the current
    unknowns."""
    return simple_config(rho=1.0, mu=1.0 / re, m_grad=1, m_diff=1, m_coef_grad=1, p_grad=1, p_diff=1,
                         p_coef_grad=1, alpha_u=0.7, alpha_p=0.3, non_orth_corr=1,
                         tol_mass=tol, tol_u=tol, tol_mom=tol, tol_p=tol,
                         mom=solver_config(solver=1, pc=1, rtol=rtol, atol=1e-14, precision=precision,
                                           warm_start=warm_start),
                         pre=solver_config(solver=0, pc=2, rtol=rtol, atol=1e-14, precision=precision,
                                           amg_strength=0.65, amg_correction=2.0, amg_omega=0.9,
                                           warm_start=warm_start))


def stream(count: int = 256, rate: float = 4.0, seed: int = 20261207, cells_min: float = 1e4,
           cells_max: float = 1e6) -> list[ServeRequestSpec]:
    """`count` requests: log-uniform size in [cells_min, cells_max] rounded to
    the nearest ladder rung (in log size), cavity or channel with equal odds,
    Re ~ U[50, 1000], exponential inter-arrival times of mean 1/rate."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for i in range(count):
        target = math.exp(rng.uniform(math.log(cells_min), math.log(cells_max)))
        kind = "cavity" if rng.random() < 0.5 else "channel"
        sides = CAVITY_SIDES if kind == "cavity" else CHANNEL_SIDES
        cells_of = (lambda n: n ** 3) if kind == "cavity" else (lambda m: 4 * m ** 3)
        side = min(sides, key=lambda s: abs(math.log(cells_of(s)) - math.log(target)))
        re = float(rng.uniform(50.0, 1000.0))
        t += float(rng.exponential(1.0 / rate))
        out.append(ServeRequestSpec(i, kind, side, cells_of(side), re, t))
    return out


def mesh_for(spec: ServeRequestSpec) -> MeshData:
    return cavity_mesh(spec.side) if spec.kind == "cavity" else channel_mesh(spec.side)


def bcs_for(spec: ServeRequestSpec):
    return cavity_bcs() if spec.kind == "cavity" else channel_bcs()
