"""Plane channel for steady RANS with the Wilcox k-omega model: a hex lattice
graded geometrically towards both walls so that the first cell sits at
y+ of about 1, bulk inflow U = 1, Re_bulk = 40,000 on the full height
2h, inlet k and omega from 5 % intensity and a 0.1 h length scale.

Patches of the hex lattice: -x inlet, +x outlet, -y and +y walls, -z and +z
symmetry planes."""
from __future__ import annotations

import math

import numpy as np

from babelsim_amd import MeshData, komega_config, simple_config, solver_config

FV, FG, ZG, IO, SYM = range(5)

H = 1.0  # half height
RE_BULK = 40_000.0
U_BULK = 1.0
NU = U_BULK * 2 * H / RE_BULK
BETA_STAR, BETA1 = 0.09, 0.075


def friction_velocity():
    """Dean's correlation Re_tau = 0.09 Re_bulk^0.88 (Re_bulk on 2h, Re_tau on h)."""
    return 0.09 * RE_BULK ** 0.88 * NU / H


def first_cell_height(y_plus=1.0):
    return y_plus * NU / friction_velocity()


def grading_ratio(n_half, first, half=H):
    """r with first (r^n - 1) / (r - 1) = half, by bisection."""
    lo, hi = 1.0 + 1e-12, 2.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        total = first * (r ** n_half - 1) / (r - 1)
        lo, hi = (r, hi) if total < half else (lo, r)
    return 0.5 * (lo + hi)


def wall_normal_coordinates(ny, y_plus=1.0):
    """ny + 1 vertex heights on [0, 2h], symmetric geometric grading."""
    if ny % 2:
        raise ValueError("the wall-normal cell count must be even")
    half = ny // 2
    first = first_cell_height(y_plus)
    r = grading_ratio(half, first)
    widths = first * r ** np.arange(half)
    widths *= H / widths.sum()
    lower = np.concatenate([[0.0], np.cumsum(widths)])
    return np.concatenate([lower, 2 * H - lower[-2::-1]])


def graded_channel_mesh(nx=512, ny=96, nz=100, length=16.0, width=3.2, y_plus=1.0) -> MeshData:
    xs = np.linspace(0.0, length, nx + 1)
    ys = wall_normal_coordinates(ny, y_plus)
    zs = np.linspace(0.0, width, nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")  # x fastest, as the lattice numbers vertices
    verts = np.stack([X, Y, Z], axis=-1).reshape(-1)
    return MeshData.hex((nx, ny, nz), (0, 0, 0), (length, 2 * H, width), vertices=verts)


def inlet_turbulence(intensity=0.05, length_scale=0.1 * H):
    k = 1.5 * (intensity * U_BULK) ** 2
    omega = math.sqrt(k) / (BETA_STAR ** 0.25 * length_scale)
    return k, omega


def wall_omega(y_plus=1.0):
    """Menter's wall value 10 * 6 nu / (beta1 y1^2) for the first cell height y1."""
    y1 = first_cell_height(y_plus)
    return 60.0 * NU / (BETA1 * y1 * y1)


def channel_bcs(y_plus=1.0):
    k_in, w_in = inlet_turbulence()
    ub = (np.array([FV, ZG, FV, FV, SYM, SYM], np.int32), np.array([[U_BULK, 0, 0]] + [[0, 0, 0]] * 5))
    pb = (np.array([ZG, FV, ZG, ZG, SYM, SYM], np.int32), np.zeros(6))
    kb = (np.array([FV, ZG, FV, FV, ZG, ZG], np.int32), np.array([k_in, 0, 0, 0, 0, 0]))
    ww = wall_omega(y_plus)
    wb = (np.array([FV, ZG, FV, FV, ZG, ZG], np.int32), np.array([w_in, 0, ww, ww, 0, 0]))
    return ub, pb, kb, wb


def channel_config(precision=1, rebuild_interval=8, warm_start=0, rtol=1e-6):
    """SIMPLE with linearUpwind momentum, relaxation 0.7 / 0.3, FP32 BiCGSTAB
    for momentum, k and omega inside FP64 refinement, the pressure correction
    by FP32-AMG CG inside FP64 refinement; warm_start sets every equation's
    warmStart (start from the current unknowns)."""
    common = dict(rtol=rtol, atol=1e-14, precision=precision, rebuild_interval=rebuild_interval, warm_start=warm_start)
    mom = solver_config(solver=1, pc=1, **common)
    # Strength threshold 0.85: on the wall-stretched cells only the wall-normal
    # couplings are strong, so aggregates follow the grading.
    pre = solver_config(solver=0, pc=2, amg_strength=0.85, **common)
    turb = solver_config(solver=1, pc=1, **common)
    cfg = simple_config(rho=1.0, mu=NU, m_conv=1, alpha_u=0.7, alpha_p=0.3, mom=mom, pre=pre)
    # Upwind transport for k and omega, as in the reference's own k-omega cases
    # (cases/naca0012/numerics/methods.bs:15,20 and
    # cases/cylinder_re3900_urans/numerics/methods.bs:12,16): linearUpwind's
    # explicit correction undershoots the steep near-wall omega, the viscosity
    # ratio then hits its bound and the segregated iteration stalls.
    kw = komega_config(k_solver=turb, w_solver=turb, k_conv=0, w_conv=0)
    return cfg, kw


def initial_fields(nc):
    """Bulk velocity everywhere and the inlet turbulence."""
    k_in, w_in = inlet_turbulence()
    U0 = np.zeros((nc, 3))
    U0[:, 0] = U_BULK
    return U0, np.full(nc, k_in), np.full(nc, w_in)
