"""Centreline comparison of a 2-D lid-driven cavity with Ghia et al. (1982),
following the reference's cases/cavity/validation/validate_cavity.py:104-131:
u on x = 0.5 and v on y = 0.5, linearly interpolated between cell centres,
wall values appended, errors at the table's coordinates, RMS over the table."""
from __future__ import annotations

from pathlib import Path

import numpy as np

_TABLE = Path(__file__).resolve().parent / "golden" / "ghia1982_re100.csv"


def table():
    d = np.loadtxt(_TABLE, delimiter=",", comments="#")
    return d[:, 0], d[:, 1], d[:, 2], d[:, 3]


def centreline_errors(centres: np.ndarray, U: np.ndarray, n: int):
    """centres (N,3) and U (N,3) on an n x n x 1 unit-square lattice; returns (u_L2, v_L2)."""
    h = 1.0 / n
    ix = np.rint(centres[:, 0] / h - 0.5).astype(int)
    iy = np.rint(centres[:, 1] / h - 0.5).astype(int)
    ug = np.zeros((n, n))
    vg = np.zeros((n, n))
    ug[iy, ix] = U[:, 0]
    vg[iy, ix] = U[:, 1]
    xs = (np.arange(n) + 0.5) * h
    u_line = np.array([np.interp(0.5, xs, ug[r]) for r in range(n)])
    v_line = np.array([np.interp(0.5, xs, vg[:, c]) for c in range(n)])
    gy, gu, gx, gv = table()
    eu = np.interp(gy, np.r_[0.0, xs, 1.0], np.r_[0.0, u_line, 1.0]) - gu
    ev = np.interp(gx, np.r_[0.0, xs, 1.0], np.r_[0.0, v_line, 0.0]) - gv
    return float(np.sqrt(np.mean(eu * eu))), float(np.sqrt(np.mean(ev * ev)))
