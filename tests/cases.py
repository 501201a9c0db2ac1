"""Case definitions shared by tests, smoke() and the benchmark: lid-driven
cavities on the six-patch hex lattice (-x, +x, -y, +y, -z, +z)."""
from __future__ import annotations

import numpy as np

from babelsim_amd import MeshData, simple_config, solver_config

FIXED_VALUE, FIXED_GRADIENT, ZERO_GRADIENT, INLET_OUTLET, SYMMETRY = range(5)


def cavity_bcs(two_d: bool):
    """Walls at rest, lid (+y) moving with (1,0,0); front/back symmetry in 2-D
    (BabelSim's 2-D idiom, cases/cavity/fields/initial)."""
    ut = [FIXED_VALUE] * 6
    uv = np.zeros((6, 3))
    uv[3] = (1.0, 0.0, 0.0)
    pt = [ZERO_GRADIENT] * 6
    if two_d:
        ut[4] = ut[5] = SYMMETRY
        pt[4] = pt[5] = SYMMETRY
    return (np.array(ut, np.int32), uv), (np.array(pt, np.int32), np.zeros(6))


def cavity2d_mesh(n: int) -> MeshData:
    return MeshData.hex((n, n, 1), (0, 0, 0), (1, 1, 1.0 / n))


def cavity3d_mesh(n: int, jitter=0.2, split=0.1, seed=20261203) -> MeshData:
    return MeshData.hex((n, n, n), jitter=jitter, split=split, seed=seed)


def c1_config(**kw):
    """Config 1: Re=100 cavity, linearUpwind, alpha_U 0.7, alpha_p 0.3, all
    SIMPLE tolerances 1e-6; p' by CG+AMG, U by BiCGSTAB."""
    base = dict(rho=1.0, mu=0.01, m_interp=0, m_grad=0, m_conv=1, m_diff=0, p_interp=0, p_grad=0, p_diff=0,
                alpha_u=0.7, alpha_p=0.3, tol_mass=1e-6, tol_u=1e-6, tol_mom=1e-6, tol_p=1e-6,
                mom=solver_config(solver=1, pc=1, rtol=1e-8, atol=1e-14),
                pre=solver_config(solver=0, pc=2, rtol=1e-8, atol=1e-14))
    base.update(kw)
    return simple_config(**base)


def c2_config(**kw):
    """Config 2: Re=400 polyhedral cavity, least-squares gradients and corrected
    diffusion on the non-orthogonal mesh."""
    base = dict(rho=1.0, mu=1.0 / 400.0, m_interp=0, m_grad=1, m_conv=1, m_diff=1, p_interp=0, p_grad=1, p_diff=1,
                m_coef_grad=1, p_coef_grad=1, alpha_u=0.7, alpha_p=0.3, non_orth_corr=1,
                tol_mass=1e-6, tol_u=1e-6, tol_mom=1e-6, tol_p=1e-6,
                mom=solver_config(solver=1, pc=1, rtol=1e-6, atol=1e-14),
                pre=solver_config(solver=0, pc=2, rtol=1e-6, atol=1e-14))
    base.update(kw)
    return simple_config(**base)
