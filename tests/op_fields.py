"""Seeded test fields and boundary-condition sets shared by the operator
parity tests (oracle against the reference build, device against the oracle).
Plain data: importing this module loads no native library."""
import numpy as np


def scalar_bcs(npatch):
    types = [0, 1, 2, 4, 3, 0][:npatch] if npatch > 1 else [0]
    vals = [1.5, 0.7, 0.0, 0.0, 2.0, -0.5][:npatch]
    return np.array(types, np.int32), np.array(vals)


def vector_bcs(npatch):
    types = [0, 1, 2, 4, 3, 0][:npatch] if npatch > 1 else [0]
    vals = np.array([[1, 0.5, -0.2], [0.3, -0.1, 0.2], [0, 0, 0], [0, 0, 0], [0.4, 0.2, 0.1], [-0.5, 0.25, 1.0]])
    return np.array(types, np.int32), vals[:npatch]


def fields(om, seed=3):
    """A smooth scalar and vector field with seeded noise on the cells of an
    oracle mesh, and a seeded face flux."""
    g = om.geometry()
    rng = np.random.default_rng(seed)
    C = g["C"]
    s = np.sin(C[:, 0]) + 0.3 * C[:, 1] ** 2 - 0.2 * C[:, 2] + 0.01 * rng.standard_normal(om.nc)
    U = np.stack([np.cos(C[:, 1]) + C[:, 2], 0.5 * C[:, 0] * C[:, 2], -0.3 * C[:, 1] + 0.1 * C[:, 0]], 1)
    U += 0.01 * rng.standard_normal(U.shape)
    phi = rng.standard_normal(om.nf)
    return s, U, phi
