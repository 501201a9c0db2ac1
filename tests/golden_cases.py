"""Operator cases evaluated on small meshes for the golden fixtures: one list
of (name, call) pairs that runs unchanged on the compiled reference, on the
oracle and on the device context (all three expose the same methods)."""
from __future__ import annotations

import numpy as np

import meshes

MESHES = ("box", "skewed", "jittered", "split_interface", "cavity2d")


def scalar_bcs(npatch):
    types = [0, 1, 2, 4, 3, 0][:npatch] if npatch > 1 else [0]
    vals = [1.5, 0.7, 0.0, 0.0, 2.0, -0.5][:npatch]
    return np.array(types, np.int32), np.array(vals)


def vector_bcs(npatch):
    types = [0, 1, 2, 4, 3, 0][:npatch] if npatch > 1 else [0]
    vals = np.array([[1, 0.5, -0.2], [0.3, -0.1, 0.2], [0, 0, 0], [0, 0, 0], [0.4, 0.2, 0.1], [-0.5, 0.25, 1.0]])
    return np.array(types, np.int32), vals[:npatch]


def inputs(arrays: dict, centres: np.ndarray, seed=3) -> dict:
    """Deterministic smooth-plus-noise fields from the cell centres."""
    rng = np.random.default_rng(seed)
    C = centres
    nc, nf = len(C), len(arrays["owner"])
    s = np.sin(C[:, 0]) + 0.3 * C[:, 1] ** 2 - 0.2 * C[:, 2] + 0.01 * rng.standard_normal(nc)
    U = np.stack([np.cos(C[:, 1]) + C[:, 2], 0.5 * C[:, 0] * C[:, 2], -0.3 * C[:, 1] + 0.1 * C[:, 0]], 1)
    U += 0.01 * rng.standard_normal(U.shape)
    phi = rng.standard_normal(nf)
    return dict(s=s, U=U, phi=phi, kf=0.5 + np.abs(phi))


def evaluate(m, npatch: int, x: dict) -> dict:
    """All operator outputs for the inputs x on mesh object m."""
    s, U, phi, kf = x["s"], x["U"], x["phi"], x["kf"]
    st, sv = scalar_bcs(npatch)
    vt, vv = vector_bcs(npatch)
    out = {f"geo_{k}": v for k, v in m.geometry().items()}
    for gm in (0, 1):
        out[f"grad_s_g{gm}"] = m.gradient(s, st, sv, gm, ctx=phi)
        out[f"grad_U_g{gm}"] = m.gradient(U, vt, vv, gm, ctx=phi)
    for im, gm in ((0, 0), (1, 0), (1, 1)):
        out[f"interp_s_i{im}g{gm}"] = m.interpolate(s, st, sv, im, gm, ctx=phi)
        out[f"interp_U_i{im}g{gm}"] = m.interpolate(U, vt, vv, im, gm, ctx=phi)
        out[f"flux_i{im}g{gm}"] = m.flux(U, vt, vv, im, gm, ctx=phi)
        out[f"flux_nctx_i{im}g{gm}"] = m.flux(U, vt, vv, im, gm, ctx=None)
    # inletOutlet needs a flux context that these operators do not bind.
    zt = np.where(st == 3, 2, st).astype(np.int32)
    for dm in (0, 1, 2):
        for gm in (0, 1):
            out[f"sngrad_d{dm}g{gm}"] = m.normal_gradient(s, zt, sv, gm, dm)
            out[f"lap_d{dm}g{gm}"] = m.explicit_laplacian(s, zt, sv, kf, 0.0, gm, dm)
            out[f"lapk_d{dm}g{gm}"] = m.explicit_laplacian(s, zt, sv, None, 0.3, gm, dm)
            for key, val in m.assemble("diffusion", s, st, sv, kf=kf, gmethod=gm, dmethod=dm, ctx=phi).items():
                out[f"diff_s_d{dm}g{gm}_{key}"] = val
            for key, val in m.assemble("diffusion", U, vt, vv, kf=kf, gmethod=gm, dmethod=dm, ctx=phi).items():
                out[f"diff_U_d{dm}g{gm}_{key}"] = val
    out["div_face"] = m.divergence_face(phi)
    for conv, interp in ((0, 0), (1, 0), (2, 0), (2, 1)):
        for gm in (0, 1):
            tag = f"c{conv}i{interp}g{gm}"
            out[f"conv_s_{tag}"] = m.explicit_convection(s, st, sv, phi, conv, interp, gm)
            out[f"conv_U_{tag}"] = m.explicit_convection(U, vt, vv, phi, conv, interp, gm)
            for key, val in m.assemble("convection", s, st, sv, phi=phi, conv=conv, interp=interp, gmethod=gm,
                                       scale=1.3).items():
                out[f"aconv_s_{tag}_{key}"] = val
            for key, val in m.assemble("convection", U, vt, vv, phi=phi, conv=conv, interp=interp, gmethod=gm,
                                       scale=0.8).items():
                out[f"aconv_U_{tag}_{key}"] = val
    return out


def mesh_arrays(name: str) -> dict:
    return getattr(meshes, name)()
