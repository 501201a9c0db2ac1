"""Test mesh fixtures (synthetic inputs; produced by the library's generator)."""
from __future__ import annotations

import numpy as np

from babelsim_amd import MeshData


def lattice(n, f):
    nx, ny, nz = n
    pts = []
    for k in range(nz + 1):
        for j in range(ny + 1):
            for i in range(nx + 1):
                pts.append(f(i, j, k))
    return np.asarray(pts, np.float64)


def box(n=(3, 3, 3), lo=(0, 0, 0), hi=(3, 3, 3)) -> dict:
    return MeshData.hex(n, lo, hi).arrays()


def skewed() -> dict:
    # Affine skew lattice (the skewed fixture of tests/operators_test.cpp:16-32 in the reference).
    pts = lattice((5, 5, 5), lambda i, j, k: (i + 0.25 * j + 0.10 * k, j + 0.15 * k, k))
    return MeshData.hex((5, 5, 5), vertices=pts).arrays()


def warped() -> dict:
    # Smooth warp (the warped fixture of tests/operators_test.cpp:34-57 in the reference).
    n = 7
    pi = np.pi

    def f(i, j, k):
        xi, eta, zeta = i / n, j / n, k / n
        env = np.sin(pi * xi) * np.sin(pi * eta)
        return (xi + 0.12 * env + 0.08 * zeta,
                eta + 0.05 * np.sin(2 * pi * xi) * np.sin(pi * eta) + 0.04 * zeta, zeta)
    return MeshData.hex((n, n, n), vertices=lattice((n, n, n), f)).arrays()


def jittered(n=(6, 5, 4), jitter=0.2, split=0.3, seed=7) -> dict:
    return MeshData.hex(n, jitter=jitter, split=split, seed=seed).arrays()


def split_interface() -> dict:
    # Two cells whose shared planar interface is split into two triangles that
    # couple the same owner/neighbour pair (fixture of tests/test_util.h:188-214).
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1],
                  [2, 0, 0], [2, 1, 0], [2, 0, 1], [2, 1, 1]], np.float64)
    faces = [([0, 4, 7, 3], 0, -1, 0), ([1, 2, 6], 0, 1, -1), ([1, 6, 5], 0, 1, -1), ([0, 1, 5, 4], 0, -1, 0),
             ([3, 7, 6, 2], 0, -1, 0), ([0, 3, 2, 1], 0, -1, 0), ([4, 5, 6, 7], 0, -1, 0),
             ([8, 10, 11, 9], 1, -1, 0), ([1, 8, 10, 5], 1, -1, 0), ([2, 6, 11, 9], 1, -1, 0),
             ([1, 2, 9, 8], 1, -1, 0), ([5, 10, 11, 6], 1, -1, 0)]
    fpo = [0]
    fpi = []
    for ring, *_ in faces:
        fpi += ring
        fpo.append(len(fpi))
    return dict(vertices=v.reshape(-1), fpo=np.array(fpo, np.int32), fpi=np.array(fpi, np.int32),
                owner=np.array([f[1] for f in faces], np.int32), neighbour=np.array([f[2] for f in faces], np.int32),
                patch=np.array([f[3] for f in faces], np.int32), npatch=1, ncells=2)


def cavity2d(n=8) -> dict:
    return MeshData.hex((n, n, 1), (0, 0, 0), (1, 1, 1.0 / n)).arrays()


ALL = {"box": box, "skewed": skewed, "warped": warped, "jittered": jittered, "split_interface": split_interface,
       "cavity2d": cavity2d}
