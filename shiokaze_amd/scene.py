"""Synthetic scenes (host numpy arrays) from the library's scene generator.

Arrays are numpy views in (k, j, i) order so that the flat index is the C-ABI's
x-fastest index: fluid (nz, ny, nx), solid (nz+1, ny+1, nx+1), u (nz, ny, nx+1),
v (nz, ny+1, nx), w (nz+1, ny, nx); us/vs/ws like u/v/w.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

POOL_SPHERE = 1   # pool y < 0.3 + sphere r = 0.15 at (0.5, 0.7, 0.5), walls only
IRREGULAR = 3     # 20 seeded solids, pool y < 0.4 + 8 seeded liquid spheres


def shapes(nx: int, ny: int, nz: int) -> dict:
    return {
        "fluid": (nz, ny, nx),
        "solid": (nz + 1, ny + 1, nx + 1),
        "u": (nz, ny, nx + 1),
        "v": (nz, ny + 1, nx),
        "w": (nz + 1, ny, nx),
        "us": (nz, ny, nx + 1),
        "vs": (nz, ny + 1, nx),
        "ws": (nz + 1, ny, nx),
    }


def generate(kind: int, n, seed: int = 0, dx: float | None = None) -> dict:
    """Scene of `kind` on an n^3 (or (nx, ny, nz)) grid with cell size dx (1/max(n))."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else tuple(n)
    if dx is None:
        dx = 1.0 / max(nx, ny, nz)
    lib = _lib.load()
    arrs = {k: np.empty(s, dtype=np.float32) for k, s in shapes(nx, ny, nz).items()}
    dropped = ctypes.c_int64(0)
    order = ["fluid", "solid", "u", "v", "w", "us", "vs", "ws"]
    rc = lib.shk_scene_generate(kind, nx, ny, nz, float(dx), int(seed),
                                *[arrs[k].ctypes.data_as(ctypes.c_void_p) for k in order],
                                ctypes.byref(dropped))
    _lib.check(rc)
    arrs["dropped"] = int(dropped.value)
    arrs["dx"] = float(dx)
    arrs["shape"] = (nx, ny, nz)
    return arrs
