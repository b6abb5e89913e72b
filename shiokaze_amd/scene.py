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


def generate2d(kind: int, n, seed: int = 0, dx: float | None = None) -> dict:
    """2D scene on an n^2 (or (nx, ny)) grid, arrays in (j, i) order: fluid (ny, nx),
    solid (ny+1, nx+1) nodal, u/us (ny, nx+1), v/vs (ny+1, nx).

    POOL_SPHERE: pool y < 0.3 and a disc r = 0.15 at (0.5, 0.7), walls only.
    IRREGULAR: 10 seeded solids (boxes and discs, r and half-extents in [0.03, 0.12]),
    pool y < 0.4 and 4 seeded discs (r in [0.05, 0.15]), the liquid clipped by the
    solid as the 3D scenes are (phi_f = max(phi_f, -(phi_s + dx))), and liquid
    components without an air cell beside them removed (they would make the system
    singular).  Velocities are uniform in [-1, 1] from a seeded generator."""
    nx, ny = (n, n) if np.isscalar(n) else tuple(n)
    if dx is None:
        dx = 1.0 / max(nx, ny)
    rng = np.random.default_rng(seed)
    cx = (np.arange(nx) + 0.5) * dx
    cy = (np.arange(ny) + 0.5) * dx
    Y, X = np.meshgrid(cy, cx, indexing="ij")
    nX, nY = np.meshgrid(np.arange(nx + 1) * dx, np.arange(ny + 1) * dx, indexing="xy")
    if kind == POOL_SPHERE:
        fluid = np.minimum(Y - 0.3, np.hypot(X - 0.5, Y - 0.7) - 0.15)
        solid = np.ones((ny + 1, nx + 1))
    elif kind == IRREGULAR:
        def solids(px, py):
            phi = np.full(px.shape, np.inf)
            for s in range(10):
                c = rng_s[s]
                if c[0] < 0.5:
                    d = np.hypot(px - c[1], py - c[2]) - c[3]
                else:
                    d = np.maximum(np.abs(px - c[1]) - c[3], np.abs(py - c[2]) - c[4])
                phi = np.minimum(phi, d)
            return phi
        rng_s = [(rng.uniform(), rng.uniform(), rng.uniform(), rng.uniform(0.03, 0.12), rng.uniform(0.03, 0.12))
                 for _ in range(10)]
        solid = solids(nX, nY)
        fluid = Y - 0.4
        for _ in range(4):
            ccx, ccy, r = rng.uniform(), rng.uniform(), rng.uniform(0.05, 0.15)
            fluid = np.minimum(fluid, np.hypot(X - ccx, Y - ccy) - r)
        fluid = np.maximum(fluid, -(solids(X, Y) + dx))
        from scipy import ndimage
        liquid = fluid < 0
        air = (fluid >= 0) & (solids(X, Y) > 0)
        lab, nlab = ndimage.label(liquid)
        near_air = ndimage.binary_dilation(air) & liquid
        keep = np.zeros(nlab + 1, bool)
        keep[np.unique(lab[near_air])] = True
        keep[0] = False
        fluid = np.where(liquid & ~keep[lab], np.abs(fluid) + dx, fluid)
    else:
        raise ValueError(f"unknown 2D scene kind {kind}")
    out = {"fluid": fluid.astype(np.float32), "solid": solid.astype(np.float32),
           "u": rng.uniform(-1, 1, (ny, nx + 1)).astype(np.float32),
           "v": rng.uniform(-1, 1, (ny + 1, nx)).astype(np.float32),
           "us": np.zeros((ny, nx + 1), np.float32), "vs": np.zeros((ny + 1, nx), np.float32),
           "dx": float(dx), "shape2": (nx, ny)}
    return out
