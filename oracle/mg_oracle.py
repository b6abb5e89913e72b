"""FP64 numpy restatement of the aggregation-multigrid preconditioner (Preconditioner=mg).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
multigrid kernels; never part of the product path.

The reference's "amg" LinSolver (src/linsolver/amg.cpp:43-166) runs smoothed
aggregation from the vendored amgcl on the CPU.  The library's counterpart is its
own design: geometric 2x2x2 aggregates of the brick grid, Galerkin coarse operators
of piecewise-constant prolongation, one damped-Jacobi pre- and post-smoothing step
and a dense solve on the coarsest level.  This module restates exactly that, on
dense (nz, ny, nx) grids of the Jacobi-scaled operator: diagonal d (0 off the
unknowns) and low-face couplings cx, cy, cz stored at the high cell of each face,
A x = d x - sum_faces c x_neighbour.  tests/test_oracle_mg.py pins it against
explicit matrix products (P0' A P0), symmetry and positivity of M, and PCG
convergence; tests/test_gpu_mg.py compares the GPU V-cycle with it.
"""
from __future__ import annotations

import numpy as np

OMEGA = 0.9   # damped-Jacobi weight
ALPHA = 1.5   # coarse-correction factor
COARSE_AXIS = 4  # coarsening stops when every axis has at most this many cells


def apply(level, x):
    d, cx, cy, cz = level
    y = d * x
    y[:, :, 1:] -= cx[:, :, 1:] * x[:, :, :-1]
    y[:, :, :-1] -= cx[:, :, 1:] * x[:, :, 1:]
    y[:, 1:, :] -= cy[:, 1:, :] * x[:, :-1, :]
    y[:, :-1, :] -= cy[:, 1:, :] * x[:, 1:, :]
    y[1:, :, :] -= cz[1:, :, :] * x[:-1, :, :]
    y[:-1, :, :] -= cz[1:, :, :] * x[1:, :, :]
    return y


def _pad2(a):
    nz, ny, nx = a.shape
    return np.pad(a, ((0, nz % 2), (0, ny % 2), (0, nx % 2)))


def _blocks(a):
    a = _pad2(a)
    nz, ny, nx = a.shape
    return a.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2)


def coarsen(level):
    """Galerkin P0' A P0 of 2x2x2 aggregation as a 7-point operator; aggregates whose
    diagonal is not above 1e-6 of their children's sum are dropped (d = 0)."""
    d, cx, cy, cz = (_blocks(a) for a in level)
    dsum = d.sum(axis=(1, 3, 5))
    inner = cx[:, :, :, :, :, 1].sum(axis=(1, 3)) + cy[:, :, :, 1, :, :].sum(axis=(1, 4)) \
        + cz[:, 1, :, :, :, :].sum(axis=(2, 4))
    dc = dsum - 2.0 * inner
    dc = np.where(dc > 1e-6 * dsum, dc, 0.0)
    return (dc, cx[:, :, :, :, :, 0].sum(axis=(1, 3)), cy[:, :, :, 0, :, :].sum(axis=(1, 4)),
            cz[:, 0, :, :, :, :].sum(axis=(2, 4)))


def hierarchy(level0):
    """Levels down to (and including) the first with at most COARSE_AXIS cells per axis."""
    levels = [tuple(np.asarray(a, dtype=np.float64) for a in level0)]
    while len(levels) == 1 or max(levels[-1][0].shape) > COARSE_AXIS:
        levels.append(coarsen(levels[-1]))
    return levels


def restrict(t):
    return _blocks(t).sum(axis=(1, 3, 5))


def prolong(xc, shape):
    return xc.repeat(2, 0).repeat(2, 1).repeat(2, 2)[:shape[0], :shape[1], :shape[2]]


def coarse_solve(level, b):
    d = level[0]
    act = np.flatnonzero(d.ravel() > 0)
    x = np.zeros(d.size)
    if act.size:
        n = d.size
        cols = []
        for j in act:
            e = np.zeros(n)
            e[j] = 1.0
            cols.append(apply(level, e.reshape(d.shape)).ravel()[act])
        M = np.array(cols).T
        x[act] = np.linalg.solve(M, b.ravel()[act])
    return x.reshape(d.shape)


def vcycle(levels, b, l=0):
    """z = M b on level l."""
    L = levels[l]
    d = L[0]
    if l == len(levels) - 1:
        return coarse_solve(L, b * (d > 0))
    inv = np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 0.0)
    x = OMEGA * inv * b
    t = np.where(d > 0, b - apply(L, x), 0.0)
    xc = vcycle(levels, restrict(t), l + 1)
    x = x + ALPHA * prolong(xc, d.shape) * (d > 0)
    return np.where(d > 0, x + OMEGA * inv * (b - apply(L, x)), 0.0)


def pcg(level0, b, levels=None, tol=1e-6, max_iterations=2000):
    """CG preconditioned by the V-cycle; max-norm stop relative to max|b|."""
    levels = levels or hierarchy(level0)
    x = np.zeros_like(b)
    r = b.copy()
    bmax = np.abs(b).max()
    if bmax == 0:
        return x, 0
    z = vcycle(levels, r)
    s = z.copy()
    rho = (r * z).sum()
    for it in range(1, max_iterations + 1):
        q = apply(level0, s)
        a = rho / (s * q).sum()
        x += a * s
        r -= a * q
        if np.abs(r).max() <= tol * bmax:
            return x, it
        z = vcycle(levels, r)
        rn = (r * z).sum()
        s = z + (rn / rho) * s
        rho = rn
    return x, max_iterations


def scaled_level0(mask, hx, hy, hz):
    """Level 0 from the exported scaled system: unit diagonal on the unknowns."""
    m = (np.asarray(mask) > 0).astype(np.float64)
    return (m, np.asarray(hx, np.float64), np.asarray(hy, np.float64), np.asarray(hz, np.float64))
