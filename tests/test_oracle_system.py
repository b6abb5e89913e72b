"""Pins the oracle's labelling, assembly, update and divergence
(macpressuresolver3.cpp:134-284) with semantics checks, M-matrix invariants,
closed-form rows and the post-projection divergence identity."""
from __future__ import annotations

import numpy as np
import pytest


def _empty(n):
    fluid = np.ones((n, n, n))
    solid = np.ones((n + 1, n + 1, n + 1))
    vel = [np.zeros((n, n, n + 1)), np.zeros((n, n + 1, n)), np.zeros((n + 1, n, n))]
    return fluid, solid, vel


def _label(oracle, fluid, solid):
    areas = oracle.area_fraction(solid)
    rhos = oracle.fluid_fraction(fluid)
    return oracle.label(fluid, areas, rhos)


def test_label_isolated_cell_is_not_an_unknown(oracle):
    fluid, solid, _ = _empty(5)
    fluid[2, 2, 2] = -0.5
    _, n = _label(oracle, fluid, solid)
    assert n == 0


def test_label_pair_and_closed_face(oracle):
    fluid, solid, _ = _empty(5)
    fluid[2, 2, 1] = fluid[2, 2, 2] = -0.5
    im, n = _label(oracle, fluid, solid)
    assert n == 2 and im[2, 2, 1] >= 0 and im[2, 2, 2] >= 0
    # close the x face between them: all nodes on the plane x = 2 inside the solid
    solid[:, :, 2] = -1.0
    im, n = _label(oracle, fluid, solid)
    assert n == 0


def test_label_relation_is_symmetric(oracle, scenes):
    sc = scenes(3, 16, 2)
    fluid = sc["fluid"].astype(np.float64)
    areas = oracle.area_fraction(sc["solid"])
    rhos = oracle.fluid_fraction(fluid)
    im, n = oracle.label(fluid, areas, rhos)
    lab = im >= 0
    assert n == lab.sum() and n > 0
    assert np.all(fluid[lab] < 0)
    # a liquid neighbour across an open face of an unknown is an unknown
    liquid = fluid < 0
    for a in range(3):
        op = (areas[a] != 0) & (rhos[a] != 0)
        sl = [slice(None)] * 3
        sl[2 - a] = slice(1, -1)
        openf = op[tuple(sl)]
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[2 - a] = slice(0, -1)
        hi[2 - a] = slice(1, None)
        pair = openf & liquid[tuple(lo)] & liquid[tuple(hi)]
        assert np.all(lab[tuple(lo)][pair]) and np.all(lab[tuple(hi)][pair])
    # numbering is a bijection onto 0..N-1 in linear order
    assert np.array_equal(np.sort(im[lab]), np.arange(n))


@pytest.fixture(scope="module")
def system(oracle, scenes):
    sc = scenes(3, 16, 4)
    out = oracle.project(sc, tol=1e-10, mode=oracle.SHK_OR_JACOBI, real_is_double=False)
    return sc, out


def test_matrix_invariants(oracle, system):
    sc, o = system
    A = oracle.to_scipy(o["rowptr"], o["cols"], o["vals"])
    assert abs(A - A.T).max() == 0
    d = A.diagonal()
    assert np.all(d > 0)
    off = A - np.diag(d) if False else A.copy()
    off.setdiag(0)
    off.eliminate_zeros()
    assert off.data.max() < 0
    rowsum = np.asarray(A.sum(axis=1)).ravel()
    assert np.all(rowsum >= -1e-9 * d)          # weakly diagonally dominant
    assert np.any(rowsum > 1e-6 * d)           # Dirichlet rows exist (free surface)
    assert np.all(np.diff(o["rowptr"]) <= 7)
    # columns sorted within each row (RCMatrix keeps them sorted)
    for r in range(0, len(o["rowptr"]) - 1, 97):
        c = o["cols"][o["rowptr"][r]:o["rowptr"][r + 1]]
        assert np.all(np.diff(c) > 0)


def test_interior_row_closed_form(oracle):
    """A fully liquid, fully open neighbourhood gives (dt/dx^2) * [6, -1 x 6]."""
    n = 6
    fluid = -np.ones((n, n, n))
    fluid[:, n - 1, :] = 0.5                      # air layer on top
    solid = np.ones((n + 1, n + 1, n + 1))
    vel = [np.zeros((n, n, n + 1)), np.zeros((n, n + 1, n)), np.zeros((n + 1, n, n))]
    dt, dx = 0.7, 0.25
    areas = oracle.area_fraction(solid)
    rhos = oracle.fluid_fraction(fluid)
    im, N = oracle.label(fluid, areas, rhos)
    rowptr, cols, vals, rhs = oracle.assemble(dt, dx, fluid, areas, rhos, vel, vel, im, N)
    r = im[2, 2, 2]
    row = vals[rowptr[r]:rowptr[r + 1]]
    c = dt / dx ** 2
    assert len(row) == 7
    assert np.allclose(sorted(row), [-c] * 6 + [6 * c], rtol=1e-15)


def test_rhs_is_minus_divergence_of_linear_field(oracle):
    n = 6
    fluid = -np.ones((n, n, n))
    fluid[:, n - 1, :] = 0.5
    solid = np.ones((n + 1, n + 1, n + 1))
    dx = 1.0 / n
    a, b, c = 0.3, -1.1, 2.0
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n + 1), indexing="ij")
    u = a * i * dx
    k, j, i = np.meshgrid(np.arange(n), np.arange(n + 1), np.arange(n), indexing="ij")
    v = b * j * dx
    k, j, i = np.meshgrid(np.arange(n + 1), np.arange(n), np.arange(n), indexing="ij")
    w = c * k * dx
    zero = [np.zeros_like(u), np.zeros_like(v), np.zeros_like(w)]
    areas = oracle.area_fraction(solid, real_is_double=True)
    rhos = oracle.fluid_fraction(fluid, real_is_double=True)
    im, N = oracle.label(fluid, areas, rhos)
    _, _, _, rhs = oracle.assemble(1.0, dx, fluid, areas, rhos, [u, v, w], zero, im, N)
    interior = np.zeros_like(im, dtype=bool)
    interior[1:-1, 1:-2, 1:-1] = True
    rows = im[interior & (im >= 0)]
    assert np.allclose(rhs[rows], -(a + b + c), rtol=1e-12)


def test_solid_velocity_enters_through_closed_area(oracle):
    """rhs gains -sgn (1 - area) u_s / dx (macpressuresolver3.cpp:210-213): moving
    domain walls with u_s inject their flux into the wall-adjacent rows."""
    n = 4
    fluid = -np.ones((n, n, n))
    fluid[:, n - 1, :] = 0.5
    solid = np.ones((n + 1, n + 1, n + 1))
    dx = 0.25
    zero = [np.zeros((n, n, n + 1)), np.zeros((n, n + 1, n)), np.zeros((n + 1, n, n))]
    us = [z.copy() for z in zero]
    us[0][:, :, 0] = 2.0                     # left wall moves in +x
    areas = oracle.area_fraction(solid, real_is_double=True)
    rhos = oracle.fluid_fraction(fluid, real_is_double=True)
    im, N = oracle.label(fluid, areas, rhos)
    _, _, _, rhs = oracle.assemble(1.0, dx, fluid, areas, rhos, zero, us, im, N)
    left = im[:, :n - 1, 0]
    assert np.allclose(rhs[left[left >= 0]], 2.0 / dx)     # -(-1)(1-0)(2)/dx
    others = np.setdiff1d(np.arange(N), left[left >= 0])
    assert np.all(rhs[others] == 0)


def test_post_divergence_identity(oracle, system):
    """D(u_new) = D(u_old) + A p = -(b - A p) on the unknowns, in FP64."""
    sc, o = system
    svel = [sc["us"], sc["vs"], sc["ws"]]
    p = oracle.scatter(o["index_map"], o["x"], real_is_double=True)
    vel64 = oracle.update_velocity(1.0, sc["dx"], p, o["areas"], o["rhos"], svel, [sc["u"], sc["v"], sc["w"]],
                                   real_is_double=True)
    _, div = oracle.divergence(sc["dx"], o["areas"], vel64, svel, o["index_map"], o["n"])
    resid = o["rhs"] - oracle.spmv(o["rowptr"], o["cols"], o["vals"], o["x"])
    assert np.allclose(div, -resid, atol=1e-9 * np.abs(o["rhs"]).max())
    _, div0 = oracle.divergence(sc["dx"], o["areas"], [sc["u"], sc["v"], sc["w"]], svel, o["index_map"], o["n"])
    assert np.allclose(div0, -o["rhs"], rtol=0, atol=0)


def test_scatter_and_float_update(oracle, system):
    sc, o = system
    p = oracle.scatter(o["index_map"], o["x"], real_is_double=False)
    lab = o["index_map"] >= 0
    assert np.all(p[~lab] == 0)
    assert np.array_equal(p[lab], o["x"][o["index_map"][lab]].astype(np.float32).astype(np.float64))
    vel = o["velocity"]
    for v in vel:
        assert np.array_equal(v, v.astype(np.float32).astype(np.float64))
    # closed faces (area 0 or rho 0) are switched off
    for a in range(3):
        closed = (o["areas"][a] == 0) | (o["rhos"][a] == 0)
        assert np.all(vel[a][closed] == 0)
