"""Pins the oracle's face fractions (utility::fraction, utility::get_area,
macutility3::compute_area_fraction / compute_fluid_fraction) against closed forms
and an independent polygon clipper -- never against a retyping of their formulas."""
from __future__ import annotations

import numpy as np
import pytest


# ---------------------------------------------------------------- utility::fraction

@pytest.mark.parametrize("a,b,want", [
    (-1.0, -2.0, 1.0), (1.0, 2.0, 0.0), (0.0, -1.0, 1.0), (0.0, 1.0, 0.0), (0.0, 0.0, 0.0),
    (-1.0, 1.0, 0.5), (-1.0, 3.0, 0.25), (3.0, -1.0, 0.25), (-3.0, 1.0, 0.75),
])
def test_fraction_closed_forms(oracle, a, b, want):
    assert oracle.fraction(a, b) == want


def test_fraction_symmetric_and_linear(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        a, b = rng.normal(size=2)
        assert oracle.fraction(a, b) == oracle.fraction(b, a)
        if a * b < 0:
            # the segment from a to b crosses zero at t = a/(a-b); the negative part's length
            lo = a / (a - b) if a < 0 else 1 - a / (a - b)
            assert oracle.fraction(a, b) == pytest.approx(lo, rel=1e-14)


def test_fraction_tiny_values(oracle):
    # phi0*phi1 underflows to -0.0, which the `>= 0` test takes as one side: an
    # inside sample then gives 1 (so the DBL_MIN floor of the gap is never reached)
    tiny = 5e-324
    assert oracle.fraction(-tiny, tiny) == 1.0
    # products that stay representable take the crossing formula
    assert oracle.fraction(-1e-160, 1e-160) == 0.5


# ---------------------------------------------------------------- utility::get_area

def _clip_area(coef) -> float:
    """Exact area of {a x + b y + c < 0} inside the unit square by Sutherland-Hodgman."""
    a, b, c = coef
    poly = [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)]
    f = lambda p: a * p[0] + b * p[1] + c  # noqa: E731
    out = []
    for i in range(4):
        p, q = poly[i], poly[(i + 1) % 4]
        fp, fq = f(p), f(q)
        if fp < 0:
            out.append(p)
        if fp * fq < 0:
            t = fp / (fp - fq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    if len(out) < 3:
        return 0.0
    xs, ys = np.array(out).T
    return 0.5 * abs(np.dot(xs, np.roll(ys, -1)) - np.dot(ys, np.roll(xs, -1)))


def test_get_area_all_in_all_out(oracle):
    assert oracle.get_area(-1, -2, -3, -4) == 1.0
    assert oracle.get_area(1, 2, 3, 4) == 0.0


def test_get_area_one_corner_triangle(oracle):
    # corner (0,0) inside: triangle with legs a/(a-b) and a/(a-d)
    a, b, c, d = -0.3, 0.9, 2.0, 0.5
    want = 0.5 * (a / (a - b)) * (a / (a - d))
    assert oracle.get_area(a, b, c, d) == pytest.approx(want, rel=1e-14)


def test_get_area_linear_fields_match_clipping(oracle):
    """For a linear field the marching-squares polygon is exact: compare with an
    independent clipper over many random half-planes (one to three corners inside)."""
    rng = np.random.default_rng(7)
    checked = 0
    for _ in range(500):
        a, b = rng.normal(size=2)
        c = rng.uniform(-1.5, 0.5)
        v = [c, a + c, a + b + c, b + c]  # (0,0), (1,0), (1,1), (0,1)
        want = _clip_area((a, b, c))
        got = oracle.get_area(v[0], v[1], v[2], v[3])
        assert got == pytest.approx(want, abs=1e-12)
        checked += 0 < want < 1
    assert checked > 200


def test_get_area_saddle(oracle):
    # opposite corners inside: the walk joins them into one hexagon (no disambiguation)
    assert oracle.get_area(-1.0, 1.0, -1.0, 1.0) == pytest.approx(0.75)
    assert oracle.get_area(1.0, -1.0, 1.0, -1.0) == pytest.approx(0.75)


# ---------------------------------------------------------------- grid fractions

def test_area_fraction_walls_and_open(oracle):
    n = 6
    solid = np.ones((n + 1, n + 1, n + 1))
    ax, ay, az = oracle.area_fraction(solid)
    for a, arr in enumerate((ax, ay, az)):
        idx = [slice(None)] * 3
        idx[2 - a] = 0
        assert np.all(arr[tuple(idx)] == 0)
        idx[2 - a] = -1
        assert np.all(arr[tuple(idx)] == 0)
        inner = [slice(None)] * 3
        inner[2 - a] = slice(1, -1)
        assert np.all(arr[tuple(inner)] == 1)


def test_area_fraction_planar_solid(oracle):
    """Solid x < c: y/z faces are open by the part of the face with x > c."""
    n = 8
    c = 3.3
    k, j, i = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    solid = (i - c).astype(np.float64)  # nodal values in cell units
    ax, ay, az = oracle.area_fraction(solid, eps_solid=0.0, real_is_double=True)
    # y faces at (i, j, k) span x in [i, i+1]
    for ii in range(n):
        want = min(max(ii + 1 - c, 0.0), 1.0)
        assert np.allclose(ay[1:-1, 1:n, ii], want, atol=1e-14)
        assert np.allclose(az[1:n, :, ii], want, atol=1e-14)
    # x faces are fully open (x > c) or closed
    assert np.all(ax[:, :, 4:n] == 1) and np.all(ax[:, :, 1:4] == 0)


def test_area_fraction_eps_clamp(oracle):
    n = 4
    k, j, i = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    solid = (i - 1.997).astype(np.float64)   # open part of y faces at i=1 is 0.003
    ay = oracle.area_fraction(solid, eps_solid=1e-2, real_is_double=True)[1]
    assert np.allclose(ay[1:-1, 1:n, 1], 1e-2)
    ay0 = oracle.area_fraction(solid, eps_solid=0.0, real_is_double=True)[1]
    assert np.allclose(ay0[1:-1, 1:n, 1], 0.003, atol=1e-12)


def test_fluid_fraction_planar_and_clamped(oracle):
    """phi = y - h (cell units): a y face between an inside and an outside cell is
    liquid over (h - y_low) of its dual segment; boundary faces clamp indices."""
    n = 6
    h = 2.3
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    phi = (j + 0.5 - h).astype(np.float64)
    rx, ry, rz = oracle.fluid_fraction(phi, eps_fluid=0.0, real_is_double=True)
    # face j between cells j-1 and j; cell j-1 = 1 is inside (1.5 < 2.3), cell 2 outside
    assert np.allclose(ry[:, 2, :], h - 1.5, atol=1e-14)
    assert np.all(ry[:, 1, :] == 1) and np.all(ry[:, 3:, :] == 0)
    # clamped boundary faces: both samples from the same cell
    assert np.all(ry[:, 0, :] == 1) and np.all(ry[:, n, :] == 0)
    assert np.all(rx[:, :2, :] == 1) and np.all(rx[:, 2:, :] == 0)


def test_fractions_float_storage(oracle):
    rng = np.random.default_rng(3)
    phi = rng.normal(size=(5, 5, 5))
    rx = oracle.fluid_fraction(phi, real_is_double=False)[0]
    assert np.array_equal(rx, rx.astype(np.float32).astype(np.float64))
    rxd = oracle.fluid_fraction(phi, real_is_double=True)[0]
    assert np.array_equal(rx, rxd.astype(np.float32).astype(np.float64))
