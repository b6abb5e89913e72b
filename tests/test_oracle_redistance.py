"""Pins the oracle's PDE redistancer (oracle.redistance, the restatement of
pderedistancer3::redistance, src/redistancer/pderedistancer3.cpp:39-183) by properties
the mathematics fixes, independent of its code:
- the band: trimmed to cells next to a sign change, dilated `width` times (checked
  against scipy's binary dilation with the 6-neighbourhood), cut at |phi| > width*dx;
- a plane signed distance is a fixed point next to the interface (the one-sided
  interface differences are exact for linear fields);
- a plane of slope 2 or 1/2 is driven to the unit-slope distance, ever closer with more
  pseudo-time steps, and a sphere's distance is kept near its surface;
- the sign of every cell next to the interface is kept;
- off the band: -/+ width*dx by the sign of the value;
- Euler and RK2 both reach the distance and agree with each other."""
from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage


def grid(n):
    dx = 1.0 / n
    x = (np.arange(n) + 0.5) * dx
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return dx, X, Y, Z


def test_band_is_trimmed_dilated_and_cut(oracle):
    n, w = 24, 3
    dx, X, Y, Z = grid(n)
    d = np.sqrt((X - 0.5) ** 2 + (Y - 0.47) ** 2 + (Z - 0.52) ** 2) - 0.27
    phi = d.astype(np.float32)
    band = (np.abs(phi) < 5 * dx).astype(np.uint8)
    out, m, cnt = oracle.redistance(dx, phi, band, w)
    # the trim keeps band cells with an opposite-sign band neighbour along an axis
    neg = (phi < 0) & (band > 0)
    pos = (phi > 0) & (band > 0)
    st = ndimage.generate_binary_structure(3, 1)
    near = (neg & ndimage.binary_dilation(pos, st)) | (pos & ndimage.binary_dilation(neg, st))
    grown = ndimage.binary_dilation(near, st, iterations=w)
    assert np.all(m <= grown) and cnt == m.sum()
    assert np.all(np.abs(out[m > 0]) <= w * dx + 1e-7)
    # every grown cell whose value stayed within the band width is on the band
    np.testing.assert_array_equal(m > 0, grown & (np.abs(out) <= w * dx) & (np.abs(out) < w * dx + 1e-12))
    # off the band: the level-set fill of the cell's side
    off = m == 0
    np.testing.assert_array_equal(out[off], np.where(out[off] < 0, -w * dx, w * dx).astype(np.float32))
    assert np.all(np.sign(out[off]) == np.where(phi[off] < 0, -1, 1))


@pytest.mark.parametrize("w", [3, 6])
def test_plane_distance_is_a_fixed_point_at_the_interface(oracle, w):
    dx, X, Y, Z = grid(32)
    d = X - 0.431
    phi = d.astype(np.float32)
    out, m, _ = oracle.redistance(dx, phi, (np.abs(phi) < w * dx).astype(np.uint8), w)
    near = np.abs(d) < dx
    assert np.abs(out[near] - phi[near]).max() <= 1e-6 * dx


@pytest.mark.parametrize("slope", [2.0, 0.5])
def test_slope_is_driven_to_the_distance(oracle, slope):
    dx, X, Y, Z = grid(32)
    d = X - 0.431
    phi = (slope * d).astype(np.float32)
    errs = []
    for w in (3, 8):
        out, m, _ = oracle.redistance(dx, phi, (np.abs(phi) < w * dx).astype(np.uint8), w)
        near = (np.abs(d) < 2.5 * dx) & (m > 0)
        errs.append(np.abs(out[near] - d[near]).max() / dx)
        assert np.all(np.sign(out[np.abs(d) < dx]) == np.sign(d[np.abs(d) < dx]))
    assert errs[1] < errs[0] and errs[1] < 0.02, errs


def test_sphere_distance_is_kept_near_its_surface(oracle):
    dx, X, Y, Z = grid(40)
    d = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2) - 0.3
    phi = d.astype(np.float32)
    out, m, _ = oracle.redistance(dx, phi, (np.abs(phi) < 3 * dx).astype(np.uint8), 3)
    near = np.abs(d) < dx
    # first order near a curved surface: a few hundredths of a cell
    assert np.abs(out[near] - d[near]).max() < 0.1 * dx


def test_euler_and_rk2_agree_to_the_step_order(oracle):
    dx, X, Y, Z = grid(32)
    d = X - 0.431
    phi = (2.0 * d).astype(np.float32)
    band = (np.abs(phi) < 6 * dx).astype(np.uint8)
    e, me, _ = oracle.redistance(dx, phi, band, 6)
    r, mr, _ = oracle.redistance(dx, phi, band, 6, rk2=True)
    both = (me > 0) & (mr > 0) & (np.abs(d) < 3 * dx)
    assert np.abs(e[both] - d[both]).max() < 0.05 * dx
    assert np.abs(r[both] - d[both]).max() < 0.05 * dx
    assert np.abs(e[both] - r[both]).max() < 0.05 * dx
