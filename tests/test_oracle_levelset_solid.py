"""Pins the oracle's restatement of gridutility3::extrapolate_levelset
(oracle/shk_liquid_oracle.c: shk_or_extrapolate_levelset, the checker of
shk_extrapolate_levelset) by closed forms of linear fields, which trilinear
interpolation reproduces exactly: a floor (the cells inside take the level set at the
mirrored point), a tilted wall with and without the horizontal variant (the minimum of
two linear values), the variant without wall extrapolation, and the band rules (one
dilation by the sign rule, the union with the solid's cells) against scipy's binary
dilation."""
from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage

N = (12, 14, 10)  # nx, ny, nz
DX = 1.0 / 14


def cell_index(shape):
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return i.astype(np.float64), j.astype(np.float64), k.astype(np.float64)


def node_index(shape):
    nx, ny, nz = shape
    return cell_index((nx + 1, ny + 1, nz + 1))


def test_no_active_solid_below_zero_changes_nothing(oracle):
    rng = np.random.default_rng(0)
    phi = rng.uniform(-0.2, 0.2, N[::-1]).astype(np.float32)
    band = (rng.random(N[::-1]) < 0.5).astype(np.uint8)
    ni, nj, nk = node_index(N)
    solid = (0.5 + 0.0 * ni).astype(np.float32)
    out, m, cnt = oracle.extrapolate_levelset(DX, solid, phi, band, 3 * DX)
    assert np.array_equal(out, phi) and np.array_equal(m, band) and cnt == int(band.sum())
    # a node below 0 that is off the solid's band does not count either
    solid[2, 3, 4] = -0.1
    sb = np.ones(solid.shape, np.uint8)
    sb[2, 3, 4] = 0
    out, m, cnt = oracle.extrapolate_levelset(DX, solid, phi, band, 3 * DX, solid_band=sb)
    assert np.array_equal(out, phi) and np.array_equal(m, band)


def test_floor_takes_the_mirrored_value(oracle):
    """Solid floor at node row j0 (value (j - j0) dx, gradient (0, dx, 0) in index units),
    level set 0.05 (j - 6.3) on every cell: a cell inside (centre j + 1/2 < j0) moves by
    -s/dx = j0 - j - 1/2 along +y to cell-array row j0 - 1/2, so it takes
    0.05 (j0 - 1/2 - 6.3); then every cell is max(value, -dx - s) with s the cell-centre
    solid value (the mean of its 8 nodes)."""
    j0 = 3
    ci, cj, ck = cell_index(N)
    ni, nj, nk = node_index(N)
    solid = ((nj - j0) * DX).astype(np.float32)
    phi = (0.05 * (cj - 6.3)).astype(np.float32)
    band = np.ones(N[::-1], np.uint8)
    out, m, cnt = oracle.extrapolate_levelset(DX, solid, phi, band, 3 * DX)
    s = (cj + 0.5 - j0) * DX
    ext = np.where(s < 0, 0.05 * (j0 - 0.5 - 6.3), 0.05 * (cj - 6.3))
    want = np.maximum(ext, -DX - s)
    np.testing.assert_allclose(out, want, rtol=0, atol=2e-7)
    assert m.all() and cnt == m.size


@pytest.mark.parametrize("horizontal", [True, False])
def test_tilted_wall_and_the_horizontal_variant(oracle, horizontal):
    """Wall with unit normal n = (2, 1, 0)/sqrt(5) (y part below sin(pi/4)); level set
    a x + b y + c.  Inside, the normal point gives phi + t (a n_x + b n_y) with t = -s/dx;
    the horizontal variant moves along (1, 0, 0) instead and keeps the minimum."""
    i0, j0 = 4.0, 5.0
    a, b, c = -0.03, 0.05, 0.01
    ci, cj, ck = cell_index(N)
    ni, nj, nk = node_index(N)
    r5 = np.sqrt(5.0)
    solid = ((2 * (ni - i0) + (nj - j0)) / r5 * DX).astype(np.float32)
    phi = (a * ci + b * cj + c).astype(np.float32)
    band = np.ones(N[::-1], np.uint8)
    out, m, _ = oracle.extrapolate_levelset(DX, solid, phi, band, 3 * DX, horizontal=horizontal)
    s = (2 * (ci + 0.5 - i0) + (cj + 0.5 - j0)) / r5 * DX
    t = -s / DX
    base = a * ci + b * cj + c
    along_n = base + t * (a * 2 / r5 + b / r5)
    along_h = base + t * a
    ext = np.minimum(along_n, along_h) if horizontal else along_n
    inside = s < 0
    # keep to cells whose moved points stay inside the grid (no clamping)
    ok = inside & (ci + t * 2 / r5 <= N[0] - 1) & (cj + t / r5 <= N[1] - 1) & (ci + t <= N[0] - 1)
    want = np.maximum(np.where(inside, ext, base), -DX - s)
    assert ok.sum() > 50
    np.testing.assert_allclose(out[ok], want[ok], rtol=0, atol=2e-7)
    np.testing.assert_allclose(out[~inside], want[~inside], rtol=0, atol=2e-7)
    if horizontal:
        assert np.any(along_h[ok] < along_n[ok] - 1e-3)  # the minimum matters on this case


def test_without_wall_extrapolation_cells_take_the_solid_value(oracle):
    j0 = 3
    ci, cj, ck = cell_index(N)
    ni, nj, nk = node_index(N)
    solid = ((nj - j0) * DX).astype(np.float32)
    phi = (0.05 * (cj - 6.3)).astype(np.float32)
    band = np.ones(N[::-1], np.uint8)
    out, _, _ = oracle.extrapolate_levelset(DX, solid, phi, band, 3 * DX, wall=False)
    s = (cj + 0.5 - j0) * DX
    want = np.maximum(np.where(s < 0, s, 0.05 * (cj - 6.3)), -DX - s)
    np.testing.assert_allclose(out, want, rtol=0, atol=2e-7)


def test_band_dilation_and_union_with_the_solid_cells(oracle):
    """The solid sits in a corner away from the band (no band cell is inside it), so the
    extrapolation changes no value: the new band is the band, plus its 6-neighbours whose
    value is >= 0 with a neighbour below 0 (set to dx), plus every cell with an active
    solid node; off it, -/+ background by the sign of max(phi, -dx - s)."""
    rng = np.random.default_rng(3)
    ci, cj, ck = cell_index(N)
    ni, nj, nk = node_index(N)
    phi = (0.02 * (cj - 7.2) + 0.004 * rng.standard_normal(N[::-1])).astype(np.float32)
    band = ((cj >= 6) & (cj <= 8)).astype(np.uint8)
    solid = (np.sqrt((ni - 1.0) ** 2 + (nj - 1.0) ** 2 + (nk - 1.0) ** 2) * DX - 1.5 * DX).astype(np.float32)
    sband = (np.abs(solid) < 2 * DX).astype(np.uint8)
    bg = np.float32(3 * DX)
    out, m, cnt = oracle.extrapolate_levelset(DX, solid, phi, band, bg, solid_band=sband)
    st = ndimage.generate_binary_structure(3, 1)
    ring = ndimage.binary_dilation(band.astype(bool), st) & ~band.astype(bool)
    below = ndimage.binary_dilation(phi < 0, st, border_value=0)  # has a 6-neighbour below 0
    new = ring & (phi >= 0) & below
    # cells with an active solid node: dilate the node mask by the 2x2x2 cell stencil
    sb = sband.astype(bool)
    anyn = sb[:-1, :-1, :-1] | sb[1:, :-1, :-1] | sb[:-1, 1:, :-1] | sb[:-1, :-1, 1:] | \
        sb[1:, 1:, :-1] | sb[1:, :-1, 1:] | sb[:-1, 1:, 1:] | sb[1:, 1:, 1:]
    want_band = band.astype(bool) | new | anyn
    assert np.array_equal(m.astype(bool), want_band) and cnt == int(want_band.sum())
    assert np.all(out[new & ~anyn] == np.float32(DX))
    assert np.array_equal(out[band.astype(bool) & ~anyn], phi[band.astype(bool) & ~anyn])
    off = ~want_band
    assert np.all(np.abs(out[off]) == bg)
