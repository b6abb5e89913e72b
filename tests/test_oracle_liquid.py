"""Pins of the oracle's velocity half of the liquid step (oracle/shk_liquid_oracle.c,
macliquid3.cpp:342-391) by closed forms, invariants and brute force on tiny grids."""
from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O


def _faces(n, rng, lo=-1.0, hi=1.0):
    nx, ny, nz = n
    shapes = [(nz, ny, nx + 1), (nz, ny + 1, nx), (nz + 1, ny, nx)]
    return [rng.uniform(lo, hi, s).astype(np.float32) for s in shapes], shapes


CROSS = ndimage.generate_binary_structure(3, 1)


@pytest.mark.parametrize("width", [1, 2, 3])
def test_extrapolate_active_set_is_the_cross_dilation(width):
    rng = np.random.default_rng(width)
    vel, shapes = _faces((9, 7, 6), rng)
    masks = [(rng.random(s) < 0.08).astype(np.uint8) for s in shapes]
    out, ms, total = O.extrapolate(vel, masks, width)
    grown = 0
    for d in range(3):
        want = ndimage.binary_dilation(masks[d].astype(bool), CROSS, iterations=width)
        np.testing.assert_array_equal(ms[d].astype(bool), want)
        grown += int(want.sum() - masks[d].sum())
        keep = masks[d].astype(bool)
        np.testing.assert_array_equal(out[d][keep], vel[d][keep])   # active values untouched
        new = want & ~keep
        if new.any():   # maximum principle: averages of active values
            assert out[d][new].min() >= vel[d][keep].min() and out[d][new].max() <= vel[d][keep].max()
    assert total == grown


def test_extrapolate_half_space_copies_the_last_layer():
    rng = np.random.default_rng(1)
    vel, shapes = _faces((12, 5, 4), rng)
    masks = [np.zeros(s, np.uint8) for s in shapes]
    c = 5
    masks[0][:, :, :c] = 1   # u active on x-faces i < c
    out, ms, _ = O.extrapolate(vel, masks, 3)
    for t in range(3):
        np.testing.assert_array_equal(out[0][:, :, c + t], vel[0][:, :, c - 1])
    assert ms[0][:, :, c + 3:].sum() == 0


def test_extrapolate_corner_average_order():
    vel = [np.zeros((3, 3, 4), np.float32), np.zeros((3, 4, 3), np.float32), np.zeros((4, 3, 3), np.float32)]
    masks = [np.zeros_like(v, np.uint8) for v in vel]
    a, b, c = np.float32(0.1), np.float32(0.7), np.float32(0.3)
    # u array: cell (0,0,0) inactive with active +x (1,0,0), +y (0,1,0) and +z (0,0,1)
    vel[0][0, 0, 1], vel[0][0, 1, 0], vel[0][1, 0, 0] = a, b, c
    masks[0][0, 0, 1] = masks[0][0, 1, 0] = masks[0][1, 0, 0] = 1
    out, _, _ = O.extrapolate(vel, masks, 1)
    s = np.float32(0.0)
    for v in (a, b, c):   # +x, then +y, then +z: the order of array_extrapolator3.h:54-55
        s = np.float32(s + v)
    assert out[0][0, 0, 0] == np.float32(s / np.float32(3))


def test_extrapolate_everything_active_is_a_no_op():
    rng = np.random.default_rng(2)
    vel, shapes = _faces((5, 4, 3), rng)
    masks = [np.ones(s, np.uint8) for s in shapes]
    out, ms, total = O.extrapolate(vel, masks, 4)
    assert total == 0
    for d in range(3):
        np.testing.assert_array_equal(out[d], vel[d])


def test_compute_max_u_closed_form_and_activity():
    n = (6, 5, 4)
    rng = np.random.default_rng(3)
    _, shapes = _faces(n, rng)
    vel = [np.full(s, u, np.float32) for s, u in zip(shapes, (1.0, 2.0, 2.0))]
    masks = [np.ones(s, np.uint8) for s in shapes]
    assert O.compute_max_u(vel, masks) == 3.0
    # one big face that is inactive does not count; the same face active does
    vel[1][2, 3, 1] = 100.0
    masks[1][2, 3, 1] = 0
    assert O.compute_max_u(vel, masks) == 3.0
    masks[1][2, 3, 1] = 1
    want = np.sqrt(1.0 + 4.0 + float(np.float32(0.5 * (100.0 + 2.0)) ** 2))
    assert abs(O.compute_max_u(vel, masks) - want) < 1e-9 * want
    assert O.compute_max_u(vel, [np.zeros(s, np.uint8) for s in shapes]) == 0.0


def test_gravity_activates_every_face():
    rng = np.random.default_rng(4)
    vel, shapes = _faces((4, 3, 5), rng)
    masks = [(rng.random(s) < 0.5).astype(np.uint8) for s in shapes]
    out, ms = O.add_gravity(0.01, (0.0, -9.8, 0.0), vel, masks)
    g = np.float32(0.01 * -9.8)
    for d in range(3):
        assert ms[d].all()
        add = g if d == 1 else np.float32(0.0)
        want = np.where(masks[d] == 1, vel[d] + add, np.float32(0.0) + add).astype(np.float32)
        np.testing.assert_array_equal(out[d], want)


def test_project_activity():
    shapes = [(2, 2, 3), (2, 3, 2), (3, 2, 2)]
    areas = [np.ones(s) for s in shapes]
    rhos = [np.ones(s) for s in shapes]
    areas[0][0, 0, 0] = 0.0
    rhos[2][1, 1, 1] = 0.0
    masks = [np.ones(s, np.uint8) for s in shapes]
    masks[1][0, 0, 0] = 0
    out = O.project_activity(areas, rhos, masks)
    assert out[0][0, 0, 0] == 0 and out[2][1, 1, 1] == 0 and out[1][0, 0, 0] == 0
    assert sum(int(m.sum()) for m in out) == sum(int(np.prod(s)) for s in shapes) - 3


def _shear(n, dx, a, b, c):
    """u = a + b*y on the x-faces, v = c, w = 0 (all faces active)."""
    nx, ny, nz = n
    u = np.empty((nz, ny, nx + 1), np.float32)
    y = (np.arange(ny) + 0.5) * dx
    u[:] = (a + b * y)[None, :, None]
    v = np.full((nz, ny + 1, nx), c, np.float32)
    w = np.zeros((nz + 1, ny, nx), np.float32)
    masks = [np.ones(x.shape, np.uint8) for x in (u, v, w)]
    return [u, v, w], masks


@pytest.mark.parametrize("maccormack", [True, False])
@pytest.mark.parametrize("deep", [True, False])
def test_advect_linear_shear_is_an_exact_shift(maccormack, deep):
    n, dx, dt = (16, 24, 8), 1.0 / 24, 0.01
    a, b, c = 0.3, 2.0, 0.9   # departure 0.216 cells along y
    vel, masks = _shear(n, dx, a, b, c)
    fluid = np.full((n[2], n[1], n[0]), -1.0 if deep else 1.0, np.float32)
    out = O.advect_velocity(dx, dt, vel, masks, fluid, maccormack=maccormack)
    y = (np.arange(n[1]) + 0.5) * dx
    want = a + b * (y - c * dt)
    inner = slice(1, n[1] - 1)
    np.testing.assert_allclose(out[0][:, inner, :], np.broadcast_to(want[None, inner, None], out[0][:, inner, :].shape),
                               rtol=2e-6, atol=2e-7)
    np.testing.assert_allclose(out[1], c, rtol=1e-6)
    assert np.all(out[2] == 0.0)


def test_advect_zero_velocity_and_pure_shear_are_invariant():
    n, dx, dt = (10, 12, 6), 1.0 / 12, 0.02
    vel, masks = _shear(n, dx, 0.0, 1.0, 0.0)   # u = y, v = w = 0: the flow moves along x, u does not vary in x
    fluid = -np.ones((n[2], n[1], n[0]), np.float32)
    out = O.advect_velocity(dx, dt, vel, masks, fluid)
    np.testing.assert_allclose(out[0], vel[0], rtol=1e-6, atol=1e-7)
    zero = [np.zeros_like(v) for v in vel]
    out = O.advect_velocity(dx, dt, zero, masks, fluid)
    for d in range(3):
        assert np.all(out[d] == 0.0)


def test_advect_limiter_and_narrow_band_switch():
    rng = np.random.default_rng(5)
    n, dx, dt = (14, 12, 10), 1.0 / 14, 0.05
    vel, shapes = _faces(n, rng)
    masks = [np.ones(s, np.uint8) for s in shapes]
    # a step in u along y: MacCormack without a limiter overshoots it
    vel[0][:] = np.where(np.arange(n[1])[None, :, None] >= n[1] // 2, 1.0, 0.0).astype(np.float32)
    deep = -np.ones((n[2], n[1], n[0]), np.float32)
    mc = O.advect_velocity(dx, dt, vel, masks, deep, maccormack=True)
    assert mc[0].min() >= 0.0 and mc[0].max() <= 1.0
    for d in range(3):
        assert mc[d].min() >= vel[d].min() and mc[d].max() <= max(vel[d].max(), 0.0)
    # the switch: near-surface faces take the semi-Lagrangian value, deep faces the MacCormack one
    x = (np.arange(n[0]) + 0.5) * dx
    fluid = np.broadcast_to((x - 0.5)[None, None, :], (n[2], n[1], n[0])).astype(np.float32).copy()
    sl = O.advect_velocity(dx, dt, vel, masks, fluid, maccormack=False)
    mixed = O.advect_velocity(dx, dt, vel, masks, fluid, maccormack=True)
    changed = 0
    for d in range(3):
        same_sl = mixed[d] == sl[d]
        same_mc = mixed[d] == mc[d]
        assert np.all(same_sl | same_mc)
        changed += int((~same_sl).sum())
        # faces well inside the liquid (x < 0.5 - 2 dx - displacement) are MacCormack
        nxf = vel[d].shape[2]
        xf = np.arange(nxf) * dx if d == 0 else (np.arange(nxf) + 0.5) * dx
        deep_cols = xf < 0.5 - 4 * dx
        np.testing.assert_array_equal(mixed[d][:, :, deep_cols], mc[d][:, :, deep_cols])
        air_cols = xf > 0.5 + 2 * dx
        np.testing.assert_array_equal(mixed[d][:, :, air_cols], sl[d][:, :, air_cols])
    assert changed > 0


def test_advect_inactive_faces_read_zero_and_come_back_zero():
    rng = np.random.default_rng(6)
    n, dx, dt = (8, 7, 6), 1.0 / 8, 0.03
    vel, shapes = _faces(n, rng)
    masks = [(rng.random(s) < 0.7).astype(np.uint8) for s in shapes]
    fluid = -np.ones((n[2], n[1], n[0]), np.float32)
    out = O.advect_velocity(dx, dt, vel, masks, fluid)
    garbage = [np.where(m == 1, v, np.float32(1e3)).astype(np.float32) for v, m in zip(vel, masks)]
    out2 = O.advect_velocity(dx, dt, garbage, masks, fluid)
    for d in range(3):
        np.testing.assert_array_equal(out[d], out2[d])
        assert np.all(out[d][masks[d] == 0] == 0.0)


def test_constrain_planar_floor():
    rng = np.random.default_rng(7)
    n, dx = (8, 10, 6), 1.0 / 10
    nx, ny, nz = n
    vel, shapes = _faces(n, rng)
    masks = [np.ones(s, np.uint8) for s in shapes]
    c = 0.33   # solid where y < c (nodal level set y - c)
    yn = np.arange(ny + 1) * dx
    solid = np.broadcast_to((yn - c)[None, :, None], (nz + 1, ny + 1, nx + 1)).astype(np.float32).copy()
    out = O.constrain_velocity(dx, solid, vel, masks)
    # v faces lie at y = j dx: inside the floor a downward v loses its normal part
    inside = (np.arange(ny + 1) * dx < c)[None, :, None]
    want_v = np.where(inside & (vel[1] < 0), np.float32(0.0), vel[1])
    want_v[:, ny, :] = np.where(want_v[:, ny, :] > 0, 0.0, want_v[:, ny, :])   # the ceiling
    # the float gradient of the float level set is (0, 1 +- 1 ulp, 0), so the removed
    # normal part leaves a residue of order 1e-7 |v|
    np.testing.assert_allclose(out[1], want_v, rtol=0, atol=1e-6)
    above = (np.arange(ny + 1) * dx > c) & (np.arange(ny + 1) < ny)
    np.testing.assert_array_equal(out[1][:, above, :], vel[1][:, above, :])
    # tangential components: the interpolation lands on the face, and the normal has no
    # x part (the x coefficients cancel pairwise) and a z part of float-rounding size (the
    # z coefficients are accumulated four negative terms first)
    np.testing.assert_array_equal(out[0], vel[0])
    np.testing.assert_allclose(out[2], vel[2], rtol=0, atol=1e-6)


def test_constrain_tilted_plane_removes_the_normal_component():
    n, dx = (12, 12, 4), 1.0 / 12
    nx, ny, nz = n
    th = 0.6
    nrm = np.array([np.cos(th), np.sin(th), 0.0])
    U = np.array([0.2, -1.0, 0.3])
    shapes = [(nz, ny, nx + 1), (nz, ny + 1, nx), (nz + 1, ny, nx)]
    vel = [np.full(s, U[d], np.float32) for d, s in enumerate(shapes)]
    masks = [np.ones(s, np.uint8) for s in shapes]
    xs = np.arange(nx + 1) * dx
    ys = np.arange(ny + 1) * dx
    phi = (xs[None, None, :] * nrm[0] + ys[None, :, None] * nrm[1] - 0.45)
    solid = np.broadcast_to(phi, (nz + 1, ny + 1, nx + 1)).astype(np.float32).copy()
    out = O.constrain_velocity(dx, solid, vel, masks)
    proj = U - nrm * (U @ nrm)   # U . n < 0: the normal part goes
    for d in range(3):
        f = np.indices(shapes[d])[::-1].astype(np.float64)   # (i, j, k) of each face
        px = f[0] + 0.5 * (d != 0)
        py = f[1] + 0.5 * (d != 1)
        inside = (px * dx * nrm[0] + py * dx * nrm[1] - 0.45) < -1e-6
        if d == 1:   # the floor and ceiling rule of the y axis comes after
            inside &= (f[1] > 0) & (f[1] < ny)
        got = out[d][inside]
        np.testing.assert_allclose(got, proj[d], rtol=0, atol=5e-6)
        outside = (px * dx * nrm[0] + py * dx * nrm[1] - 0.45) > 1e-6
        keep = out[d][outside]
        if d == 1:   # the y walls still apply to v
            keep = keep[(f[1][outside] > 0) & (f[1][outside] < ny)]
        np.testing.assert_array_equal(keep, np.float32(U[d]))


def _uniform(n, U):
    nx, ny, nz = n
    shapes = [(nz, ny, nx + 1), (nz, ny + 1, nx), (nz + 1, ny, nx)]
    return [np.full(s, U[d], np.float32) for d, s in enumerate(shapes)], [np.ones(s, np.uint8) for s in shapes]


@pytest.mark.parametrize("maccormack", [True, False])
def test_levelset_linear_field_shifts_exactly(maccormack):
    n, dx, dt = (20, 18, 16), 1.0 / 20, 0.01
    U = (0.7, -0.4, 0.3)
    vel, masks = _uniform(n, U)
    k, j, i = np.indices((n[2], n[1], n[0])).astype(np.float64)
    phi = (0.3 * i + 0.5 * j - 0.2 * k - 4.0) * dx
    band = np.ones(phi.shape, np.uint8)
    out = O.advect_levelset(dx, dt, phi.astype(np.float32), band, vel, masks, background=3 * dx,
                            maccormack=maccormack)
    want = ((0.3 * (i - dt * U[0] / dx) + 0.5 * (j - dt * U[1] / dx) - 0.2 * (k - dt * U[2] / dx)) - 4.0) * dx
    inner = (slice(1, -1),) * 3
    np.testing.assert_allclose(out[inner], want[inner], rtol=0, atol=2e-6)


def test_levelset_still_cells_and_inactive_cells_keep_their_values():
    rng = np.random.default_rng(8)
    n, dx, dt = (10, 9, 8), 1.0 / 10, 0.02
    vel, shapes = _faces(n, rng)
    masks = [np.ones(s, np.uint8) for s in shapes]
    masks[0][3, 4, :] = 0    # the cells next to these faces do not move
    phi = rng.uniform(-0.3, 0.3, (n[2], n[1], n[0])).astype(np.float32)
    band = (rng.random(phi.shape) < 0.6).astype(np.uint8)
    out = O.advect_levelset(dx, dt, phi, band, vel, masks, background=0.3)
    np.testing.assert_array_equal(out[band == 0], phi[band == 0])
    still = np.zeros(phi.shape, bool)
    still[3, 4, :] = True
    np.testing.assert_array_equal(out[still & (band == 1)], phi[still & (band == 1)])
    zero = [np.zeros_like(v) for v in vel]
    np.testing.assert_array_equal(O.advect_levelset(dx, dt, phi, band, zero, masks, background=0.3), phi)


def test_levelset_backward_pass_reads_the_background_outside_the_band():
    rng = np.random.default_rng(9)
    n, dx, dt = (12, 12, 12), 1.0 / 12, 0.04
    vel, shapes = _faces(n, rng)
    masks = [np.ones(s, np.uint8) for s in shapes]
    phi = rng.uniform(-0.3, 0.3, (n[2], n[1], n[0])).astype(np.float32)
    band = np.zeros(phi.shape, np.uint8)
    band[:, :, 4:8] = 1
    a = O.advect_levelset(dx, dt, phi, band, vel, masks, background=0.25)
    b = O.advect_levelset(dx, dt, phi, band, vel, masks, background=-0.25)
    differs = a != b
    assert differs.any()
    assert not differs[:, :, 5:7].all()   # only cells whose backward departure leaves the band
    assert np.all(band[differs] == 1)
    sa = O.advect_levelset(dx, dt, phi, band, vel, masks, background=0.25, maccormack=False)
    sb = O.advect_levelset(dx, dt, phi, band, vel, masks, background=-0.25, maccormack=False)
    np.testing.assert_array_equal(sa, sb)   # the semi-Lagrangian pass never reads it


def test_weno_advection_is_exact_for_cubics(oracle):
    """WENO=1 (WENO3::interpolate, order 6): every candidate stencil of interp6 reproduces
    a cubic and the weights sum to one, so a tensor-product cubic level set moved by a
    constant velocity lands on the cubic at the departure points (to float rounding of
    the samples), while trilinear interpolation errs by its second-order term."""
    n = 16
    dx = 1.0 / n
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")

    def cubic(x, y, z):
        return 1e-4 * ((x - 7.3) ** 3 - 2 * (y - 8.1) ** 2 * (y - 3.0) + (z - 6.6) ** 3 + x * y * z)

    phi = cubic(i, j, k).astype(np.float32)
    band = np.ones_like(phi, dtype=np.uint8)
    shift = (0.37, -0.61, 0.23)  # cells travelled per step along x, y, z
    dt = 1.0
    vel, masks = [], []
    for d in range(3):
        shp = (n + (d == 2), n + (d == 1), n + (d == 0))
        vel.append(np.full(shp, shift[d] * dx / dt, dtype=np.float32))
        masks.append(np.ones(shp, dtype=np.uint8))
    inner = (i >= 3) & (i < n - 3) & (j >= 3) & (j < n - 3) & (k >= 3) & (k < n - 3)
    exact = cubic(i - shift[0], j - shift[1], k - shift[2])
    w = oracle.advect_levelset(dx, dt, phi, band, vel, masks, 1.0, maccormack=False, weno=True)
    t = oracle.advect_levelset(dx, dt, phi, band, vel, masks, 1.0, maccormack=False, weno=False)
    scale = np.abs(exact[inner]).max()
    err_w = np.abs(w[inner] - exact[inner]).max() / scale
    err_t = np.abs(t[inner] - exact[inner]).max() / scale
    assert err_w < 1e-5, err_w
    assert err_t > 100 * err_w, (err_t, err_w)
