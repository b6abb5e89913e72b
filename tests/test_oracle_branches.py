"""Pins of the oracle's project() branches (macpressuresolver3.cpp:75-124, 220-260,
287-293): solid density, first-order switches, surface tension, volume correction
and warm start.  Each is checked against something other than its own formula:
closed forms (quadratic Laplacian, Laplace's law of a static drop), identities of
the projection (uniform divergence under a constant rhs shift), endpoints of the
density blend, and agreement of warm and cold starts."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O


def _drop_scene(n=32, radius=0.25):
    """A liquid sphere at rest in the middle of a closed box, no solids."""
    dx = 1.0 / n
    c = (np.arange(n) + 0.5) * dx
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2)
    sc = {"fluid": (r - radius).astype(np.float32), "solid": np.ones((n + 1,) * 3, np.float32), "dx": dx,
          "shape": (n, n, n)}
    for k, s in zip(("u", "v", "w"), O.face_shapes(n, n, n)):
        sc[k] = np.zeros(s, np.float32)
    for k, s in zip(("us", "vs", "ws"), O.face_shapes(n, n, n)):
        sc[k] = np.zeros(s, np.float32)
    return sc, r


def test_curvature_exact_on_quadratic():
    # phi = x^2 + y^2 + z^2 on cell centres with dx = 1/16 (all values exact in float):
    # the 7-point Laplacian is exactly 6 inside
    n, dx = 12, 1.0 / 16
    i = np.arange(n) * dx
    z, y, x = np.meshgrid(i, i, i, indexing="ij")
    phi = x ** 2 + y ** 2 + z ** 2
    c = O.curvature(phi, dx)
    np.testing.assert_array_equal(c[1:-1, 1:-1, 1:-1], 6.0)
    # at i = n-1 the +x neighbour is clamped to the cell itself, so the x part of the
    # Laplacian is (phi(i-1) + phi(i) - 2 phi(i)) / dx^2 while y and z still give 4
    xs = (n - 1) * dx
    want_x = ((xs - dx) ** 2 + xs ** 2 - 2 * xs ** 2) / dx ** 2
    np.testing.assert_allclose(c[5, 5, n - 1], 4.0 + want_x, rtol=1e-6)


def test_curvature_of_sphere_distance():
    sc, r = _drop_scene(32, 0.25)
    c = O.curvature(sc["fluid"], sc["dx"])
    band = np.abs(r - 0.25) < 2 * sc["dx"]
    ratio = c[band] / (2.0 / r[band])   # the Laplacian of |x| in 3D is 2/|x|
    assert 0.99 < ratio.min() and ratio.max() < 1.01


def test_surface_tension_laplace_law():
    """A static drop: the ghost-fluid jump makes the pressure inside 2*sigma/R and the
    projected velocity vanishes up to the curvature's discretisation error."""
    sigma, radius = 0.1, 0.25
    sc, _ = _drop_scene(32, radius)
    o = O.project(sc, tol=1e-10, mode=O.SHK_OR_JACOBI, surface_tension_coef=sigma)
    lab = o["index_map"] >= 0
    p = o["pressure"][lab]
    np.testing.assert_allclose(p.mean(), 2 * sigma / radius, rtol=0.01)
    assert p.max() - p.min() < 0.01 * 2 * sigma / radius
    jump = max(np.abs(v).max() for v in o["velocity_in"])
    left = max(np.abs(v).max() for v in o["velocity"])
    assert jump > 100 and left < 1e-3 * jump
    # dt cancels: the same pressure at dt = 0.5
    o2 = O.project(sc, dt=0.5, tol=1e-10, mode=O.SHK_OR_JACOBI, surface_tension_coef=sigma)
    np.testing.assert_allclose(o2["pressure"][lab].mean(), p.mean(), rtol=1e-4)


def test_surface_tension_only_on_interface_faces(scenes):
    sc = scenes(1, 16, 0)
    o = O.project(sc, tol=1e-6, mode=O.SHK_OR_CG, surface_tension_coef=0.05)
    for a in range(3):
        changed = o["velocity_in"][a] != np.asarray(sc["uvw"[a]], np.float64)
        rho = o["rhos"][a]
        assert not np.any(changed & ((rho == 0) | (rho == 1)))
        assert np.any(changed)


def test_volume_correction_uniform_divergence(scenes):
    """Solving A p = b + c makes the projected (FP64) divergence equal c on every
    unknown: D(u_new) = D(u) + A p = -b + (b + c)."""
    sc = scenes(3, 16, 2)
    c = -0.37
    o = O.project(sc, tol=1e-13, mode=O.SHK_OR_JACOBI, real_is_double=True, rhs_correct=c)
    svel = [sc[k] for k in ("us", "vs", "ws")]
    _, div = O.divergence(float(sc["dx"]), o["areas"], o["velocity"], svel, o["index_map"], o["n"])
    np.testing.assert_allclose(div, c, atol=1e-8)


def test_volume_correction_controller():
    # off without gain or target; zero error and zero integral give zero
    assert O.volume_correction(0.0, 1.1, 1.0, 0.1, 0.0) == (0.0, 0.0)
    assert O.volume_correction(1.0, 1.1, 0.0, 0.1, 0.0) == (0.0, 0.0)
    assert O.volume_correction(1.0, 1.0, 1.0, 0.1, 0.0)[0] == 0.0
    # too much liquid -> negative divergence (the liquid contracts), and vice versa
    c_hi, y1 = O.volume_correction(1.0, 1.1, 1.0, 0.1, 0.0)
    c_lo, _ = O.volume_correction(1.0, 0.9, 1.0, 0.1, 0.0)
    assert c_hi < 0 < c_lo
    # the integral state accumulates the relative error times dt
    np.testing.assert_allclose(y1, 0.1 * 0.1)
    c2, y2 = O.volume_correction(1.0, 1.1, 1.0, 0.1, y1)
    np.testing.assert_allclose(y2, 2 * 0.1 * 0.1)
    assert c2 < c_hi   # the integral term strengthens a persisting error
    # proportional part: at zero integral, the correction grows with the gain
    c_g2, _ = O.volume_correction(2.0, 1.1, 1.0, 0.0, 0.0)
    c_g1, _ = O.volume_correction(1.0, 1.1, 1.0, 0.0, 0.0)
    np.testing.assert_allclose(c_g2, 2 * c_g1)


@pytest.mark.parametrize("seed", [2, 5])
def test_solid_density_endpoints(scenes, seed):
    sc = scenes(3, 16, seed)
    shapes = O.face_shapes(16, 16, 16)
    base = O.project(sc, tol=1e-8, mode=O.SHK_OR_JACOBI)
    # density 1: a full solid -- areas unchanged, updated velocity scaled by the area
    one = O.project(sc, tol=1e-8, mode=O.SHK_OR_JACOBI, solid_density=[np.ones(s) for s in shapes])
    for a in range(3):
        np.testing.assert_array_equal(one["areas"][a], base["areas"][a])
        want = (np.float32(base["areas"][a]) * np.float32(base["velocity"][a])).astype(np.float64)
        np.testing.assert_array_equal(one["velocity"][a], want)
    # density 0: every face open, walls included; no velocity scaling
    zero = O.project(sc, tol=1e-8, mode=O.SHK_OR_JACOBI, solid_density=[np.zeros(s) for s in shapes])
    for a in range(3):
        np.testing.assert_array_equal(zero["areas"][a], 1.0)
    # in between the blend is affine in the density: the midpoint of the endpoints
    half = O.solid_density_blend(base["areas"], [np.full(s, 0.5) for s in shapes], real_is_double=True)
    for a in range(3):
        np.testing.assert_allclose(half[a], 0.5 * (base["areas"][a] + 1.0), rtol=0, atol=1e-15)


def test_first_order_switches(scenes):
    sc = scenes(3, 16, 2)
    so = O.project(sc, tol=1e-6, mode=O.SHK_OR_CG)
    fo = O.project(sc, tol=1e-6, mode=O.SHK_OR_CG, second_order_solid=False, second_order_fluid=False)
    for a in range(3):
        for key in ("areas", "rhos"):
            x, y = so[key][a], fo[key][a]
            np.testing.assert_array_equal(y == 0, x == 0)
            np.testing.assert_array_equal(y[y != 0], 1.0)
    assert fo["n"] == so["n"]


def test_warm_start(scenes):
    sc = scenes(1, 16, 0)
    cold = O.project(sc, tol=1e-8, mode=O.SHK_OR_CG)
    zero = O.project(sc, tol=1e-8, mode=O.SHK_OR_CG, warm_prev=np.zeros(sc["fluid"].shape))
    np.testing.assert_array_equal(zero["x"], cold["x"])
    assert zero["iters"] == cold["iters"]
    # from a perturbed previous pressure the warm start reaches the same solution
    rng = np.random.default_rng(1)
    prev = cold["prev"] + 0.01 * rng.standard_normal(cold["prev"].shape) * np.abs(cold["x"]).max()
    warm = O.project(sc, tol=1e-10, mode=O.SHK_OR_CG, warm_prev=prev)
    truth = O.project(sc, tol=1e-12, mode=O.SHK_OR_CG)
    np.testing.assert_allclose(warm["x"], truth["x"], rtol=0, atol=1e-6 * np.abs(truth["x"]).max())
    # and the next warm start's state is the solution per cell
    lab = warm["index_map"] >= 0
    np.testing.assert_array_equal(warm["prev"][lab], warm["x"][warm["index_map"][lab]])
