"""Pins the FP64 restatement of the learned preconditioner (oracle/unet_oracle.py).

Each check takes an independent route: torch's own conv3d and autograd (CPU,
float64), scipy.ndimage.correlate, brute-force loops, adjoint identities, central
finite differences of the whole network's loss, and closed forms of the loss.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import unet_oracle as U
from unet_cases import init_params as _init_params
from unet_cases import random_system as _random_system
from unet_cases import training_sample as _training_sample


def _rng(seed=0):
    return np.random.default_rng(seed)


def test_param_layout():
    offs, total = U.layer_offsets()
    assert total == sum(27 * ci * co + co for ci, co, _ in U.LAYERS) == 736496
    p = np.arange(total, dtype=np.float64)
    views = U.split_params(p)
    assert views[0][0].shape == (27, 8, 16) and views[3][1].shape == (128,)
    assert views[1][0][0, 0, 0] == offs[1][0] and views[7][1][-1] == total - 1


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5e-3, 3.14159, 65504.0, 1e-30, -7.77e8],
                     dtype=np.float32)
    vals = np.concatenate([vals, _rng(1).standard_normal(4096).astype(np.float32) * 10])
    ref = torch.from_numpy(vals).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(U.bf16_round(vals), ref)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6
    assert U.bf16_round(1.0 + 2 ** -8) == 1.0 and U.bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6


def _torch_conv(x, w, b=None):
    import torch
    xt = torch.from_numpy(np.ascontiguousarray(x.transpose(3, 0, 1, 2)))[None]
    # W[tap][ci][co] -> torch weight [co][ci][kz][ky][kx] (cross-correlation, tap = kz*9 + ky*3 + kx)
    wt = torch.from_numpy(np.ascontiguousarray(w.reshape(3, 3, 3, w.shape[1], w.shape[2]).transpose(4, 3, 0, 1, 2)))
    bt = None if b is None else torch.from_numpy(b)
    return xt, wt, bt, torch.nn.functional.conv3d(xt, wt, bt, padding=1)


def test_conv3d_matches_torch_and_loops():
    pytest.importorskip("torch")
    rng = _rng(2)
    x = rng.standard_normal((3, 4, 5, 8))
    w = rng.standard_normal((27, 8, 16))
    b = rng.standard_normal(16)
    y = U.conv3d(x, w, b)
    _, _, _, yt = _torch_conv(x, w, b)
    np.testing.assert_allclose(y, yt[0].numpy().transpose(1, 2, 3, 0), rtol=1e-12, atol=1e-12)
    # brute force at a few voxels, including corners (zero padding)
    for (z, yy, xx) in [(0, 0, 0), (2, 3, 4), (1, 2, 0), (1, 1, 2)]:
        acc = b.copy()
        for t in range(27):
            dz, dy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
            zz, y2, x2 = z + dz, yy + dy, xx + dx
            if 0 <= zz < 3 and 0 <= y2 < 4 and 0 <= x2 < 5:
                acc += x[zz, y2, x2] @ w[t]
        np.testing.assert_allclose(y[z, yy, xx], acc, rtol=1e-12, atol=1e-12)


def test_conv3d_single_channel_is_scipy_correlate():
    from scipy import ndimage
    rng = _rng(3)
    x = rng.standard_normal((6, 5, 7))
    k = rng.standard_normal((3, 3, 3))
    ref = ndimage.correlate(x, k, mode="constant", cval=0.0)
    y = U.conv3d(x[..., None], k.reshape(27, 1, 1))[..., 0]
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_conv3d_gradients_match_autograd():
    torch = pytest.importorskip("torch")
    rng = _rng(4)
    x = rng.standard_normal((4, 3, 5, 16))
    w = rng.standard_normal((27, 16, 32))
    b = rng.standard_normal(32)
    dy = rng.standard_normal((4, 3, 5, 32))
    xt, wt, bt, yt = _torch_conv(x, w, b)
    for t in (xt, wt, bt):
        t.requires_grad_(True)
    yt = torch.nn.functional.conv3d(xt, wt, bt, padding=1)
    yt.backward(torch.from_numpy(np.ascontiguousarray(dy.transpose(3, 0, 1, 2)))[None])
    dx = U.conv3d_dgrad(dy, w)
    dw, db = U.conv3d_wgrad(x, dy)
    np.testing.assert_allclose(dx, xt.grad[0].numpy().transpose(1, 2, 3, 0), rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(dw, wt.grad.numpy().transpose(2, 3, 4, 1, 0).reshape(27, 16, 32), rtol=1e-11,
                               atol=1e-11)
    np.testing.assert_allclose(db, bt.grad.numpy(), rtol=1e-11, atol=1e-11)


def test_conv3d_adjoint_identities():
    rng = _rng(5)
    x = rng.standard_normal((4, 4, 4, 8))
    w = rng.standard_normal((27, 8, 16))
    dy = rng.standard_normal((4, 4, 4, 16))
    lhs = (U.conv3d(x, w) * dy).sum()
    assert np.isclose(lhs, (x * U.conv3d_dgrad(dy, w)).sum(), rtol=1e-12)
    assert np.isclose(lhs, (w * U.conv3d_wgrad(x, dy)[0]).sum(), rtol=1e-12)


def test_resampling_adjoints():
    rng = _rng(6)
    f = rng.standard_normal((4, 6, 8, 3))
    c = rng.standard_normal((2, 3, 4, 3))
    assert np.isclose((U.pool(f) * c).sum(), (f * U.pool_bwd(c)).sum(), rtol=1e-12)
    assert np.isclose((U.upsample(c) * f).sum(), (c * U.upsample_bwd(f)).sum(), rtol=1e-12)
    # pool of a constant is the constant; upsample then pool is the identity
    np.testing.assert_allclose(U.pool(np.full((4, 4, 4), 2.5)), 2.5)
    np.testing.assert_allclose(U.pool(U.upsample(c)), c, rtol=1e-14)


def _dense_A(cx, cy, cz):
    """The operator as an explicit matrix, assembled face by face."""
    shape = cx.shape
    n = int(np.prod(shape))
    A = np.eye(n)
    idx = np.arange(n).reshape(shape)
    for (c, ax) in ((cx, 2), (cy, 1), (cz, 0)):
        for v in np.ndindex(*shape):
            if v[ax] == 0:
                continue
            u = list(v)
            u[ax] -= 1
            a, b = idx[v], idx[tuple(u)]
            A[a, b] -= c[v]
            A[b, a] -= c[v]
    return A


def test_stencil_is_the_assembled_matrix():
    shape = (3, 4, 5)
    mask, (cx, cy, cz) = _random_system(shape, 7)
    A = _dense_A(cx, cy, cz)
    x = _rng(8).standard_normal(shape)
    np.testing.assert_allclose(U.stencil(cx, cy, cz, x).ravel(), A @ x.ravel(), rtol=1e-13, atol=1e-13)
    assert np.allclose(A, A.T)


def test_prolongation():
    rng = _rng(9)
    u = rng.standard_normal((3, 4, 5))
    g = rng.standard_normal((6, 8, 10))
    # adjoint pair
    assert np.isclose((U.prolong(u) * g).sum(), (u * U.prolong_T(g)).sum(), rtol=1e-12)
    # constants stay constant (clamped edges), and the weights sum to one
    np.testing.assert_allclose(U.prolong(np.full((3, 4, 5), 1.5)), 1.5, rtol=1e-15)
    # a linear function of the cell centres is reproduced exactly away from the edges
    Z, Y, X = np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij")
    lin = 0.3 * X - 1.1 * Y + 0.7 * Z + 2.0
    fz, fy, fx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    fine = 0.3 * ((fx + 0.5) / 2 - 0.5) - 1.1 * ((fy + 0.5) / 2 - 0.5) + 0.7 * ((fz + 0.5) / 2 - 0.5) + 2.0
    np.testing.assert_allclose(U.prolong(lin)[1:-1, 1:-1, 1:-1], fine[1:-1, 1:-1, 1:-1], rtol=1e-13, atol=1e-13)
    # brute force at one interior fine cell: weights (3/4, 1/4) per axis toward the nearer neighbour
    f = (3, 4, 5)  # z odd -> neighbour +1; y even -> -1; x odd -> +1
    ref = 0.0
    for dz, wz in ((0, .75), (1, .25)):
        for dy, wy in ((0, .75), (-1, .25)):
            for dx, wx in ((0, .75), (1, .25)):
                ref += wz * wy * wx * u[1 + dz, 2 + dy, 2 + dx]
    assert np.isclose(U.prolong(u)[f], ref, rtol=1e-14)


def test_loss_closed_forms():
    r, mask, cx, cy, cz, xstar = _training_sample((16, 16, 16), 11)
    p = _init_params(12)
    wo, _ = U.layer_offsets()[0][7]
    p[wo:] = 0.0  # head weights and bias zero: U == 0
    # omega = 0: z = 0, L = x*'A x* / x*'r = 1 because r = A x*
    loss, _, z = U.loss_and_grad(p, r, mask, cx, cy, cz, xstar, 0.0, need_grad=False)
    assert np.all(z == 0) and abs(loss - 1.0) < 1e-12
    # U == 0, omega = w: two weighted-Jacobi steps from zero, z = w M r + w M (r - A w M r),
    # evaluated with the explicit matrix
    w = 0.6
    loss1, _, z1 = U.loss_and_grad(p, r, mask, cx, cy, cz, xstar, w, need_grad=False)
    A = _dense_A(cx, cy, cz)
    M = mask.ravel()
    rv = r.ravel()
    za = w * M * rv
    zb = za + w * M * (rv - A @ za)
    np.testing.assert_allclose(z1.ravel(), zb, rtol=1e-12, atol=1e-14)
    e = zb - xstar.ravel()
    assert abs(loss1 - e @ A @ e / (xstar.ravel() @ rv)) < 1e-12
    # positive homogeneity: M(2r) = 2 M(r)
    p2 = _init_params(13)
    za, _, _ = U.precondition(p2, r, mask, cx, cy, cz, 0.7)
    zb, _, _ = U.precondition(p2, 2.0 * r, mask, cx, cy, cz, 0.7)
    np.testing.assert_allclose(zb, 2.0 * za, rtol=1e-12, atol=1e-12)


def test_network_gradient_matches_finite_differences():
    r, mask, cx, cy, cz, xstar = _training_sample((16, 16, 16), 21)
    p = _init_params(22)
    loss, g, _ = U.loss_and_grad(p, r, mask, cx, cy, cz, xstar, 0.8)
    rng = _rng(23)
    offs, total = U.layer_offsets()
    picks = []
    for wo, bo in offs:  # two weights and one bias of every layer
        picks += list(rng.integers(wo, bo, 2)) + [bo]
    h = 1e-6
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = U.loss_and_grad(pp, r, mask, cx, cy, cz, xstar, 0.8, need_grad=False)[0]
        lm = U.loss_and_grad(pm, r, mask, cx, cy, cz, xstar, 0.8, need_grad=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])) + 1e-7, (i, fd, g[i])
    # the gradient is not trivially zero anywhere it is checked as a whole
    assert np.linalg.norm(g) > 0
