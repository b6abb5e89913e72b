"""FP64 numpy restatement of the learned preconditioner's network and loss.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's BF16
MFMA convolution kernels and of the network's loss and gradients; never part of
the product path.

The reference has no learned preconditioner: its pcg plugin applies none
(pcg_solver.h:379-384, the preconditioner's result is discarded and z = r).  The
network here is this library's own design; what it must agree with is the
mathematics, so every function below is pinned in tests/test_oracle_unet.py by an
independent route (brute-force loops, scipy.ndimage.correlate, adjoint identities,
central finite differences, closed forms), not by a second copy of itself.

Layouts match the library: grids are numpy arrays of shape (nz, ny, nx[, C]) so
the flat voxel index is x + nx*(y + ny*z); convolution weights are W[tap][ci][co]
with tap = (dz+1)*9 + (dy+1)*3 + (dx+1); the parameter vector is, layer by layer,
W (27*cin*cout) followed by the bias (cout).

With bf16=True the restatement rounds (round-to-nearest-even) at the points where
the GPU stores BF16: weights, features, every activation, every activation
gradient.  Accumulations stay in FP64 (the GPU's are FP32), so the two agree to a
few BF16 ulps, not bit for bit.
"""
from __future__ import annotations

import numpy as np

# (cin, cout, relu) of the 8 convolutions: enc0..enc2, bottleneck, dec2..dec0, head
LAYERS = [(8, 16, True), (16, 32, True), (32, 64, True), (64, 128, True),
          (192, 64, True), (96, 32, True), (48, 16, True), (16, 16, False)]


def layer_offsets():
    """(w_offset, b_offset) of each layer in the flat parameter vector, and its length."""
    offs, o = [], 0
    for cin, cout, _ in LAYERS:
        offs.append((o, o + 27 * cin * cout))
        o += 27 * cin * cout + cout
    return offs, o


def num_params() -> int:
    return layer_offsets()[1]


def split_params(p):
    """Views (W[27][cin][cout], b[cout]) of each layer in the flat vector p."""
    offs, _ = layer_offsets()
    out = []
    for (cin, cout, _), (wo, bo) in zip(LAYERS, offs):
        out.append((p[wo:bo].reshape(27, cin, cout), p[bo:bo + cout]))
    return out


def bf16_round(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def _q(a, bf16):
    return bf16_round(a) if bf16 else a


TAPS = [((t // 9) - 1, ((t // 3) % 3) - 1, (t % 3) - 1) for t in range(27)]  # (dz, dy, dx)


def _shifted(xp, dz, dy, dx, shape):
    nz, ny, nx = shape
    return xp[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]


def conv3d(x, w, b=None):
    """y[v, co] = b[co] + sum_{tap, ci} w[tap, ci, co] * x[v + off(tap), ci] (zero padding)."""
    nz, ny, nx, cin = x.shape
    cout = w.shape[2]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((nz, ny, nx, cout))
    for t, (dz, dy, dx) in enumerate(TAPS):
        y += _shifted(xp, dz, dy, dx, (nz, ny, nx)) @ w[t]
    if b is not None:
        y += b
    return y


def conv3d_dgrad(dy, w):
    """dx[v, ci] = sum_{tap, co} w[tap, ci, co] * dy[v - off(tap), co]."""
    nz, ny, nx, cout = dy.shape
    cin = w.shape[1]
    dp = np.pad(dy, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dx_ = np.zeros((nz, ny, nx, cin))
    for t, (dz, dyy, ddx) in enumerate(TAPS):
        dx_ += _shifted(dp, -dz, -dyy, -ddx, (nz, ny, nx)) @ w[t].T
    return dx_


def conv3d_wgrad(x, dy):
    """(dw[tap, ci, co] = sum_v x[v + off(tap), ci] dy[v, co], db[co] = sum_v dy[v, co])."""
    nz, ny, nx, cin = x.shape
    cout = dy.shape[3]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    d2 = dy.reshape(-1, cout)
    dw = np.zeros((27, cin, cout))
    for t, (dz, dyy, ddx) in enumerate(TAPS):
        dw[t] = _shifted(xp, dz, dyy, ddx, (nz, ny, nx)).reshape(-1, cin).T @ d2
    return dw, d2.sum(axis=0)


def pool(f):
    """2x2x2 average of a fine grid (nz, ny, nx[, C]) with even dims."""
    nz, ny, nx = f.shape[:3]
    return f.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, *f.shape[3:]).mean(axis=(1, 3, 5))


def pool_bwd(g):
    """Adjoint of pool: each child receives 1/8 of its parent's gradient."""
    return upsample(g) / 8.0


def upsample(c):
    """Nearest-neighbour 2x upsampling (each coarse value copied to its 8 children)."""
    return c.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def upsample_bwd(g):
    """Adjoint of upsample: the sum over the 8 children."""
    return pool(g) * 8.0


def features(r, mask, cx, cy, cz, inv_scale):
    """8-channel network input on the half grid: [R(r)*inv_scale, R(mask), R(cx), R(cy), R(cz), 0, 0, 0]."""
    out = np.zeros(tuple(s // 2 for s in r.shape) + (8,))
    out[..., 0] = pool(r) * inv_scale
    out[..., 1] = pool(mask)
    out[..., 2] = pool(cx)
    out[..., 3] = pool(cy)
    out[..., 4] = pool(cz)
    return out


def stencil(cx, cy, cz, x):
    """y = x - sum over faces of c * x_neighbour: the unit-diagonal 7-point operator whose
    low-face couplings sit at the high cell of each face (cx[v] couples v and v - 1 in x)."""
    y = x.copy()
    y[:, :, 1:] -= cx[:, :, 1:] * x[:, :, :-1]
    y[:, :, :-1] -= cx[:, :, 1:] * x[:, :, 1:]
    y[:, 1:, :] -= cy[:, 1:, :] * x[:, :-1, :]
    y[:, :-1, :] -= cy[:, 1:, :] * x[:, 1:, :]
    y[1:, :, :] -= cz[1:, :, :] * x[:-1, :, :]
    y[:-1, :, :] -= cz[1:, :, :] * x[1:, :, :]
    return y


def rms_over_mask(r, mask):
    cnt = float(mask.sum())
    return float(np.sqrt((r.astype(np.float64) ** 2).sum() / cnt)) if cnt > 0 else 1.0


def forward(p, feat, bf16=False):
    """Network forward on the 8-channel features; returns (u, cache) with u the head's
    channel 0 on the half grid."""
    P = split_params(np.asarray(p, dtype=np.float64))
    W = [(_q(w, bf16), b) for w, b in P]
    relu = lambda a: np.maximum(a, 0.0)
    c = {"feat": _q(feat, bf16)}
    c["enc0"] = _q(relu(conv3d(c["feat"], *W[0])), bf16)
    c["p1"] = _q(pool(c["enc0"]), bf16)
    c["enc1"] = _q(relu(conv3d(c["p1"], *W[1])), bf16)
    c["p2"] = _q(pool(c["enc1"]), bf16)
    c["enc2"] = _q(relu(conv3d(c["p2"], *W[2])), bf16)
    c["p3"] = _q(pool(c["enc2"]), bf16)
    c["bott"] = _q(relu(conv3d(c["p3"], *W[3])), bf16)
    c["cat2"] = np.concatenate([upsample(c["bott"]), c["enc2"]], axis=3)
    c["dec2"] = _q(relu(conv3d(c["cat2"], *W[4])), bf16)
    c["cat1"] = np.concatenate([upsample(c["dec2"]), c["enc1"]], axis=3)
    c["dec1"] = _q(relu(conv3d(c["cat1"], *W[5])), bf16)
    c["cat0"] = np.concatenate([upsample(c["dec1"]), c["enc0"]], axis=3)
    c["dec0"] = _q(relu(conv3d(c["cat0"], *W[6])), bf16)
    u = conv3d(c["dec0"], *W[7])[..., 0]
    c["W"] = W
    return u, c


def backward(c, du, bf16=False):
    """Gradient of the flat parameter vector given dL/du (half grid)."""
    W = c["W"]
    grads = [None] * 8

    def layer(l, x, dz, need_dx=True):
        dw, db = conv3d_wgrad(x, dz)
        grads[l] = (dw, db)
        return _q(conv3d_dgrad(dz, W[l][0]), bf16) if need_dx else None

    dout = np.zeros(du.shape + (16,))
    dout[..., 0] = _q(du, bf16)
    d = layer(7, c["dec0"], dout) * (c["dec0"] > 0)
    dcat0 = layer(6, c["cat0"], d)
    dd1 = _q(upsample_bwd(dcat0[..., :32]), bf16) * (c["dec1"] > 0)
    de0 = dcat0[..., 32:]
    dcat1 = layer(5, c["cat1"], dd1)
    dd2 = _q(upsample_bwd(dcat1[..., :64]), bf16) * (c["dec2"] > 0)
    de1 = dcat1[..., 64:]
    dcat2 = layer(4, c["cat2"], dd2)
    dbott = _q(upsample_bwd(dcat2[..., :128]), bf16) * (c["bott"] > 0)
    de2 = dcat2[..., 128:]
    dp3 = layer(3, c["p3"], dbott)
    de2 = _q(de2 + _q(pool_bwd(dp3), bf16), bf16) * (c["enc2"] > 0)
    dp2 = layer(2, c["p2"], de2)
    de1 = _q(de1 + _q(pool_bwd(dp2), bf16), bf16) * (c["enc1"] > 0)
    dp1 = layer(1, c["p1"], de1)
    de0 = _q(de0 + _q(pool_bwd(dp1), bf16), bf16) * (c["enc0"] > 0)
    layer(0, c["feat"], de0, need_dx=False)
    return np.concatenate([np.concatenate([dw.ravel(), db]) for dw, db in grads])


SMOOTH_STEPS = 2  # weighted-Jacobi steps after the coarse correction


def _taps_1d(n_fine):
    """Per fine index x: the two coarse taps (x >> 1 with weight 3/4, and its clamped
    neighbour on x's side with weight 1/4) of cell-centred linear interpolation."""
    x = np.arange(n_fine)
    c = n_fine // 2
    t0 = x >> 1
    t1 = np.clip(t0 + np.where(x & 1, 1, -1), 0, c - 1)
    return t0, t1


def prolong(u):
    """Cell-centred trilinear prolongation of a coarse grid (nz, ny, nx) to 2x, with
    clamped indexing at the box edges (a constant stays constant)."""
    nz, ny, nx = u.shape
    out = np.zeros((2 * nz, 2 * ny, 2 * nx))
    tz, ty, tx = _taps_1d(2 * nz), _taps_1d(2 * ny), _taps_1d(2 * nx)
    for a, wa in ((0, 0.75), (1, 0.25)):
        for b, wb in ((0, 0.75), (1, 0.25)):
            for c, wc in ((0, 0.75), (1, 0.25)):
                out += wa * wb * wc * u[np.ix_(tz[a], ty[b], tx[c])]
    return out


def prolong_T(g):
    """Adjoint of prolong (the fine-to-coarse transpose)."""
    nz, ny, nx = (s // 2 for s in g.shape)
    out = np.zeros((nz, ny, nx))
    tz, ty, tx = _taps_1d(2 * nz), _taps_1d(2 * ny), _taps_1d(2 * nx)
    for a, wa in ((0, 0.75), (1, 0.25)):
        for b, wb in ((0, 0.75), (1, 0.25)):
            for c, wc in ((0, 0.75), (1, 0.25)):
                np.add.at(out, np.ix_(tz[a], ty[b], tx[c]), wa * wb * wc * g)
    return out


def precondition(p, r, mask, cx, cy, cz, omega, bf16=False, rms=None):
    """z = M(r): z0 = mask * s * P U(R r / s, R mask, R cx, R cy, R cz), then SMOOTH_STEPS
    steps z += omega * mask * (r - A z); s = RMS of r over the mask (or the given rms).
    Returns (z, cache, s)."""
    s = rms_over_mask(r, mask) if rms is None else rms
    s = s if s > 0 else 1.0
    u, cache = forward(p, features(r, mask, cx, cy, cz, 1.0 / s), bf16)
    z = mask * (s * prolong(u))
    for _ in range(SMOOTH_STEPS):
        z = z + omega * mask * (r - stencil(cx, cy, cz, z))
    return z, cache, s


def loss_and_grad(p, r, mask, cx, cy, cz, xstar, omega, bf16=False, need_grad=True):
    """L = (z - x*)' A (z - x*) / (x*' r) and dL/dp (the training sample of the library)."""
    z, cache, s = precondition(p, r, mask, cx, cy, cz, omega, bf16)
    e = z - xstar
    ae = stencil(cx, cy, cz, e)
    den = float((xstar * r).sum())
    loss = float((e * ae).sum()) / den
    if not need_grad:
        return loss, None, z
    gz = 2.0 * ae / den
    for _ in range(SMOOTH_STEPS):  # z_{i+1} = z_i + w M (r - A z_i):  g_i = g_{i+1} - w A (M g_{i+1})
        gz = gz - omega * stencil(cx, cy, cz, mask * gz)
    du = s * prolong_T(mask * gz)
    return loss, backward(cache, du, bf16), z
