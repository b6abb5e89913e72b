"""The learned preconditioner's HIP kernels against the FP64 oracle (oracle/unet_oracle.py):
BF16 MFMA convolution forward / input gradient / weight gradient, the whole
network's loss and parameter gradients, Adam, weights files, and flexible CG
inside shk_project (converging to the same residual as the plain solver)."""
from __future__ import annotations

import numpy as np
import pytest

from gpu_helpers import BAR_STOP, assert_pressure_bar, log, oracle_truth, run_projection
from oracle import unet_oracle as U
from unet_cases import init_params, training_sample

pytestmark = pytest.mark.gpu


def _dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def _bf16_tensor(a):
    import torch
    return _dev(np.asarray(a, dtype=np.float32), torch.bfloat16)


def _host(t):
    import torch
    return t.to(torch.float64).cpu().numpy()


CONV_CASES = [  # (nz, ny, nx), cin, cout, relu
    ((5, 6, 7), 8, 16, True),
    ((8, 8, 8), 16, 32, True),
    ((4, 9, 6), 48, 16, False),
    ((6, 4, 8), 64, 128, True),
    ((4, 4, 4), 192, 64, True),
    ((3, 5, 4), 32, 48, False),
]


@pytest.mark.parametrize("shape,cin,cout,relu", CONV_CASES)
def test_conv3d_forward(shape, cin, cout, relu):
    from shiokaze_amd import learned
    log(f"[gpu] conv3d forward {shape} {cin}->{cout} relu={relu}")
    rng = np.random.default_rng(cin * 1000 + cout)
    x = U.bf16_round(rng.standard_normal(shape + (cin,)))
    w = rng.standard_normal((27, cin, cout)) * np.sqrt(2.0 / (27 * cin))
    b = rng.standard_normal(cout) * 0.1
    y = _host(learned.conv3d(_bf16_tensor(x), _dev(w.astype(np.float32)), _dev(b.astype(np.float32)), relu))
    wq = U.bf16_round(w)
    ref = U.conv3d(x, wq, b.astype(np.float32).astype(np.float64))
    if relu:
        ref = np.maximum(ref, 0.0)
    scale = U.conv3d(np.abs(x), np.abs(wq))
    # BF16 output rounding (2^-9 relative) plus FP32 accumulation over 27*cin terms
    tol = 2.0 ** -8 * np.abs(ref) + 1e-5 * scale + 1e-30
    err = np.abs(y - ref)
    log(f"[gpu]   max err/tol = {np.max(err / tol):.3f}")
    assert np.all(err <= tol)


CAT_CASES = [  # (nz, ny, nx), c1 (upsampled coarse channels), cin, cout, relu
    ((10, 20, 40), 0, 8, 16, True),      # partial tiles in x, y and z (tiles of 16 x 8 x 4)
    ((12, 16, 32), 0, 16, 32, True),
    ((8, 24, 48), 32, 48, 16, True),     # the last decoder layer: 32 + 16 channels
    ((6, 10, 34), 64, 96, 32, False),    # 64 + 32 channels; the last stage starts in the skip input
    ((4, 8, 16), 0, 24, 16, False),      # one 24-channel stage (3 groups: 9 entries in 3 chunks)
    ((8, 8, 16), 0, 40, 32, True),       # stages of 32 and 8 channels
]


# the direct-load kernel on the network's small-grid layers and on ragged shapes:
# cout of 48 (3 output tiles), 64 and 128; stages of 8, 16, 24 and 32 channels; the
# last N-tile partly outside the grid
DIRECT_CASES = [  # (nz, ny, nx), c1, cin, cout, relu
    ((16, 16, 16), 0, 16, 32, True),     # encoder level 1 at 16^3
    ((8, 8, 8), 0, 32, 64, True),        # encoder level 2 at 8^3
    ((8, 8, 8), 0, 64, 128, True),       # bottleneck at 8^3
    ((4, 4, 4), 0, 64, 128, True),       # bottleneck at 4^3 (4 N-tiles)
    ((16, 16, 16), 128, 192, 64, True),  # decoder level 2 at 16^3
    ((32, 32, 32), 64, 96, 32, True),    # decoder level 1 at 32^3
    ((5, 6, 7), 0, 8, 16, True),         # 210 voxels: a partial N-tile
    ((3, 5, 4), 0, 32, 48, False),
    ((4, 9, 6), 0, 48, 16, False),       # stages of 24 channels
    ((8, 8, 16), 0, 40, 32, True),       # stages of 8 channels
]


@pytest.mark.parametrize("shape,c1,cin,cout,relu", CAT_CASES[:4] + DIRECT_CASES)
def test_conv3d_direct_kernel(shape, c1, cin, cout, relu):
    """The direct-load kernel (K split over the waves of a block, partial sums added in
    LDS) against the oracle, on concatenated and plain inputs."""
    _check_conv_cat(shape, c1, cin, cout, relu, 3)


@pytest.mark.parametrize("shape,c1,cin,cout,relu", CAT_CASES)
@pytest.mark.parametrize("kernel", [1, 2])
def test_conv3d_concatenated_input(shape, c1, cin, cout, relu, kernel):
    """Both convolution kernels (the tiled one and the row-reuse one) on the decoder's
    input [nearest upsampling of a half-size grid, skip] against the oracle's
    convolution of the explicitly concatenated FP64 array."""
    _check_conv_cat(shape, c1, cin, cout, relu, kernel)


def test_conv3d_rows_several_tiles_per_block():
    """The 16-output row-reuse kernel is persistent (two blocks per CU): 1024 tiles of
    16 x 4 x 8 voxels give every block several tiles and both stages of each in turn."""
    _check_conv_cat((64, 64, 128), 32, 48, 16, True, 2)


def _check_conv_cat(shape, c1, cin, cout, relu, kernel):
    from shiokaze_amd import learned
    log(f"[gpu] conv3d cat {shape} c1={c1} {cin}->{cout} relu={relu} kernel={kernel}")
    rng = np.random.default_rng(cin * 100 + cout + c1)
    nz, ny, nx = shape
    skip = U.bf16_round(rng.standard_normal(shape + (cin - c1,)))
    coarse = U.bf16_round(rng.standard_normal((nz // 2, ny // 2, nx // 2, c1))) if c1 else None
    w = rng.standard_normal((27, cin, cout)) * np.sqrt(2.0 / (27 * cin))
    b = rng.standard_normal(cout) * 0.1
    y = _host(learned.conv3d_cat(None if coarse is None else _bf16_tensor(coarse), _bf16_tensor(skip),
                                 _dev(w.astype(np.float32)), _dev(b.astype(np.float32)), relu, kernel))
    if c1:
        up = coarse.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)
        x = np.concatenate([up, skip], axis=3)
    else:
        x = skip
    wq = U.bf16_round(w)
    ref = U.conv3d(x, wq, b.astype(np.float32).astype(np.float64))
    if relu:
        ref = np.maximum(ref, 0.0)
    scale = U.conv3d(np.abs(x), np.abs(wq))
    tol = 2.0 ** -8 * np.abs(ref) + 1e-5 * scale + 1e-30
    err = np.abs(y - ref)
    log(f"[gpu]   max err/tol = {np.max(err / tol):.3f}")
    assert np.all(err <= tol)


def test_conv3d_kernel_choice():
    """Kernel 0 takes the row-reuse kernel on a grid of >= 512 tiles (bit-identical to
    kernel 2), the direct-load kernel where the tiled kernel would have fewer than 1024
    blocks (bit-identical to kernel 3) and the tiled kernel otherwise (kernel 1)."""
    import torch
    from shiokaze_amd import learned
    rng = np.random.default_rng(7)
    for shape, cin, pick in (((64, 64, 64), 16, 2), ((32, 32, 32), 16, 3), ((64, 64, 64), 8, 1)):
        w = _dev((rng.standard_normal((27, cin, 16)) * 0.1).astype(np.float32))
        x = _bf16_tensor(rng.standard_normal(shape + (cin,)))
        y0 = learned.conv3d_cat(None, x, w, None, True, 0)
        yp = learned.conv3d_cat(None, x, w, None, True, pick)
        assert torch.equal(y0, yp), shape


@pytest.mark.parametrize("shape,cin,cout", [((5, 6, 7), 16, 16), ((4, 5, 6), 32, 64), ((4, 4, 4), 192, 64),
                                            ((3, 4, 5), 8, 16), ((4, 4, 3), 96, 32)])
def test_conv3d_backward(shape, cin, cout):
    from shiokaze_amd import learned
    log(f"[gpu] conv3d backward {shape} {cin}->{cout}")
    rng = np.random.default_rng(cin * 7 + cout)
    x = U.bf16_round(rng.standard_normal(shape + (cin,)))
    w = rng.standard_normal((27, cin, cout)) * np.sqrt(2.0 / (27 * cin))
    dz = U.bf16_round(rng.standard_normal(shape + (cout,)))
    need_dx = cin % 16 == 0
    dx, dw, db = learned.conv3d_backward(_bf16_tensor(x), _dev(w.astype(np.float32)), _bf16_tensor(dz), need_dx)
    wq = U.bf16_round(w)
    if need_dx:
        ref = U.conv3d_dgrad(dz, wq)
        scale = U.conv3d_dgrad(np.abs(dz), np.abs(wq))
        err = np.abs(_host(dx) - ref)
        tol = 2.0 ** -8 * np.abs(ref) + 1e-5 * scale + 1e-30
        log(f"[gpu]   dx max err/tol = {np.max(err / tol):.3f}")
        assert np.all(err <= tol)
    rdw, rdb = U.conv3d_wgrad(x, dz)
    sdw, sdb = U.conv3d_wgrad(np.abs(x), np.abs(dz))
    edw = np.abs(dw.cpu().numpy().astype(np.float64) - rdw)
    edb = np.abs(db.cpu().numpy().astype(np.float64) - rdb)
    log(f"[gpu]   dw max err/scale = {np.max(edw / sdw):.2e}, db {np.max(edb / sdb):.2e}")
    assert np.all(edw <= 1e-5 * sdw + 1e-30)
    assert np.all(edb <= 1e-5 * sdb + 1e-30)


def _sample_dev(shape, seed):
    r, mask, cx, cy, cz, xstar = training_sample(shape, seed)
    f32 = [a.astype(np.float32) for a in (r, mask, cx, cy, cz, xstar)]
    return f32, [_dev(a) for a in f32]


def test_network_loss_and_gradients():
    """The GPU network's loss and parameter gradients against the FP64 oracle with BF16
    rounding emulated where the GPU stores BF16 (activations and their gradients).
    Error model: each stored BF16 value carries a relative rounding of at most 2^-9; the
    FP32 MFMA accumulation over up to 27*128 terms adds ~1e-6 relative; so a layer's
    gradient, an FP32 sum over the voxels of BF16 products whose rounding the oracle
    reproduces only at the stored points, differs by O(1e-3) relative L2 (observed
    <= 1.2e-3).  The per-layer bound is 5e-3."""
    from shiokaze_amd import UNet
    log("[gpu] network loss/gradients 32^3")
    (r, mask, cx, cy, cz, xs), dev = _sample_dev((32, 32, 32), 5)
    p = init_params(6).astype(np.float32)
    with UNet(seed=1) as net:
        assert net.num_params == U.num_params()
        net.set_params(p)
        np.testing.assert_array_equal(net.get_params(), p)
        loss = net.train_step(*dev, omega=0.8, lr=0.0)
        g = net.get_params(grads=True).astype(np.float64)
    rl, rg, _ = U.loss_and_grad(p.astype(np.float64), *[a.astype(np.float64) for a in (r, mask, cx, cy, cz, xs)],
                                0.8, bf16=True)
    log(f"[gpu]   loss gpu {loss:.6f} oracle {rl:.6f}")
    assert loss == pytest.approx(rl, rel=1e-2)
    cos = float(g @ rg / (np.linalg.norm(g) * np.linalg.norm(rg)))
    log(f"[gpu]   gradient cosine {cos:.6f}")
    for l, (wo, bo) in enumerate(U.layer_offsets()[0]):
        cout = U.LAYERS[l][1]
        seg = slice(wo, bo + cout)
        rel = np.linalg.norm(g[seg] - rg[seg]) / np.linalg.norm(rg[seg])
        log(f"[gpu]   layer {l}: rel L2 grad error {rel:.3e}")
        assert rel < 5e-3
    assert cos > 0.998


def test_apply_matches_oracle_and_is_homogeneous():
    from shiokaze_amd import UNet
    log("[gpu] apply 32x48x16")
    (r, mask, cx, cy, cz, xs), dev = _sample_dev((16, 48, 32), 8)
    p = init_params(9).astype(np.float32)
    with UNet() as net:
        net.set_params(p)
        z = net.apply(*dev[:5], omega=0.5).cpu().numpy().astype(np.float64)
        z2 = net.apply(dev[0] * 4.0, *dev[1:5], omega=0.5).cpu().numpy().astype(np.float64)
    rz, _, _ = U.precondition(p.astype(np.float64), *[a.astype(np.float64) for a in (r, mask, cx, cy, cz)], 0.5,
                              bf16=True)
    rel = np.linalg.norm(z - rz) / np.linalg.norm(rz)
    log(f"[gpu]   rel L2 error {rel:.3e}")
    # BF16 storage emulated in the oracle; the remaining FP32 accumulation-order error
    # is ~1e-5 (observed <= 1e-4)
    assert rel < 1e-3
    assert np.all(z[mask == 0] == 0)
    # the RMS normalisation makes M positively homogeneous (exactly, up to FP32 scaling)
    np.testing.assert_allclose(z2, 4.0 * z, rtol=1e-5, atol=1e-6 * np.abs(z).max())


def test_unet_smoothing_setting():
    """shk_unet_set_smoothing: the creation default (two omega steps) is what apply()
    does without it; more steps or Chebyshev weights change M; bad values are refused."""
    import torch
    from shiokaze_amd import UNet, _lib
    rng = np.random.default_rng(11)
    shape = (16, 16, 16)
    mask = _dev(np.ones(shape, dtype=np.float32))
    hx, hy, hz = (_dev((rng.random(shape) * 0.15).astype(np.float32)) for _ in range(3))
    r = _dev(rng.standard_normal(shape).astype(np.float32))
    with UNet(seed=3) as net:
        z0 = net.apply(r, mask, hx, hy, hz, omega=0.8)
        net.set_smoothing(2, 0.0)
        assert torch.equal(net.apply(r, mask, hx, hy, hz, omega=0.8), z0)
        net.set_smoothing(10, 16.0)
        z10 = net.apply(r, mask, hx, hy, hz, omega=0.8)
        assert torch.isfinite(z10).all() and not torch.equal(z10, z0)
        for steps, cheb in ((0, 0.0), (65, 0.0), (4, 0.5), (4, -1.0)):
            with pytest.raises(_lib.ShkError):
                net.set_smoothing(steps, cheb)


def test_residual_kernel():
    from shiokaze_amd import learned
    (r, mask, cx, cy, cz, xs), dev = _sample_dev((16, 24, 32), 12)
    x = (np.random.default_rng(1).standard_normal(r.shape) * mask).astype(np.float32)
    res = learned.residual(dev[2], dev[3], dev[4], _dev(x), dev[0]).cpu().numpy().astype(np.float64)
    ref = r.astype(np.float64) - U.stencil(cx.astype(np.float64), cy.astype(np.float64), cz.astype(np.float64),
                                           x.astype(np.float64))
    np.testing.assert_allclose(res, ref, rtol=0, atol=1e-5 * np.abs(ref).max())


def test_adam_step_and_weights_file(tmp_path):
    from shiokaze_amd import ShkError, UNet
    log("[gpu] adam + save/load")
    _, dev = _sample_dev((16, 16, 16), 3)
    p0 = init_params(4).astype(np.float32)
    lr = 1e-3
    with UNet() as net:
        net.set_params(p0)
        net.train_step(*dev, omega=1.0, lr=0.0)
        g = net.get_params(grads=True)
        net.train_step(*dev, omega=1.0, lr=lr)   # accumulates g again, then Adam step 1
        p1 = net.get_params()
        assert np.all(net.get_params(grads=True) == 0)
        path = tmp_path / "w.bin"
        net.save(path)
        with UNet(seed=99) as other:
            assert not np.array_equal(other.get_params(), p1)
            other.load(path)
            np.testing.assert_array_equal(other.get_params(), p1)
        bad = tmp_path / "bad.bin"
        bad.write_bytes(b"not a weights file")
        with pytest.raises(ShkError):
            net.load(bad)
    # first Adam step from zero moments (m^ = G, v^ = G^2): p1 = p0 - lr * G / (|G| + eps), with
    # G = the gradient accumulated twice (~ 2g) and eps = 1e-8
    big = np.abs(g) > 1e-3 * np.abs(g).max()
    G = 2.0 * g[big].astype(np.float64)
    step = (p0 - p1)[big].astype(np.float64)
    np.testing.assert_allclose(step, lr * G / (np.abs(G) + 1e-8), rtol=2e-4, atol=2e-9)


def _scene_config(weights):
    return {"Residual": 1e-6, "Preconditioner": "unet", "UnetWeights": str(weights)}


@pytest.mark.parametrize("kind,n,seed", [(1, 32, 0), (3, (24, 40, 16), 3), (3, 48, 5)])
def test_projection_with_learned_preconditioner(tmp_path, scenes, oracle, kind, n, seed):
    """Flexible CG with the (untrained) network converges to the residual asked for,
    falling back to Jacobi-PCG if it stalls; at the bar stop its pressure meets the
    pressure bar against the oracle's FP64 solve to 1e-12."""
    from shiokaze_amd import UNet
    log(f"[gpu] unet projection kind={kind} n={n} seed={seed}")
    path = tmp_path / "net.bin"
    with UNet(seed=seed) as net:
        net.save(path)
    sc = scenes(kind, n, seed)
    g = run_projection(sc, _scene_config(path), want_system=False)
    r = g["result"]
    log(f"[gpu]   status={r['status_name']} count={r['count']} outer={r['outer']} fallback={r['fallback']} "
        f"relres={r['true_relres_inf']:.2e}")
    assert r["status"] == 0
    assert r["true_relres_inf"] <= 1e-6
    gb = run_projection(sc, dict(_scene_config(path), Residual=BAR_STOP["unet"]))
    assert gb["result"]["status"] == 0
    assert gb["div_after_fp64"] / gb["div_before"] <= 1e-6
    truth = oracle_truth(oracle, sc, ("3d", str(kind), str(n), str(seed)))
    assert_pressure_bar(gb["system"]["x"], truth, f"untrained unet kind={kind} n={n} at {BAR_STOP['unet']:g}")


def test_unet_config_errors(scenes, tmp_path):
    import shiokaze_amd
    from gpu_helpers import to_device
    sc = scenes(1, 16, 0)
    d = to_device(sc)
    with shiokaze_amd.MacPressureSolver3(sc["shape"], sc["dx"], config={"Preconditioner": "unet"}) as ps:
        with pytest.raises(shiokaze_amd.ShkError):
            ps.project(1.0, (d["u"], d["v"], d["w"]), d["solid"], (d["us"], d["vs"], d["ws"]), d["fluid"])
    with shiokaze_amd.MacPressureSolver3(sc["shape"], sc["dx"]) as ps:
        for key, val in (("UnetSmoothing", "0"), ("UnetSmoothing", "65"), ("UnetChebyshev", "0.5"),
                         ("UnetChebyshev", "-2"), ("UnetChebyshev", "x")):
            with pytest.raises(shiokaze_amd.ShkError):
                ps.configure(key, val)
        ps.configure("UnetSmoothing", "3")
        ps.configure("UnetChebyshev", "0")
    missing = tmp_path / "missing.bin"
    with shiokaze_amd.MacPressureSolver3(sc["shape"], sc["dx"], config=_scene_config(missing)) as ps:
        with pytest.raises(shiokaze_amd.ShkError):
            ps.project(1.0, (d["u"], d["v"], d["w"]), d["solid"], (d["us"], d["vs"], d["ws"]), d["fluid"])


def test_export_scaled_matches_system(scenes):
    """b^ = D^-1/2 b, couplings c / sqrt(d_i d_j), x^ = D^1/2 x from the exported system."""
    import shiokaze_amd
    from gpu_helpers import to_device
    sc = scenes(3, 32, 2)
    d = to_device(sc)
    with shiokaze_amd.MacPressureSolver3(sc["shape"], sc["dx"], config={"Residual": 1e-8}) as ps:
        ps.project(1.0, (d["u"], d["v"], d["w"]), d["solid"], (d["us"], d["vs"], d["ws"]), d["fluid"])
        s = {k: v.cpu().numpy().astype(np.float64) for k, v in ps.export_scaled().items()}
        sy = {k: v.cpu().numpy() for k, v in ps.export_system().items()}
    lab = sy["label"] != 0
    np.testing.assert_array_equal(s["mask"], lab.astype(np.float64))
    dg = np.where(lab, sy["diag"], 1.0)
    sig = np.where(lab, 1.0 / np.sqrt(dg), 0.0)
    np.testing.assert_allclose(s["rhat"], (sig * sy["rhs"]).astype(np.float32), rtol=1e-6, atol=0)
    np.testing.assert_allclose(s["xhat"], np.where(lab, sy["x"] / np.where(lab, sig, 1.0), 0.0).astype(np.float32),
                               rtol=1e-6, atol=0)
    for key, ax in (("hx", 2), ("hy", 1), ("hz", 0)):
        lo = np.roll(sig, 1, axis=ax)
        sl = [slice(None)] * 3
        sl[ax] = 0
        lo[tuple(sl)] = 0.0
        c = sy["c" + key[1]]
        np.testing.assert_allclose(s[key], (c * sig * lo).astype(np.float32), rtol=1e-6, atol=1e-30)


WEIGHTS = __import__("pathlib").Path(__file__).resolve().parents[1] / "shiokaze_amd" / "weights" / \
    "unet_irregular_128.bin"


def test_trained_network_gain(scenes, oracle, tmp_path):
    """What training buys: on held-out irregular scenes at 128^3 (seeds after the
    training range), flexible CG with the trained network needs a fraction of the
    iterations of Jacobi-PCG and of the same network untrained, never falls back,
    and reaches the same residual and pressure.  The solver's smoothing after each
    application (10 Chebyshev-weighted steps) takes fewer iterations than the
    training-time two damped-Jacobi steps."""
    from shiokaze_amd import UNet
    untrained = tmp_path / "untrained.bin"
    with UNet(seed=0) as net:
        net.save(untrained)
    rows = []
    for seed in (2000, 2001, 2002, 2003):
        sc = scenes(3, 128, seed)
        plain = run_projection(sc, {"Residual": 1e-6, "Precision": "mixed", "Preconditioner": "jacobi"},
                               want_system=False)
        tr = run_projection(sc, _scene_config(WEIGHTS), want_system=False)
        un = run_projection(sc, _scene_config(untrained), want_system=False)
        t2 = run_projection(sc, dict(_scene_config(WEIGHTS), UnetSmoothing=2, UnetChebyshev=0), want_system=False)
        p, t, u, s2 = plain["result"], tr["result"], un["result"], t2["result"]
        log(f"[gpu] gain seed {seed}: jacobi {p['count']} it, trained {t['count']} it (fallback {t['fallback']}), "
            f"untrained {u['count']} it (fallback {u['fallback']}), trained with 2 plain steps {s2['count']} it")
        for r in (p, t, u, s2):
            assert r["status"] == 0 and r["true_relres_inf"] <= 1e-6
        assert t["fallback"] == 0 and s2["fallback"] == 0
        rows.append((p["count"], t["count"], u["count"], s2["count"]))
        if seed == 2000:
            # the trained network at the bar stop against the oracle's FP64 truth
            gb = run_projection(sc, dict(_scene_config(WEIGHTS), Residual=BAR_STOP["unet"]))
            assert gb["result"]["status"] == 0 and gb["result"]["fallback"] == 0
            truth = oracle_truth(oracle, sc, ("3d", "3", "128", str(seed)))
            assert_pressure_bar(gb["system"]["x"], truth, f"trained unet 128^3 seed {seed} at {BAR_STOP['unet']:g}")
    jac, trained, raw, two = (np.mean([r[i] for r in rows]) for i in range(4))
    log(f"[gpu] gain: mean iterations jacobi {jac:.1f}, trained {trained:.1f}, untrained {raw:.1f}, "
        f"trained with 2 plain steps {two:.1f}")
    assert trained <= 0.25 * jac
    assert trained <= 0.5 * raw
    assert trained <= 0.5 * two


@pytest.mark.parametrize("kind,n,seed", [(3, 32, 2), (3, 48, 7)])
def test_brick_preconditioner_matches_dense_apply(scenes, kind, n, seed):
    """Inside flexible CG the preconditioner runs on the projection's bricks: features
    from brick vectors, prolongation into the bricks and brick smoothing.  One
    application (shk_unet_precondition) equals the dense network application
    (shk_unet_apply) on the exported scaled system: bit for bit off the unknowns (zero)
    and to the input scale's last-bit rounding on them (the two sum r.r in different
    orders)."""
    import torch

    import shiokaze_amd
    from gpu_helpers import to_device
    from shiokaze_amd import UNet
    raw = WEIGHTS.read_bytes()   # magic, layer count, (cin, cout) per layer, nparam, omega, weights
    nl = int(np.frombuffer(raw[8:12], dtype=np.int32)[0])
    at = 12 + 8 * nl + 8
    omega = float(np.frombuffer(raw[at:at + 8], dtype=np.float64)[0])
    sc = scenes(kind, n, seed)
    d = to_device(sc)
    # the dense application is the training-time network: two smoothing steps with the
    # weights file's omega
    cfg = dict(_scene_config(WEIGHTS), UnetSmoothing=2, UnetChebyshev=0)
    with shiokaze_amd.MacPressureSolver3(sc["shape"], sc["dx"], config=cfg) as ps:
        ps.project(1.0, (d["u"], d["v"], d["w"]), d["solid"], (d["us"], d["vs"], d["ws"]), d["fluid"])
        s = ps.export_scaled()
        g = torch.Generator(device="cuda").manual_seed(seed)
        r = (torch.randn(s["rhat"].shape, device="cuda", generator=g) * s["mask"]).contiguous()
        zb = ps.unet_precondition(r).cpu().numpy()
    with UNet(seed=0) as net:
        net.load(WEIGHTS)
        zd = net.apply(r, s["mask"], s["hx"], s["hy"], s["hz"], omega=omega).cpu().numpy()
    lab = s["mask"].cpu().numpy() > 0
    err = np.abs(zb - zd).max() / np.abs(zd).max()
    log(f"[gpu] brick vs dense preconditioner kind={kind} n={n}: max rel diff {err:.2e}, omega {omega}")
    assert np.all(zb[~lab] == 0) and np.all(zd[~lab] == 0)
    assert err <= 2e-3
