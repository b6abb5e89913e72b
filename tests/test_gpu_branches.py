"""GPU parity of project()'s branches (macpressuresolver3.cpp:75-124, 220-260,
287-293) through the C-ABI against the CPU oracle: surface tension, solid density,
first-order switches, volume correction and warm start.

Bit-exact: fractions, system (diag, couplings, rhs), the surface-tension velocity,
the volume-correction constant, and the velocity update given the same pressure.
Within tolerance: solutions, and the warm-start rhs (A p_prev summed in another order)."""
from __future__ import annotations

import numpy as np
import pytest

from gpu_helpers import log, oracle_dense_system, to_device

pytestmark = pytest.mark.gpu

DT = 1.0


def _drop_scene(n=32, radius=0.25):
    from oracle import oracle as O
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
    return sc


def _project(sc, config, surface_tension=0.0, density=None, volumes=None, repeat=1):
    """Runs shk_project `repeat` times on fresh copies of the scene's velocity in one
    context; returns per-call dicts of host arrays."""
    import torch

    from shiokaze_amd import MacPressureSolver3
    nx, ny, nz = sc["shape"]
    outs = []
    with MacPressureSolver3((nx, ny, nz), sc["dx"], config=config) as ps:
        dens = None
        if density is not None:
            dens = [torch.from_numpy(np.ascontiguousarray(d, dtype=np.float32)).cuda() for d in density]
            ps.set_solid_density(dens)
        for _ in range(repeat):
            d = to_device(sc)
            if volumes is not None:
                ps.set_target_volume(*volumes)
            res = ps.project(DT, (d["u"], d["v"], d["w"]), d["solid"], (d["us"], d["vs"], d["ws"]), d["fluid"],
                             surface_tension=surface_tension)
            out = {"result": res, "pressure": ps.get_pressure().cpu().numpy(),
                   "velocity": [d[k].cpu().numpy() for k in ("u", "v", "w")]}
            fr = ps.export_fractions()
            out["areas"] = [fr[f"area{a}"].cpu().numpy() for a in range(3)]
            out["rhos"] = [fr[f"rho{a}"].cpu().numpy() for a in range(3)]
            if res["unknowns"]:
                out["system"] = {k: v.cpu().numpy() for k, v in ps.export_system().items()}
                out["div_after_fp64"] = ps.divergence(DT, [torch.from_numpy(np.ascontiguousarray(sc[k])).cuda()
                                                           for k in ("u", "v", "w")],
                                                      (d["us"], d["vs"], d["ws"]), fp64_update=True)
            outs.append(out)
        torch.cuda.synchronize()
    return outs


def _assert_system_bitexact(oracle, g, o):
    lab = o["index_map"] >= 0
    s = g["system"]
    np.testing.assert_array_equal(s["label"].astype(bool), lab)
    ref = oracle_dense_system(oracle, o)
    np.testing.assert_array_equal(s["diag"][lab], ref["diag"][lab])
    np.testing.assert_array_equal(s["rhs"][lab], ref["rhs"][lab])
    both_x = lab[:, :, 1:] & lab[:, :, :-1]
    both_y = lab[:, 1:, :] & lab[:, :-1, :]
    both_z = lab[1:, :, :] & lab[:-1, :, :]
    np.testing.assert_array_equal(s["cx"][:, :, 1:][both_x], ref["cx"][:, :, 1:][both_x])
    np.testing.assert_array_equal(s["cy"][:, 1:, :][both_y], ref["cy"][:, 1:, :][both_y])
    np.testing.assert_array_equal(s["cz"][1:, :, :][both_z], ref["cz"][1:, :, :][both_z])


def _assert_fractions_bitexact(g, o):
    for a in range(3):
        np.testing.assert_array_equal(g["areas"][a].astype(np.float64), o["areas"][a])
        np.testing.assert_array_equal(g["rhos"][a].astype(np.float64), o["rhos"][a])


# ---------------------------------------------------------------- surface tension

@pytest.mark.parametrize("kind,n,seed", [(1, 24, 0), (3, (24, 40, 16), 3)])
def test_surface_tension_parity(scenes, oracle, kind, n, seed):
    sc = scenes(kind, n, seed)
    sigma = 0.05
    log(f"[gpu] surface tension kind={kind} n={n} seed={seed}")
    g = _project(sc, {"Residual": 1e-6, "Preconditioner": "none"}, surface_tension=sigma)[0]
    o = oracle.project(sc, tol=1e-6, mode=oracle.SHK_OR_CG, surface_tension_coef=sigma)
    _assert_fractions_bitexact(g, o)
    _assert_system_bitexact(oracle, g, o)
    # the velocity update from the GPU pressure, applied by the oracle to the
    # oracle's surface-tension velocity, is the GPU's velocity bit for bit
    vel = oracle.update_velocity(DT, sc["dx"], g["pressure"].astype(np.float64), o["areas"], o["rhos"],
                                 [sc["us"], sc["vs"], sc["ws"]], o["velocity_in"], real_is_double=False)
    for a in range(3):
        np.testing.assert_array_equal(g["velocity"][a].astype(np.float64), vel[a])
    assert g["result"]["true_relres_inf"] <= 1e-6


def test_surface_tension_laplace_law(oracle):
    """A static drop projects to the Laplace pressure 2*sigma/R with almost no velocity."""
    sigma, radius = 0.1, 0.25
    sc = _drop_scene(32, radius)
    g = _project(sc, {"Residual": 1e-8, "Preconditioner": "jacobi"}, surface_tension=sigma)[0]
    lab = g["system"]["label"].astype(bool)
    p = g["pressure"][lab].astype(np.float64)
    log(f"[gpu] drop: p mean {p.mean():.6f} (2 sigma/R = {2 * sigma / radius}), "
        f"max |u| after {max(np.abs(v).max() for v in g['velocity']):.3e}")
    np.testing.assert_allclose(p.mean(), 2 * sigma / radius, rtol=0.01)
    assert max(np.abs(v).max() for v in g["velocity"]) < 0.2


@pytest.mark.parametrize("host", [False, True])
def test_surface_tension_without_unknowns(oracle, host):
    """An isolated liquid cell has no unknowns; its faces still receive the jump."""
    n = 12
    sc = _drop_scene(n, 0.0)
    fluid = np.ones((n, n, n), np.float32)
    fluid[5, 6, 7] = -0.1
    sc["fluid"] = fluid
    sigma = 0.07
    o = oracle.project(sc, tol=1e-6, mode=oracle.SHK_OR_CG, surface_tension_coef=sigma)
    assert o["n"] == 0
    if host:
        import ctypes

        import shiokaze_amd
        hostf = {k: np.ascontiguousarray(sc[k]).copy() for k in ("fluid", "solid", "u", "v", "w", "us", "vs", "ws")}
        ps = shiokaze_amd.MacPressureSolver3((n, n, n), sc["dx"], config={"HostPointers": 1})
        res = shiokaze_amd.ShkResult()
        p = lambda arr: arr.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        assert ps.lib.shk_project(ps.ctx, DT, p(hostf["u"]), p(hostf["v"]), p(hostf["w"]), p(hostf["solid"]),
                                  p(hostf["us"]), p(hostf["vs"]), p(hostf["ws"]), p(hostf["fluid"]), sigma,
                                  ctypes.byref(res)) == 0
        ps.close()
        vel = [hostf[k] for k in ("u", "v", "w")]
        assert res.unknowns == 0
    else:
        g = _project(sc, {}, surface_tension=sigma)[0]
        assert g["result"]["unknowns"] == 0
        vel = g["velocity"]
    changed = 0
    for a in range(3):
        np.testing.assert_array_equal(vel[a].astype(np.float64), o["velocity_in"][a])
        changed += int(np.count_nonzero(vel[a]))
    assert changed == 6


# ---------------------------------------------------------------- solid density

def _density(sc, seed):
    from oracle import oracle as O
    nx, ny, nz = sc["shape"]
    rng = np.random.default_rng(seed)
    return [rng.uniform(0.0, 1.0, s).astype(np.float32) for s in O.face_shapes(nx, ny, nz)]


@pytest.mark.parametrize("kind,n,seed", [(3, 24, 3), (3, (24, 40, 16), 4)])
def test_solid_density_parity(scenes, oracle, kind, n, seed):
    sc = scenes(kind, n, seed)
    dens = _density(sc, seed)
    log(f"[gpu] solid density kind={kind} n={n} seed={seed}")
    g = _project(sc, {"Residual": 1e-6, "Preconditioner": "none"}, density=dens)[0]
    d64 = [d.astype(np.float64) for d in dens]
    o = oracle.project(sc, tol=1e-6, mode=oracle.SHK_OR_CG, solid_density=d64)
    _assert_fractions_bitexact(g, o)
    _assert_system_bitexact(oracle, g, o)
    vel = oracle.update_velocity(DT, sc["dx"], g["pressure"].astype(np.float64), o["areas"], o["rhos"],
                                 [sc["us"], sc["vs"], sc["ws"]], [sc["u"], sc["v"], sc["w"]], real_is_double=False)
    vel = oracle.density_scale(d64, o["areas"], vel, real_is_double=False)
    for a in range(3):
        np.testing.assert_array_equal(g["velocity"][a].astype(np.float64), vel[a])


def test_solid_density_host_pointers_and_off(scenes):
    """Host density arrays (staged by the library) give the device result; switching
    the density off restores the plain projection."""
    import ctypes

    import shiokaze_amd
    sc = scenes(3, 24, 3)
    dens = _density(sc, 3)
    cfg = {"Residual": 1e-6, "Preconditioner": "jacobi"}
    dev = _project(sc, cfg, density=dens)[0]
    plain = _project(sc, cfg)[0]
    nx, ny, nz = sc["shape"]
    host = {k: np.ascontiguousarray(sc[k]).copy() for k in ("fluid", "solid", "u", "v", "w", "us", "vs", "ws")}
    hd = [np.ascontiguousarray(d) for d in dens]
    p = lambda arr: arr.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    ps = shiokaze_amd.MacPressureSolver3(sc["shape"], sc["dx"], config=dict(cfg, HostPointers=1))
    assert ps.lib.shk_set_solid_density(ps.ctx, p(hd[0]), p(hd[1]), p(hd[2])) == 0
    res = shiokaze_amd.ShkResult()
    assert ps.lib.shk_project(ps.ctx, DT, p(host["u"]), p(host["v"]), p(host["w"]), p(host["solid"]),
                              p(host["us"]), p(host["vs"]), p(host["ws"]), p(host["fluid"]), 0.0,
                              ctypes.byref(res)) == 0
    for a, k in enumerate(("u", "v", "w")):
        np.testing.assert_array_equal(host[k], dev["velocity"][a])
    # one or two NULL arrays are refused; three NULLs switch it off
    assert ps.lib.shk_set_solid_density(ps.ctx, p(hd[0]), None, None) == shiokaze_amd._lib.SHK_BAD_ARG
    assert ps.lib.shk_set_solid_density(ps.ctx, None, None, None) == 0
    host = {k: np.ascontiguousarray(sc[k]).copy() for k in ("fluid", "solid", "u", "v", "w", "us", "vs", "ws")}
    assert ps.lib.shk_project(ps.ctx, DT, p(host["u"]), p(host["v"]), p(host["w"]), p(host["solid"]),
                              p(host["us"]), p(host["vs"]), p(host["ws"]), p(host["fluid"]), 0.0,
                              ctypes.byref(res)) == 0
    ps.close()
    for a, k in enumerate(("u", "v", "w")):
        np.testing.assert_array_equal(host[k], plain["velocity"][a])


# ---------------------------------------------------------------- first-order switches

def test_first_order_parity(scenes, oracle):
    sc = scenes(3, 24, 3)
    cfg = {"Residual": 1e-6, "Preconditioner": "none", "SecondOrderAccurateFluid": 0,
           "SecondOrderAccurateSolid": 0}
    g = _project(sc, cfg)[0]
    o = oracle.project(sc, tol=1e-6, mode=oracle.SHK_OR_CG, second_order_solid=False, second_order_fluid=False)
    _assert_fractions_bitexact(g, o)
    _assert_system_bitexact(oracle, g, o)


# ---------------------------------------------------------------- volume correction

def test_volume_correction_parity(scenes, oracle):
    sc = scenes(1, 24, 0)
    volumes = (1.08, 1.0)   # 8% too much liquid
    gs = _project(sc, {"Residual": 1e-8, "Preconditioner": "jacobi", "Gain": 1.0}, volumes=volumes, repeat=2)
    y = 0.0
    for call, g in enumerate(gs):
        c, y = oracle.volume_correction(1.0, volumes[0], volumes[1], DT, y)
        log(f"[gpu] volume correction call {call}: rhs constant gpu={g['result']['volume_correct_rhs']!r} "
            f"oracle={c!r}")
        assert g["result"]["volume_correct_rhs"] == c
        o = oracle.project(sc, tol=1e-8, mode=oracle.SHK_OR_JACOBI, rhs_correct=c)
        _assert_system_bitexact(oracle, g, o)
        # the projected divergence is the constant on every unknown (max |D| == |c| up
        # to the solver residual)
        b_inf = g["result"]["b_inf"]
        assert abs(g["div_after_fp64"] - abs(c)) <= 2e-8 * b_inf
    assert c < 0


def test_volume_correction_off_without_target(scenes):
    sc = scenes(1, 16, 0)
    g = _project(sc, {"Residual": 1e-6, "Gain": 1.0})[0]
    assert g["result"]["volume_correct_rhs"] == 0.0
    g = _project(sc, {"Residual": 1e-6, "Gain": 0.0}, volumes=(1.2, 1.0))[0]
    assert g["result"]["volume_correct_rhs"] == 0.0


# ---------------------------------------------------------------- warm start

@pytest.mark.parametrize("precond,precision", [("none", "fp64"), ("jacobi", "mixed"), ("mg", "mixed")])
def test_warm_start(scenes, oracle, precond, precision):
    sc = scenes(1, 24, 0)
    cfg = {"Residual": 1e-6, "Preconditioner": precond, "Precision": precision}
    cold = _project(sc, cfg)[0]
    w1, w2 = _project(sc, dict(cfg, WarmStart=1), repeat=2)
    # a warm start from the zero initial state is the cold start bit for bit
    np.testing.assert_array_equal(w1["system"]["x"], cold["system"]["x"])
    assert w1["result"]["count"] == cold["result"]["count"]
    # the second call solves for the change from the first call's pressure:
    # rhs' = rhs - A p_prev (the oracle's CSR product sums in another order)
    lab = cold["system"]["label"].astype(bool)
    o = oracle.project(sc, tol=1e-6, mode=oracle.SHK_OR_CG, warm_prev=w1["system"]["x"])
    ref = oracle_dense_system(oracle, o)
    # rounding scale of the product: its largest term, diag * p_prev
    scale = np.abs(cold["system"]["diag"][lab] * w1["system"]["x"][lab]).max()
    err = np.abs(w2["system"]["rhs"][lab] - ref["rhs"][lab]).max()
    log(f"[gpu] warm start {precond}/{precision}: iterations cold {cold['result']['count']} "
        f"second warm {w2['result']['count']}; rhs' err {err:.2e} of max|diag p| {scale:.2e}")
    assert err <= 1e-13 * scale
    # and reaches the same pressure as the cold solve, to the solvers' tolerance
    truth = oracle.project(sc, tol=1e-12, mode=oracle.SHK_OR_JACOBI)
    xt = truth["x"][truth["index_map"][lab]]
    rel = np.linalg.norm(w2["system"]["x"][lab] - xt) / np.linalg.norm(xt)
    log(f"[gpu] warm start {precond}/{precision}: second solve pressure rel-L2 vs 1e-12 = {rel:.2e}")
    # the second call solves for the change from the first pressure, to 1e-6 of the
    # change's own rhs: its pressure meets the bar
    assert rel <= 1e-5
    assert w2["result"]["status"] == 0


# ---------------------------------------------------------------- serialized state

def test_saved_state_resumes_bit_for_bit(scenes):
    """macpressuresolver3::serialize / initialize(filestream) (macpressuresolver3.cpp:
    319-344): a context that loads the state another context saved after a warm-started,
    volume-corrected projection continues exactly as that context does (same volume
    correction constant, iterations, pressure and velocity)."""
    import torch

    import shiokaze_amd
    from gpu_helpers import to_device
    sc1, sc2 = scenes(3, 24, 4), scenes(3, 24, 5)
    cfg = {"Residual": 1e-6, "WarmStart": 1, "Gain": 1.0}

    def run(ps, sc):
        d = to_device(sc)
        res = ps.project(1.0, (d["u"], d["v"], d["w"]), d["solid"], (d["us"], d["vs"], d["ws"]), d["fluid"])
        torch.cuda.synchronize()
        return res, ps.get_pressure().cpu().numpy(), [d[k].cpu().numpy() for k in ("u", "v", "w")]

    with shiokaze_amd.MacPressureSolver3(sc1["shape"], sc1["dx"], config=cfg) as a, \
            shiokaze_amd.MacPressureSolver3(sc1["shape"], sc1["dx"], config=cfg) as b:
        a.set_target_volume(0.95, 1.0)
        run(a, sc1)
        state = a.save_state()
        b.load_state(state)
        assert b.save_state() == state
        a.set_target_volume(0.97, 1.0)
        b.set_target_volume(0.97, 1.0)
        ra, pa, va = run(a, sc2)
        rb, pb, vb = run(b, sc2)
        log(f"[gpu] resumed state ({len(state)} bytes): rhs constant {rb['volume_correct_rhs']:.6g}, "
            f"{rb['count']} its (uninterrupted {ra['count']})")
        assert ra["volume_correct_rhs"] == rb["volume_correct_rhs"] != 0.0
        assert ra["count"] == rb["count"]
        np.testing.assert_array_equal(pa, pb)
        for x, y in zip(va, vb):
            np.testing.assert_array_equal(x, y)
        # a state of another shape, or a damaged one, is refused
        with shiokaze_amd.MacPressureSolver3((16, 24, 24), sc1["dx"], config=cfg) as c:
            with pytest.raises(shiokaze_amd.ShkError):
                c.load_state(state)
        with pytest.raises(shiokaze_amd.ShkError):
            b.load_state(state[:-8])
        with pytest.raises(shiokaze_amd.ShkError):
            b.load_state(b"XXXXXXXX" + state[8:])
