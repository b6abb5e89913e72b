"""GPU parity of the velocity half of the liquid step (include/shk_liquid.h; the
reference's macliquid3.cpp:342-391) through the C-ABI against the CPU oracle
(oracle/shk_liquid_oracle.c).  Every stage is bit-exact: the same float/double
rounding points, no reductions whose order matters (max |u| is a max)."""
from __future__ import annotations

import numpy as np
import pytest

from gpu_helpers import log

pytestmark = pytest.mark.gpu


def _t(a, dtype=None):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a if dtype is None else a.astype(dtype))).cuda()


def _h(t):
    return t.cpu().numpy()


def _solver(n, dx, config=None):
    from shiokaze_amd import MacLiquidVelocity, MacPressureSolver3
    ps = MacPressureSolver3(n, dx, device=0, config=config or {})
    return ps, MacLiquidVelocity(ps)


def _active_masks(sc, O):
    """The activity a projection leaves: faces with non-zero area and liquid fraction."""
    areas = O.area_fraction(sc["solid"])
    rhos = O.fluid_fraction(sc["fluid"])
    return [((a != 0) & (r != 0)).astype(np.uint8) for a, r in zip(areas, rhos)]


def _velocity(shapes, rng, scale=1.0):
    return [(rng.uniform(-scale, scale, s)).astype(np.float32) for s in shapes]


def _scene(scenes, kind, n, seed=0):
    from shiokaze_amd import scene as S
    return scenes(getattr(S, kind), n, seed)


CASES = [("POOL_SPHERE", 24, 0), ("IRREGULAR", 32, 1), ("IRREGULAR", (40, 28, 36), 3)]


@pytest.mark.parametrize("kind,n,seed", CASES)
@pytest.mark.parametrize("width", [1, 2, 4])
def test_extrapolate_matches_oracle(scenes, oracle, kind, n, seed, width):
    sc = _scene(scenes, kind, n, seed)
    O = oracle
    shape = tuple(sc["fluid"].shape[::-1])
    rng = np.random.default_rng(seed + 10 * width)
    masks = _active_masks(sc, O)
    vel = _velocity([m.shape for m in masks], rng)
    want, want_m, want_n = O.extrapolate(vel, masks, width)
    ps, lv = _solver(shape, sc["dx"])
    with ps:
        dv = [_t(v) for v in vel]
        dm = [_t(m) for m in masks]
        got_n = lv.extrapolate(dv, dm, width)
        log(f"[gpu] extrapolate {kind} {shape} width {width}: activated {got_n} (oracle {want_n})")
        assert got_n == want_n
        for d in range(3):
            np.testing.assert_array_equal(_h(dm[d]), want_m[d])
            act = want_m[d] == 1
            np.testing.assert_array_equal(_h(dv[d])[act], want[d][act])


def test_extrapolate_takes_any_nonzero_mask_as_active(scenes, oracle):
    sc = _scene(scenes, "IRREGULAR", 32, 1)
    O = oracle
    shape = tuple(sc["fluid"].shape[::-1])
    rng = np.random.default_rng(5)
    masks = _active_masks(sc, O)
    vel = _velocity([m.shape for m in masks], rng)
    want, want_m, want_n = O.extrapolate(vel, masks, 3)
    ps, lv = _solver(shape, sc["dx"])
    with ps:
        dv = [_t(v) for v in vel]
        dm = [_t((m * rng.integers(1, 256, m.shape)).astype(np.uint8)) for m in masks]   # 1..255 where active
        assert lv.extrapolate(dv, dm, 3) == want_n
        for d in range(3):
            np.testing.assert_array_equal(_h(dm[d]), want_m[d])
            np.testing.assert_array_equal(_h(dv[d])[want_m[d] == 1], want[d][want_m[d] == 1])


@pytest.mark.parametrize("kind,n,seed", CASES)
def test_constrain_matches_oracle(scenes, oracle, kind, n, seed):
    sc = _scene(scenes, kind, n, seed)
    O = oracle
    shape = tuple(sc["fluid"].shape[::-1])
    rng = np.random.default_rng(seed + 1)
    masks = [(rng.random(m.shape) < 0.9).astype(np.uint8) for m in _active_masks(sc, O)]   # walls included
    vel = _velocity([m.shape for m in masks], rng)
    want = O.constrain_velocity(sc["dx"], sc["solid"], vel, masks)
    ps, lv = _solver(shape, sc["dx"])
    with ps:
        dv = [_t(v) for v in vel]
        dm = [_t(m) for m in masks]
        lv.constrain(_t(sc["solid"]), dv, dm)
        changed = 0
        for d in range(3):
            got = _h(dv[d])
            act = masks[d] == 1
            np.testing.assert_array_equal(got[act], want[d][act])
            np.testing.assert_array_equal(got[~act], vel[d][~act])   # inactive faces are not touched
            changed += int((want[d] != vel[d]).sum())
        log(f"[gpu] constrain {kind} {shape}: {changed} faces changed, bit-exact")
        assert changed > 0


@pytest.mark.parametrize("kind,n,seed", CASES)
def test_compute_max_u_matches_oracle(scenes, oracle, kind, n, seed):
    sc = _scene(scenes, kind, n, seed)
    O = oracle
    shape = tuple(sc["fluid"].shape[::-1])
    rng = np.random.default_rng(seed + 2)
    masks = _active_masks(sc, O)
    vel = _velocity([m.shape for m in masks], rng, 3.0)
    want = O.compute_max_u(vel, masks)
    ps, lv = _solver(shape, sc["dx"])
    with ps:
        got = lv.compute_max_u([_t(v) for v in vel], [_t(m) for m in masks])
        log(f"[gpu] max_u {kind} {shape}: {got!r} vs {want!r}")
        assert got == want
        none = lv.compute_max_u([_t(v) for v in vel], [_t(np.zeros_like(m)) for m in masks])
        assert none == 0.0


@pytest.mark.parametrize("kind,n,seed", CASES)
@pytest.mark.parametrize("maccormack,trim", [(True, 1), (False, 1), (True, 3)])
def test_advect_matches_oracle(scenes, oracle, kind, n, seed, maccormack, trim):
    sc = _scene(scenes, kind, n, seed)
    O = oracle
    shape = tuple(sc["fluid"].shape[::-1])
    rng = np.random.default_rng(seed + 3)
    masks = _active_masks(sc, O)
    _, masks, _ = O.extrapolate(_velocity([m.shape for m in masks], rng), masks, 3)
    vel = _velocity([m.shape for m in masks], rng, 2.0)
    dt = 1.5 * sc["dx"] / 2.0   # departures up to 1.5 cells
    want = O.advect_velocity(sc["dx"], dt, vel, masks, sc["fluid"], maccormack=maccormack, trim=trim)
    ps, lv = _solver(shape, sc["dx"], {"MacCormack": maccormack, "TrimNarrowBand": trim})
    with ps:
        dv = [_t(v) for v in vel]
        lv.advect(dt, dv, [_t(m) for m in masks], _t(sc["fluid"]))
        for d in range(3):
            np.testing.assert_array_equal(_h(dv[d]), want[d])
        log(f"[gpu] advect {kind} {shape} maccormack={maccormack} trim={trim}: bit-exact")


@pytest.mark.parametrize("kind,n,seed", CASES)
@pytest.mark.parametrize("maccormack,trim", [(True, 1), (False, 1), (True, 2)])
def test_advect_levelset_matches_oracle(scenes, oracle, kind, n, seed, maccormack, trim):
    sc = _scene(scenes, kind, n, seed)
    O = oracle
    shape = tuple(sc["fluid"].shape[::-1])
    dx = sc["dx"]
    rng = np.random.default_rng(seed + 4)
    masks = _active_masks(sc, O)
    _, masks, _ = O.extrapolate(_velocity([m.shape for m in masks], rng), masks, 3)
    vel = _velocity([m.shape for m in masks], rng, 2.0)
    bw = 3 * dx
    phi = np.clip(sc["fluid"], -bw, bw).astype(np.float32)     # the band values and the fill / background reads
    band = (np.abs(sc["fluid"]) < bw).astype(np.uint8)
    dt = 1.5 * dx / 2.0
    want = O.advect_levelset(dx, dt, phi, band, vel, masks, bw, maccormack=maccormack, trim=trim)
    ps, lv = _solver(shape, dx, {"MacCormack": maccormack, "TrimNarrowBand": trim})
    with ps:
        dphi = _t(phi)
        lv.advect_levelset(dt, dphi, _t(band), [_t(v) for v in vel], [_t(m) for m in masks], bw)
        got = _h(dphi)
        np.testing.assert_array_equal(got, want)
        log(f"[gpu] advect_levelset {kind} {shape} maccormack={maccormack} trim={trim}: "
            f"{int((got != phi).sum())} of {int(band.sum())} band cells moved, bit-exact")


def test_gravity_and_projection_activity_match_oracle(scenes, oracle):
    sc = _scene(scenes, "IRREGULAR", 32, 1)
    O = oracle
    shape = tuple(sc["fluid"].shape[::-1])
    rng = np.random.default_rng(7)
    masks = [(rng.random(m.shape) < 0.5).astype(np.uint8) for m in _active_masks(sc, O)]
    vel = _velocity([m.shape for m in masks], rng)
    dt = 0.004
    want, want_m = O.add_gravity(dt, (0.0, -9.8, 0.0), vel, masks)
    ps, lv = _solver(shape, sc["dx"], {"Residual": 1e-6})
    with ps:
        dv = [_t(v) for v in vel]
        dm = [_t(m) for m in masks]
        lv.add_gravity(dt, dv, dm)
        for d in range(3):
            np.testing.assert_array_equal(_h(dv[d]), want[d])
            np.testing.assert_array_equal(_h(dm[d]), want_m[d])
        svel = [_t(np.zeros(m.shape, np.float32)) for m in masks]
        res = ps.project(dt, dv, _t(sc["solid"]), svel, _t(sc["fluid"]))
        assert res["unknowns"] > 0
        lv.project_activity(dm)
        fr = ps.export_fractions()
        areas = [fr[f"area{a}"].cpu().numpy().astype(np.float64) for a in range(3)]
        rhos = [fr[f"rho{a}"].cpu().numpy().astype(np.float64) for a in range(3)]
        want_act = O.project_activity(areas, rhos, want_m)
        for d in range(3):
            np.testing.assert_array_equal(_h(dm[d]), want_act[d])
        log(f"[gpu] gravity + projection activity: {sum(int(_h(m).sum()) for m in dm)} active faces")


def test_liquid_steps_stagewise_against_oracle(scenes, oracle):
    """Three steps of macliquid3's velocity half with the projection; every stage of
    every step is checked bit-for-bit against the oracle applied to the GPU's state."""
    from shiokaze_amd import scene as S
    O = oracle
    sc = scenes(S.POOL_SPHERE, 32, 0)
    shape = tuple(sc["fluid"].shape[::-1])
    dx = sc["dx"]
    masks = _active_masks(sc, O)
    rng = np.random.default_rng(11)
    vel = [np.where(m == 1, v, 0).astype(np.float32) for m, v in zip(masks, _velocity([m.shape for m in masks], rng))]
    ps, lv = _solver(shape, dx, {"Residual": 1e-6, "Preconditioner": "mg", "Precision": "mixed"})
    with ps:
        dv = [_t(v) for v in vel]
        dm = [_t(m) for m in masks]
        solid, fluid = _t(sc["solid"]), _t(sc["fluid"])
        svel = [_t(np.zeros(m.shape, np.float32)) for m in masks]
        for step in range(3):
            hv, hm = [_h(v) for v in dv], [_h(m) for m in dm]
            max_u = lv.compute_max_u(dv, dm)
            assert max_u == O.compute_max_u(hv, hm)
            dt = lv.timestepper.advance(max_u, dx)
            width = int(lv.extend_width + lv.timestepper.current_cfl)
            ev, em, _ = O.extrapolate(hv, hm, width)
            lv.extrapolate(dv, dm, width)
            for d in range(3):
                np.testing.assert_array_equal(_h(dm[d]), em[d])
                np.testing.assert_array_equal(_h(dv[d])[em[d] == 1], ev[d][em[d] == 1])
            hv, hm = [_h(v) for v in dv], [_h(m) for m in dm]
            cv = O.constrain_velocity(dx, sc["solid"], hv, hm)
            lv.constrain(solid, dv, dm)
            for d in range(3):
                np.testing.assert_array_equal(_h(dv[d])[hm[d] == 1], cv[d][hm[d] == 1])
            hv = [_h(v) for v in dv]
            av = O.advect_velocity(dx, dt, hv, hm, sc["fluid"])
            lv.advect(dt, dv, dm, fluid)
            for d in range(3):
                np.testing.assert_array_equal(_h(dv[d]), av[d])
            gv, gm = O.add_gravity(dt, lv.gravity, av, hm)
            lv.add_gravity(dt, dv, dm)
            for d in range(3):
                np.testing.assert_array_equal(_h(dv[d]), gv[d])
                np.testing.assert_array_equal(_h(dm[d]), gm[d])
            res = ps.project(dt, dv, solid, svel, fluid)
            lv.project_activity(dm)
            log(f"[gpu] step {step}: dt {dt:.4g} CFL {lv.timestepper.current_cfl:.3f} width {width} "
                f"its {res['count']} relres {res['true_relres_inf']:.2e}")
            assert res["status"] == 0 and res["true_relres_inf"] < 1e-5
            assert max(int(_h(m).sum()) for m in dm) > 0


def test_step_method_runs_and_keeps_the_activity_consistent(scenes, oracle):
    from shiokaze_amd import scene as S
    O = oracle
    sc = scenes(S.IRREGULAR, 32, 1)
    shape = tuple(sc["fluid"].shape[::-1])
    masks = _active_masks(sc, O)
    ps, lv = _solver(shape, sc["dx"], {"Residual": 1e-6, "Precision": "mixed", "Preconditioner": "mg"})
    with ps:
        dv = [_t(np.zeros(m.shape, np.float32)) for m in masks]
        dm = [_t(m) for m in masks]
        svel = [_t(np.zeros(m.shape, np.float32)) for m in masks]
        for step in range(4):
            res = lv.step(dv, dm, _t(sc["solid"]), svel, _t(sc["fluid"]))
            log(f"[gpu] step() {step}: dt {res['dt']:.4g} max_u {res['max_u']:.3f} width {res['extend_width']} "
                f"extended {res['extended']} its {res['count']}")
            assert res["status"] == 0
        fr = ps.export_fractions()
        for d in range(3):
            a = fr[f"area{d}"].cpu().numpy()
            r = fr[f"rho{d}"].cpu().numpy()
            m = _h(dm[d]).astype(bool)
            assert np.all(m == ((a != 0) & (r != 0)))   # after gravity every face was active
            assert np.all(_h(dv[d])[~m] == 0.0)         # and the projection zeroed what it switched off


def test_liquid_calls_refuse_2d_and_host_pointers():
    import ctypes

    import torch

    from shiokaze_amd import MacLiquidVelocity, MacPressureSolver2, MacPressureSolver3, ShkError
    from shiokaze_amd import _lib
    ps2 = MacPressureSolver2((16, 16), 1.0 / 16, device=0)
    with ps2:
        buf = torch.zeros(64, device="cuda")
        p = ctypes.c_void_p(buf.data_ptr())
        assert ps2.lib.shk_add_gravity(ps2.ctx, 0.01, 0.0, -9.8, 0.0, p, p, p, p, p, p) == _lib.SHK_BAD_ARG
    ps = MacPressureSolver3((8, 8, 8), 1.0 / 8, device=0, config={"HostPointers": True})
    with ps:
        lv = MacLiquidVelocity(ps)
        u = [torch.zeros(ps.face_count(d), device="cuda") for d in range(3)]
        m = [torch.zeros(ps.face_count(d), device="cuda", dtype=torch.uint8) for d in range(3)]
        with pytest.raises(ShkError):
            lv.add_gravity(0.01, u, m)
    ps = MacPressureSolver3((8, 8, 8), 1.0 / 8, device=0)
    with ps:
        with pytest.raises(ShkError):
            ps.configure("WENO", "maybe")


def test_liquid_edge_cases(scenes, oracle):
    """Width 0, dt = 0, all-inactive masks and the smallest grid (2 cells per axis),
    each against the oracle."""
    O = oracle
    rng = np.random.default_rng(12)
    for shape in ((2, 2, 2), (9, 5, 3)):
        nx, ny, nz = shape
        dx = 1.0 / nx
        fshape = [(nz, ny, nx + 1), (nz, ny + 1, nx), (nz + 1, ny, nx)]
        vel = [rng.uniform(-1, 1, s).astype(np.float32) for s in fshape]
        masks = [(rng.random(s) < 0.5).astype(np.uint8) for s in fshape]
        fluid = rng.uniform(-0.2, 0.2, (nz, ny, nx)).astype(np.float32)
        solid = rng.uniform(-0.2, 0.2, (nz + 1, ny + 1, nx + 1)).astype(np.float32)
        ps, lv = _solver(shape, dx)
        with ps:
            dv, dm = [_t(v) for v in vel], [_t(m) for m in masks]
            assert lv.extrapolate(dv, dm, 0) == 0
            for d in range(3):
                np.testing.assert_array_equal(_h(dv[d]), vel[d])
                np.testing.assert_array_equal(_h(dm[d]), masks[d])
            for dt in (0.0, 0.3 * dx):
                want = O.advect_velocity(dx, dt, vel, masks, fluid)
                dv = [_t(v) for v in vel]
                lv.advect(dt, dv, dm, _t(fluid))
                for d in range(3):
                    np.testing.assert_array_equal(_h(dv[d]), want[d])
            want = O.constrain_velocity(dx, solid, vel, masks)
            dv = [_t(v) for v in vel]
            lv.constrain(_t(solid), dv, dm)
            for d in range(3):
                np.testing.assert_array_equal(_h(dv[d])[masks[d] == 1], want[d][masks[d] == 1])
            none = [np.zeros(s, np.uint8) for s in fshape]
            ev, em, n = O.extrapolate(vel, none, 3)
            dv, dn = [_t(v) for v in vel], [_t(m) for m in none]
            assert lv.extrapolate(dv, dn, 3) == n == 0
            assert lv.compute_max_u(dv, dn) == 0.0
            want = O.advect_velocity(dx, 0.1, vel, none, fluid)
            lv.advect(0.1, dv, dn, _t(fluid))
            for d in range(3):
                np.testing.assert_array_equal(_h(dv[d]), want[d])
    # one cell per axis is refused (trilinear interpolation needs two)
    ps, lv = _solver((1, 4, 4), 0.25)
    with ps:
        import torch
        from shiokaze_amd import ShkError
        v = [torch.zeros(ps.face_count(d), device="cuda") for d in range(3)]
        m = [torch.zeros(ps.face_count(d), device="cuda", dtype=torch.uint8) for d in range(3)]
        with pytest.raises(ShkError):
            lv.add_gravity(0.01, v, m)


def _offset_view(a, dtype=None):
    """A device copy of `a` that starts one element into a larger buffer: contiguous but
    not 16-byte aligned, so the kernels take their element-by-element path."""
    import torch
    h = np.ascontiguousarray(a if dtype is None else a.astype(dtype)).ravel()
    buf = torch.zeros(h.size + 1, dtype=torch.from_numpy(h[:1]).dtype, device="cuda")
    view = buf[1:h.size + 1]
    view.copy_(torch.from_numpy(h).cuda())
    assert view.data_ptr() % 16 != 0
    return view


@pytest.mark.parametrize("shape", [(2, 2, 3), (13, 9, 11), (64, 64, 64)])
def test_unaligned_arrays_match_oracle(scenes, oracle, shape):
    """Velocities and masks passed as views that are not 16-byte aligned (slices of a
    packed buffer): gravity, the mask normalisation of the extrapolation and the
    projection's activity are bit-exact against the oracle, every face is incremented
    exactly once, and each thread only touches its own four faces (the 64^3 case would
    take quadratic time otherwise)."""
    import time
    O = oracle
    nx, ny, nz = shape
    dx = 1.0 / max(shape)
    fshape = [(nz, ny, nx + 1), (nz, ny + 1, nx), (nz + 1, ny, nx)]
    rng = np.random.default_rng(23)
    vel = [rng.uniform(-1, 1, s).astype(np.float32) for s in fshape]
    # tags > 1 exercise the normalisation (any nonzero mask is active)
    masks = [(rng.integers(0, 3, s) * (rng.random(s) < 0.4)).astype(np.uint8) for s in fshape]
    dt = 0.004
    ps, lv = _solver(shape, dx)
    with ps:
        dv = [_offset_view(v) for v in vel]
        dm = [_offset_view(m) for m in masks]
        t0 = time.perf_counter()
        lv.add_gravity(dt, dv, dm)
        want, want_m = O.add_gravity(dt, lv.gravity, vel, [(m != 0).astype(np.uint8) for m in masks])
        for d in range(3):
            np.testing.assert_array_equal(_h(dv[d]).reshape(fshape[d]), want[d])
            np.testing.assert_array_equal(_h(dm[d]).reshape(fshape[d]), want_m[d])
        # extrapolation normalises arbitrary nonzero masks to 1 before its rounds
        dv = [_offset_view(v) for v in vel]
        dm = [_offset_view(m) for m in masks]
        ev, em, n = O.extrapolate(vel, [(m != 0).astype(np.uint8) for m in masks], 2)
        assert lv.extrapolate(dv, dm, 2) == n
        for d in range(3):
            np.testing.assert_array_equal(_h(dm[d]).reshape(fshape[d]), em[d])
            np.testing.assert_array_equal(_h(dv[d]).reshape(fshape[d])[em[d] == 1], ev[d][em[d] == 1])
        # the projection's activity from a projection of a pool scene of this shape
        from shiokaze_amd import scene as S
        sc = S.generate(S.POOL_SPHERE, shape, 0)
        d = {k: _t(sc[k]) for k in ("fluid", "solid", "u", "v", "w", "us", "vs", "ws")}
        ps.project(1.0, (d["u"], d["v"], d["w"]), d["solid"], (d["us"], d["vs"], d["ws"]), d["fluid"])
        dm = [_offset_view(m) for m in masks]
        lv.project_activity(dm)
        fr = ps.export_fractions()
        areas = [fr[f"area{a}"].cpu().numpy().astype(np.float64).reshape(fshape[a]) for a in range(3)]
        rhos = [fr[f"rho{a}"].cpu().numpy().astype(np.float64).reshape(fshape[a]) for a in range(3)]
        want_act = O.project_activity(areas, rhos, masks)
        for a in range(3):
            np.testing.assert_array_equal(_h(dm[a]).reshape(fshape[a]) != 0, want_act[a] != 0)
        elapsed = time.perf_counter() - t0
    log(f"[gpu] unaligned views {shape}: gravity, normalisation and activity bit-exact in {elapsed:.2f} s")
    assert elapsed < 30.0


# ---------------------------------------------------------------- PDE redistancing

def _distorted_levelset(n, seed):
    """A sphere level set with its distance distorted (slopes 0.5..2 by region), its
    narrow band at 3 cells, and the background values outside it (-/+ 3 dx)."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / n
    x = (np.arange(n) + 0.5) * dx
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    c = rng.uniform(0.35, 0.65, 3)
    d = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - rng.uniform(0.15, 0.3)
    d = np.minimum(d, Y - rng.uniform(0.2, 0.35))
    phi = (d * (1.25 + 0.75 * np.sin(7 * X + 3 * Z))).astype(np.float32)
    band = (np.abs(phi) < 3 * dx).astype(np.uint8)
    phi = np.where(band > 0, phi, np.where(phi < 0, -3 * dx, 3 * dx)).astype(np.float32)
    return dx, phi, band


@pytest.mark.parametrize("n,seed,width", [(24, 0, 3), (40, 1, 3), (33, 2, 5), (17, 3, 1)])
@pytest.mark.parametrize("scheme", ["Euler", "RK2"])
def test_redistance_matches_oracle(oracle, n, seed, width, scheme):
    """shk_redistance against the oracle's pderedistancer3 restatement: band, values
    and count bit for bit."""
    dx, phi, band = _distorted_levelset(n, seed)
    ps, lv = _solver((n, n, n), dx, {"RedistTemporalScheme": scheme})
    dphi, dband = _t(phi), _t(band)
    cnt = lv.redistance(dphi, dband, width)
    ps.close()
    ophi, oband, ocnt = oracle.redistance(dx, phi, band, width, rk2=scheme == "RK2")
    log(f"[gpu] redistance n={n} seed={seed} width={width} {scheme}: band {cnt} (oracle {ocnt})")
    assert cnt == ocnt
    np.testing.assert_array_equal(_h(dband), oband)
    np.testing.assert_array_equal(_h(dphi), ophi)


def test_redistance_rate_and_refusals(oracle):
    import shiokaze_amd
    dx, phi, band = _distorted_levelset(20, 4)
    ps, lv = _solver((20, 20, 20), dx, {"IntegrationRate": 0.5})
    dphi, dband = _t(phi), _t(band)
    lv.redistance(dphi, dband, 2)
    ophi, oband, _ = oracle.redistance(dx, phi, band, 2, rate=0.5)
    np.testing.assert_array_equal(_h(dphi), ophi)
    for key, val in (("IntegrationRate", "0"), ("IntegrationRate", "1.5"), ("RedistTemporalScheme", "RK4")):
        with pytest.raises(shiokaze_amd.ShkError):
            ps.configure(key, val)
    with pytest.raises(shiokaze_amd.ShkError):
        lv.redistance(dphi, dband, 0)
    ps.close()


@pytest.mark.parametrize("kind,n,seed", CASES[:2])
@pytest.mark.parametrize("maccormack", [True, False])
def test_advect_weno_matches_oracle(scenes, oracle, kind, n, seed, maccormack):
    """WENO=1: the semi-Lagrangian values of the velocity and of the level set by 6th-order
    WENO interpolation (macadvection3.cpp:85-86, 193-194), bit for bit with the oracle."""
    sc = _scene(scenes, kind, n, seed)
    O = oracle
    shape = tuple(sc["fluid"].shape[::-1])
    dx = sc["dx"]
    rng = np.random.default_rng(seed + 7)
    masks = _active_masks(sc, O)
    _, masks, _ = O.extrapolate(_velocity([m.shape for m in masks], rng), masks, 3)
    vel = _velocity([m.shape for m in masks], rng, 2.0)
    dt = 1.5 * dx / 2.0
    want = O.advect_velocity(dx, dt, vel, masks, sc["fluid"], maccormack=maccormack, weno=True)
    bw = 3 * dx
    phi = np.clip(sc["fluid"], -bw, bw).astype(np.float32)
    band = (np.abs(sc["fluid"]) < bw).astype(np.uint8)
    want_phi = O.advect_levelset(dx, dt, phi, band, vel, masks, bw, maccormack=maccormack, weno=True)
    plain = O.advect_velocity(dx, dt, vel, masks, sc["fluid"], maccormack=maccormack)
    ps, lv = _solver(shape, dx, {"MacCormack": maccormack, "WENO": 1})
    with ps:
        dv = [_t(v) for v in vel]
        lv.advect(dt, dv, [_t(m) for m in masks], _t(sc["fluid"]))
        for d in range(3):
            np.testing.assert_array_equal(_h(dv[d]), want[d])
        dphi = _t(phi)
        lv.advect_levelset(dt, dphi, _t(band), [_t(v) for v in vel], [_t(m) for m in masks], bw)
        np.testing.assert_array_equal(_h(dphi), want_phi)
    moved = sum(int((w != p).sum()) for w, p in zip(want, plain))
    log(f"[gpu] advect WENO {kind} {shape} maccormack={maccormack}: bit-exact ({moved} faces differ from trilinear)")
    assert moved > 0


def _tracker_inputs(scenes, oracle, kind, n, seed, width=3):
    """The level set as the tracker holds it before extrapolate_levelset: the scene's liquid
    redistanced by the oracle (band and -/+ width*dx off it), the scene's nodal solid and
    its band |solid| < 3 dx with -/+ 3 dx off it (as the reference's solid array reads)."""
    sc = _scene(scenes, kind, n, seed)
    dx = sc["dx"]
    bw = width * dx
    phi = np.clip(sc["fluid"], -bw, bw).astype(np.float32)
    band = (np.abs(sc["fluid"]) < bw).astype(np.uint8)
    phi, band, _ = oracle.redistance(dx, phi, band, width)
    solid = sc["solid"].astype(np.float32)
    sband = (np.abs(solid) < 3 * dx).astype(np.uint8)
    solid = np.where(sband > 0, solid, np.where(solid < 0, -3 * dx, 3 * dx)).astype(np.float32)
    return sc, dx, phi, band, solid, sband, np.float32(bw)


@pytest.mark.parametrize("kind,n,seed", CASES)
@pytest.mark.parametrize("keys", [{}, {"HorizontalSolidWallExtrapolation": 0}, {"SolidWallExtrapolation": 0}])
def test_extrapolate_levelset_matches_oracle(scenes, oracle, kind, n, seed, keys):
    """shk_extrapolate_levelset against the oracle's gridutility3::extrapolate_levelset:
    values, band and count bit for bit (the level set redistanced first, as the tracker
    runs them)."""
    sc, dx, phi, band, solid, sband, bg = _tracker_inputs(scenes, oracle, kind, n, seed)
    shape = tuple(phi.shape[::-1])
    wall = bool(keys.get("SolidWallExtrapolation", 1))
    horizontal = bool(keys.get("HorizontalSolidWallExtrapolation", 1))
    ophi, oband, ocnt = oracle.extrapolate_levelset(dx, solid, phi, band, bg, solid_band=sband, wall=wall,
                                                    horizontal=horizontal)
    ps, lv = _solver(shape, dx, keys)
    with ps:
        dphi, dband = _t(phi), _t(band)
        cnt = lv.extrapolate_levelset(_t(solid), dphi, dband, float(bg), solid_band=_t(sband))
        log(f"[gpu] extrapolate_levelset {kind} {n} seed={seed} {keys}: band {cnt} (oracle {ocnt}), "
            f"{int((ophi != phi).sum())} values changed")
        assert cnt == ocnt
        np.testing.assert_array_equal(_h(dband), oband)
        np.testing.assert_array_equal(_h(dphi), ophi)


def test_extrapolate_levelset_without_solid_and_all_nodes_active(scenes, oracle):
    """No active solid node below 0: nothing changes.  A NULL node mask: every node is
    active (every cell joins the band)."""
    sc, dx, phi, band, solid, sband, bg = _tracker_inputs(scenes, oracle, "POOL_SPHERE", 24, 0)
    shape = tuple(phi.shape[::-1])
    ps, lv = _solver(shape, dx)
    with ps:
        dphi, dband = _t(phi), _t(band)
        cnt = lv.extrapolate_levelset(_t(np.abs(solid) + np.float32(dx)), dphi, dband, float(bg))
        assert cnt == int(band.sum())
        np.testing.assert_array_equal(_h(dphi), phi)
        np.testing.assert_array_equal(_h(dband), band)
        sc3, dx3, phi3, band3, solid3, _, bg3 = _tracker_inputs(scenes, oracle, "IRREGULAR", 32, 1)
    ophi, oband, ocnt = oracle.extrapolate_levelset(dx3, solid3, phi3, band3, bg3)
    ps, lv = _solver(tuple(phi3.shape[::-1]), dx3)
    with ps:
        dphi, dband = _t(phi3), _t(band3)
        cnt = lv.extrapolate_levelset(_t(solid3), dphi, dband, float(bg3))
        assert cnt == ocnt == phi3.size
        np.testing.assert_array_equal(_h(dphi), ophi)
        np.testing.assert_array_equal(_h(dband), oband)
