"""GPU parity of the 2D projection (shk_create2d / shk_project2d, the counterpart of
macpressuresolver2.cpp:52-281) against the oracle's 2D restatement.

Bit-exact: fractions, labels, system, the surface-tension and solid-density
branches, and the velocity update given the same pressure.  Within tolerance: the
Krylov solution (iteration counts inside the spread that rounding order gives the
oracle) and the FP64 post-projection divergence."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from gpu_helpers import BAR_STOP_2D, PRESSURE_BAR, iteration_spread, log, within_spread

pytestmark = pytest.mark.gpu

CASES = [(1, (64, 48), 0), (3, (48, 64), 2), (3, (70, 37), 5)]
FIELDS2 = ("fluid", "solid", "u", "v", "us", "vs")


def _ids(c):
    return f"kind{c[0]}-{c[1][0]}x{c[1][1]}-seed{c[2]}"


def _scene(kind, shape, seed):
    from shiokaze_amd import scene as S
    return S.generate2d(kind, shape, seed)


def _dev(sc):
    import torch
    return {k: torch.from_numpy(np.ascontiguousarray(sc[k])).cuda() for k in FIELDS2}


def run2d(sc, config, surface_tension=0.0, density=None, repeat=1, dt=1.0):
    import torch

    from shiokaze_amd import MacPressureSolver2
    nx, ny = sc["fluid"].shape[1], sc["fluid"].shape[0]
    outs = []
    with MacPressureSolver2((nx, ny), sc["dx"], config=config) as ps:
        if density is not None:
            ps.set_solid_density([torch.from_numpy(np.ascontiguousarray(d, np.float32)).cuda() for d in density])
        for _ in range(repeat):
            d = _dev(sc)
            vel0 = [d["u"].clone(), d["v"].clone()]
            res = ps.project(dt, (d["u"], d["v"]), d["solid"], (d["us"], d["vs"]), d["fluid"],
                             surface_tension=surface_tension)
            out = {"result": res, "pressure": ps.get_pressure().cpu().numpy(),
                   "velocity": [d["u"].cpu().numpy(), d["v"].cpu().numpy()]}
            fr = ps.export_fractions()
            out["areas"] = [fr["area0"].cpu().numpy(), fr["area1"].cpu().numpy()]
            out["rhos"] = [fr["rho0"].cpu().numpy(), fr["rho1"].cpu().numpy()]
            if res["unknowns"]:
                out["system"] = {k: v.cpu().numpy() for k, v in ps.export_system().items()}
                svel = (d["us"], d["vs"])
                out["div_before"] = ps.divergence(dt, vel0, svel, fp64_update=False)
                out["div_after_fp64"] = ps.divergence(dt, vel0, svel, fp64_update=True)
            outs.append(out)
        torch.cuda.synchronize()
    return outs


def _dense_2d(o):
    """Oracle CSR -> dense (ny, nx) diag, low-face couplings along x and y, rhs, x."""
    im = o["index_map2"]
    ny, nx = im.shape
    n = o["n"]
    flat = im.ravel()
    cells = np.nonzero(flat >= 0)[0]
    cell_of_row = np.empty(n, dtype=np.int64)
    cell_of_row[flat[cells]] = cells
    rr = np.repeat(np.arange(n), np.diff(o["rowptr"]))
    cr, cc = cell_of_row[rr], cell_of_row[o["cols"]]
    delta = cc - cr
    out = {k: np.zeros(nx * ny) for k in ("diag", "cx", "cy", "rhs", "x")}
    out["diag"][cr[delta == 0]] = o["vals"][delta == 0]
    out["cx"][cr[delta == -1]] = -o["vals"][delta == -1]
    out["cy"][cr[delta == -nx]] = -o["vals"][delta == -nx]
    out["rhs"][cell_of_row] = o["rhs"]
    out["x"][cell_of_row] = o["x"]
    return {k: v.reshape(ny, nx) for k, v in out.items()}


def _check_fractions(g, o):
    for a in range(2):
        np.testing.assert_array_equal(g["areas"][a].astype(np.float64), o["areas2"][a])
        np.testing.assert_array_equal(g["rhos"][a].astype(np.float64), o["rhos2"][a])


def _check_system(g, o):
    lab = o["index_map2"] >= 0
    s = g["system"]
    np.testing.assert_array_equal(s["label"].astype(bool), lab)
    ref = _dense_2d(o)
    np.testing.assert_array_equal(s["diag"][lab], ref["diag"][lab])
    np.testing.assert_array_equal(s["rhs"][lab], ref["rhs"][lab])
    both_x = lab[:, 1:] & lab[:, :-1]
    both_y = lab[1:, :] & lab[:-1, :]
    np.testing.assert_array_equal(s["cx"][:, 1:][both_x], ref["cx"][:, 1:][both_x])
    # the 2D y axis is the library's z axis
    np.testing.assert_array_equal(s["cz"][1:, :][both_y], ref["cy"][1:, :][both_y])
    assert not np.any(s["cy"])


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_2d_parity(oracle, case):
    kind, shape, seed = case
    sc = _scene(kind, shape, seed)
    log(f"[gpu] 2D kind={kind} shape={shape} seed={seed}")
    g = run2d(sc, {"Residual": 1e-6, "Preconditioner": "none", "Precision": "fp64"})[0]
    o = oracle.project2d(sc, tol=1e-6, mode=oracle.SHK_OR_CG)
    _check_fractions(g, o)
    assert g["result"]["unknowns"] == o["n"]
    _check_system(g, o)
    r = g["result"]
    spread = iteration_spread(oracle, o["rowptr"], o["cols"], o["vals"], o["rhs"], 1e-6, oracle.SHK_OR_CG)
    log(f"[gpu] 2D iters gpu={r['count']} oracle={o['iters']} renumbered={spread} true={r['true_relres_inf']:.2e}")
    assert r["status"] == 0 and within_spread(r["count"], spread)
    assert r["true_relres_inf"] <= 1e-6 * 1.0001
    # the velocity update from the GPU pressure is the GPU's velocity bit for bit
    ny, nx = sc["fluid"].shape
    p = g["pressure"].astype(np.float64).reshape(ny, 1, nx)
    svel = [np.asarray(sc["us"], np.float64).reshape(ny, 1, nx + 1), np.zeros((ny, 2, nx)),
            np.asarray(sc["vs"], np.float64).reshape(ny + 1, 1, nx)]
    vel = oracle.update_velocity(1.0, sc["dx"], p, o["areas"], o["rhos"], svel, o["velocity_in"])
    np.testing.assert_array_equal(g["velocity"][0].astype(np.float64), vel[0].reshape(ny, nx + 1))
    np.testing.assert_array_equal(g["velocity"][1].astype(np.float64), vel[2].reshape(ny + 1, nx))
    # FP64 post-projection divergence = the true residual
    rel = g["div_after_fp64"] / g["div_before"]
    assert rel <= 1e-6 * 1.0001
    assert rel == pytest.approx(r["true_relres_inf"], rel=1e-6, abs=1e-12)


@pytest.mark.parametrize("precond,precision", [("jacobi", "fp64"), ("jacobi", "mixed"), ("none", "mixed")])
def test_2d_solvers(oracle, precond, precision):
    """Each 2D solver path reaches the reference's stop, and at the 2D bar stop meets
    the pressure bar against the oracle's FP64 solve to 1e-12."""
    sc = _scene(3, (96, 80), 3)
    truth = oracle.project2d(sc, tol=1e-12, mode=oracle.SHK_OR_JACOBI)
    lab = truth["index_map2"] >= 0
    xt = _dense_2d(truth)["x"][lab]
    for stop in (1e-6, BAR_STOP_2D[precond]):
        g = run2d(sc, {"Residual": stop, "Preconditioner": precond, "Precision": precision})[0]
        rel = np.linalg.norm(g["system"]["x"][lab] - xt) / np.linalg.norm(xt)
        r = g["result"]
        log(f"[gpu] 2D {precond}/{precision} stop {stop:g}: iters {r['count']} outer {r['outer']} "
            f"true {r['true_relres_inf']:.2e} x rel {rel:.2e}")
        assert r["status"] == 0 and r["true_relres_inf"] <= stop * 1.0001
    assert rel <= PRESSURE_BAR


def test_2d_branches(oracle):
    """Surface tension, solid density, and a warm start in 2D."""
    sc = _scene(3, (64, 64), 4)
    ny, nx = sc["fluid"].shape
    rng = np.random.default_rng(4)
    dens = [rng.uniform(0, 1, (ny, nx + 1)).astype(np.float32), rng.uniform(0, 1, (ny + 1, nx)).astype(np.float32)]
    g = run2d(sc, {"Residual": 1e-6, "Preconditioner": "none"}, surface_tension=0.03, density=dens)[0]
    o = oracle.project2d(sc, tol=1e-6, mode=oracle.SHK_OR_CG, surface_tension_coef=0.03,
                         solid_density=[d.astype(np.float64) for d in dens])
    _check_fractions(g, o)
    _check_system(g, o)
    p = g["pressure"].astype(np.float64).reshape(ny, 1, nx)
    svel = [np.asarray(sc["us"], np.float64).reshape(ny, 1, nx + 1), np.zeros((ny, 2, nx)),
            np.asarray(sc["vs"], np.float64).reshape(ny + 1, 1, nx)]
    vel = oracle.update_velocity(1.0, sc["dx"], p, o["areas"], o["rhos"], svel, o["velocity_in"])
    d3 = [dens[0].astype(np.float64).reshape(ny, 1, nx + 1), np.zeros((ny, 2, nx)),
          dens[1].astype(np.float64).reshape(ny + 1, 1, nx)]
    vel = oracle.density_scale(d3, o["areas"], vel)
    np.testing.assert_array_equal(g["velocity"][0].astype(np.float64), vel[0].reshape(ny, nx + 1))
    np.testing.assert_array_equal(g["velocity"][1].astype(np.float64), vel[2].reshape(ny + 1, nx))
    # warm start: the first call is the cold start, the second (which solves for the
    # change from the first pressure) meets the pressure bar
    cold = run2d(sc, {"Residual": 1e-6, "Preconditioner": "jacobi"})[0]
    w1, w2 = run2d(sc, {"Residual": 1e-6, "Preconditioner": "jacobi", "WarmStart": 1}, repeat=2)
    np.testing.assert_array_equal(w1["system"]["x"], cold["system"]["x"])
    truth = oracle.project2d(sc, tol=1e-12, mode=oracle.SHK_OR_JACOBI)
    lab = truth["index_map2"] >= 0
    xt = _dense_2d(truth)["x"][lab]
    rel = np.linalg.norm(w2["system"]["x"][lab] - xt) / np.linalg.norm(xt)
    log(f"[gpu] 2D warm start: second solve pressure rel-L2 vs 1e-12 = {rel:.2e}")
    assert rel <= PRESSURE_BAR


def test_2d_laplace_law():
    n, sigma, radius = 64, 0.1, 0.25
    dx = 1.0 / n
    gc = (np.arange(n) + 0.5) * dx
    y, x = np.meshgrid(gc, gc, indexing="ij")
    sc = {"fluid": (np.hypot(x - 0.5, y - 0.5) - radius).astype(np.float32),
          "solid": np.ones((n + 1, n + 1), np.float32), "u": np.zeros((n, n + 1), np.float32),
          "v": np.zeros((n + 1, n), np.float32), "us": np.zeros((n, n + 1), np.float32),
          "vs": np.zeros((n + 1, n), np.float32), "dx": dx}
    g = run2d(sc, {"Residual": 1e-8, "Preconditioner": "jacobi"}, surface_tension=sigma)[0]
    lab = g["system"]["label"].astype(bool)
    p = g["pressure"][lab].astype(np.float64)
    log(f"[gpu] 2D disc: p mean {p.mean():.6f} (sigma/R = {sigma / radius})")
    np.testing.assert_allclose(p.mean(), sigma / radius, rtol=0.01)


def test_2d_host_pointers_and_errors():
    import shiokaze_amd
    sc = _scene(1, (64, 48), 0)
    dev = run2d(sc, {"Residual": 1e-6})[0]
    ny, nx = sc["fluid"].shape
    host = {k: np.ascontiguousarray(sc[k]).copy() for k in FIELDS2}
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    ps = shiokaze_amd.MacPressureSolver2((nx, ny), sc["dx"], config={"Residual": 1e-6, "HostPointers": 1})
    res = shiokaze_amd.ShkResult()
    assert ps.lib.shk_project2d(ps.ctx, 1.0, p(host["u"]), p(host["v"]), p(host["solid"]), p(host["us"]),
                                p(host["vs"]), p(host["fluid"]), 0.0, ctypes.byref(res)) == 0
    np.testing.assert_array_equal(host["u"], dev["velocity"][0])
    np.testing.assert_array_equal(host["v"], dev["velocity"][1])
    # a 3D call on a 2D context, and the 3D-only learned preconditioner, are refused
    bad = shiokaze_amd._lib.SHK_BAD_ARG
    assert ps.lib.shk_project(ps.ctx, 1.0, *([p(host["u"])] * 8), 0.0, ctypes.byref(res)) == bad
    ps.configure("Preconditioner", "unet")
    assert ps.lib.shk_project2d(ps.ctx, 1.0, p(host["u"]), p(host["v"]), p(host["solid"]), p(host["us"]),
                                p(host["vs"]), p(host["fluid"]), 0.0, ctypes.byref(res)) == bad
    ps.close()
    ps3 = shiokaze_amd.MacPressureSolver3((8, 8, 8), 1 / 8)
    assert ps3.lib.shk_project2d(ps3.ctx, 1.0, *([p(host["u"])] * 6), 0.0, ctypes.byref(res)) == bad
    ps3.close()


def test_2d_large_converges():
    """A 2048^2 pool: the size-independent properties (true residual, divergence)."""
    sc = _scene(1, (2048, 2048), 0)
    for precision in ("mixed", "fp64"):
        g = run2d(sc, {"Residual": 1e-6, "Preconditioner": "jacobi", "Precision": precision})[0]
        r = g["result"]
        log(f"[gpu] 2D 2048^2 {precision}: unknowns {r['unknowns']} iters {r['count']} "
            f"true {r['true_relres_inf']:.2e} solve {r['ms_solve']:.1f} ms")
        assert r["status"] == 0 and r["true_relres_inf"] <= 1e-6
        assert g["div_after_fp64"] / g["div_before"] <= 1e-6 * 1.0001


# ---------------------------------------------------------------- multigrid in 2D

def _scaled_2d(system):
    """The Jacobi-scaled 2D system exactly as the library forms it (k_scale_coeff:
    c * sig_cell * sig_low in double, stored as float) from the exported one."""
    lab = system["label"].astype(bool)
    d = system["diag"]
    sig = np.where(lab, 1.0 / np.sqrt(np.where(lab, d, 1.0)), 0.0)
    hx = np.zeros_like(d)
    hz = np.zeros_like(d)
    hx[:, 1:] = system["cx"][:, 1:] * sig[:, 1:] * sig[:, :-1]
    hz[1:, :] = system["cz"][1:, :] * sig[1:, :] * sig[:-1, :]
    return lab.astype(np.float64), hx.astype(np.float32).astype(np.float64), hz.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("case", [(1, (64, 48), 0), (3, (96, 80), 3), (3, (70, 37), 5)], ids=_ids)
def test_2d_mg_vcycle_matches_oracle(case):
    """One 2D V-cycle (2x1x2 aggregation of the nx x 1 x ny layout) against the numpy
    multigrid restatement on the same scaled system held as (ny, 1, nx)."""
    import torch

    from oracle import mg_oracle as MG
    from shiokaze_amd import MacPressureSolver2
    kind, shape, seed = case
    sc = _scene(kind, shape, seed)
    ny, nx = sc["fluid"].shape
    d = _dev(sc)
    with MacPressureSolver2((nx, ny), sc["dx"], config={"Residual": 1e-6, "Preconditioner": "mg"}) as ps:
        res = ps.project(1.0, (d["u"], d["v"]), d["solid"], (d["us"], d["vs"]), d["fluid"])
        system = {k: v.cpu().numpy() for k, v in ps.export_system().items()}
        mask, hx, hz = _scaled_2d(system)
        rng = np.random.default_rng(seed)
        r = (rng.standard_normal((ny, nx)) * mask).astype(np.float32)
        z = ps.mg_apply(torch.from_numpy(r).cuda()).cpu().numpy().astype(np.float64)
    emb = lambda a: a.reshape(ny, 1, nx)  # noqa: E731
    levels = MG.hierarchy(MG.scaled_level0(emb(mask), emb(hx), np.zeros((ny, 1, nx)), emb(hz)))
    ref = MG.vcycle(levels, emb(r.astype(np.float64))).reshape(ny, nx)
    rel = np.linalg.norm(z - ref) / np.linalg.norm(ref)
    log(f"[gpu] 2D mg v-cycle {shape}: levels {[lv[0].shape for lv in levels]}, rel L2 {rel:.2e}, "
        f"projection {res['count']} its")
    assert res["status"] == 0
    # FP32 V-cycle against the FP64 restatement
    assert rel < 1e-5
    assert np.all(z[mask == 0] == 0)


@pytest.mark.parametrize("case", [(1, (256, 192), 0), (3, (256, 256), 2), (3, (70, 37), 5)], ids=_ids)
def test_2d_projection_with_multigrid(oracle, case):
    kind, shape, seed = case
    sc = _scene(kind, shape, seed)
    g = run2d(sc, {"Residual": 1e-6, "Preconditioner": "mg"})[0]
    jac = run2d(sc, {"Residual": 1e-6, "Precision": "mixed", "Preconditioner": "jacobi"})[0]
    truth = oracle.project2d(sc, tol=1e-12, mode=oracle.SHK_OR_JACOBI)
    r, rj = g["result"], jac["result"]
    lab = truth["index_map2"] >= 0
    xt = _dense_2d(truth)["x"][lab]
    rel = np.linalg.norm(g["system"]["x"][lab] - xt) / np.linalg.norm(xt)
    log(f"[gpu] 2D mg {shape}: {r['count']} its (outer {r['outer']}) vs jacobi {rj['count']}, "
        f"relres {r['true_relres_inf']:.2e}, x rel {rel:.2e}")
    assert r["status"] == 0 and r["true_relres_inf"] <= 1e-6
    assert r["count"] * 4 <= rj["count"]
    # at the 2D bar stop the pressure meets the bar
    gb = run2d(sc, {"Residual": BAR_STOP_2D["mg"], "Preconditioner": "mg"})[0]
    relb = np.linalg.norm(gb["system"]["x"][lab] - xt) / np.linalg.norm(xt)
    log(f"[gpu] 2D mg {shape} at {BAR_STOP_2D['mg']:g}: {gb['result']['count']} its, x rel {relb:.2e}")
    assert gb["result"]["status"] == 0
    assert relb <= PRESSURE_BAR
