"""The multigrid preconditioner (Preconditioner=mg) on the GPU: one V-cycle against
the numpy restatement (oracle/mg_oracle.py) on the exported scaled system, its
symmetry, and whole projections that reach the residual asked for in far fewer
iterations than Jacobi-PCG."""
from __future__ import annotations

import numpy as np
import pytest

from gpu_helpers import BAR_STOP, assert_pressure_bar, log, oracle_truth, run_projection, to_device
from oracle import mg_oracle as MG

pytestmark = pytest.mark.gpu


def _solver_after_projection(sc, config):
    import shiokaze_amd
    d = to_device(sc)
    ps = shiokaze_amd.MacPressureSolver3(sc["shape"], sc["dx"], config=config)
    res = ps.project(1.0, (d["u"], d["v"], d["w"]), d["solid"], (d["us"], d["vs"], d["ws"]), d["fluid"])
    return ps, res


CASES = [(3, 32, 2), (1, (13, 21, 9), 0), (3, (24, 40, 16), 3)]


@pytest.mark.parametrize("kind,n,seed", CASES)
def test_vcycle_matches_oracle(scenes, kind, n, seed):
    import torch
    log(f"[gpu] mg v-cycle kind={kind} n={n} seed={seed}")
    sc = scenes(kind, n, seed)
    ps, res = _solver_after_projection(sc, {"Residual": 1e-6, "Preconditioner": "mg"})
    try:
        s = {k: v.cpu().numpy().astype(np.float64) for k, v in ps.export_scaled().items()}
        levels = MG.hierarchy(MG.scaled_level0(s["mask"], s["hx"], s["hy"], s["hz"]))
        rng = np.random.default_rng(seed)
        r = (rng.standard_normal(s["mask"].shape) * s["mask"]).astype(np.float32)
        z = ps.mg_apply(torch.from_numpy(r).cuda()).cpu().numpy().astype(np.float64)
    finally:
        ps.close()
    ref = MG.vcycle(levels, r.astype(np.float64))
    rel = np.linalg.norm(z - ref) / np.linalg.norm(ref)
    log(f"[gpu]   levels {[lv[0].shape for lv in levels]}: rel L2 error {rel:.2e}")
    # FP32 V-cycle against the FP64 restatement: a few float roundings per level
    assert rel < 1e-5
    assert np.abs(z - ref).max() <= 1e-3 * np.abs(ref).max()
    assert np.all(z[s["mask"] == 0] == 0)


def test_vcycle_after_another_scene_on_the_same_context(scenes):
    """Coarse cells without a fine brick are written only by the per-projection
    coarsening (zeros); a context that projected another scene before, with another
    brick set and then the same one, must give the V-cycle of the new system."""
    import torch

    import shiokaze_amd
    first, second = scenes(1, 40, 0), scenes(3, 40, 6)
    cfg = {"Residual": 1e-6, "Preconditioner": "mg"}
    ps = shiokaze_amd.MacPressureSolver3(second["shape"], second["dx"], config=cfg)
    try:
        for sc in (first, second, first, second):
            d = to_device(sc)
            ps.project(1.0, (d["u"], d["v"], d["w"]), d["solid"], (d["us"], d["vs"], d["ws"]), d["fluid"])
        s = {k: v.cpu().numpy().astype(np.float64) for k, v in ps.export_scaled().items()}
        r = (np.random.default_rng(7).standard_normal(s["mask"].shape) * s["mask"]).astype(np.float32)
        z = ps.mg_apply(torch.from_numpy(r).cuda()).cpu().numpy().astype(np.float64)
    finally:
        ps.close()
    ref = MG.vcycle(MG.hierarchy(MG.scaled_level0(s["mask"], s["hx"], s["hy"], s["hz"])), r.astype(np.float64))
    rel = np.linalg.norm(z - ref) / np.linalg.norm(ref)
    log(f"[gpu] mg v-cycle after another scene: rel L2 error {rel:.2e}")
    assert rel < 1e-5
    assert np.all(z[s["mask"] == 0] == 0)


def test_vcycle_is_symmetric(scenes):
    import torch
    sc = scenes(3, 32, 4)
    ps, _ = _solver_after_projection(sc, {"Residual": 1e-6, "Preconditioner": "mg"})
    try:
        m = ps.export_scaled()["mask"]
        g = torch.Generator(device="cuda").manual_seed(1)
        u = torch.randn(m.shape, device="cuda", generator=g) * m
        v = torch.randn(m.shape, device="cuda", generator=g) * m
        Mu, Mv = ps.mg_apply(u.contiguous()), ps.mg_apply(v.contiguous())
    finally:
        ps.close()
    a, b = float((Mu.double() * v.double()).sum()), float((u.double() * Mv.double()).sum())
    log(f"[gpu] mg symmetry: <Mu,v> {a:.8e} <u,Mv> {b:.8e}")
    assert abs(a - b) <= 1e-4 * abs(a)
    assert float((Mu.double() * u.double()).sum()) > 0


@pytest.mark.parametrize("kind,n,seed", [(1, 32, 0), (3, (24, 40, 16), 3), (3, 64, 5), (1, (13, 21, 9), 0)])
def test_projection_with_multigrid(scenes, oracle, kind, n, seed):
    """Multigrid PCG reaches the reference's stop in at most a quarter of Jacobi-PCG's
    iterations, and at its bar stop meets the pressure bar against the oracle."""
    log(f"[gpu] mg projection kind={kind} n={n} seed={seed}")
    sc = scenes(kind, n, seed)
    g = run_projection(sc, {"Residual": 1e-6, "Preconditioner": "mg"}, want_system=False)
    jac = run_projection(sc, {"Residual": 1e-6, "Precision": "mixed", "Preconditioner": "jacobi"}, want_system=False)
    r, rj = g["result"], jac["result"]
    log(f"[gpu]   mg {r['count']} it (outer {r['outer']}), jacobi {rj['count']} it, relres {r['true_relres_inf']:.2e}")
    assert r["status"] == 0 and r["true_relres_inf"] <= 1e-6
    assert r["count"] * 4 <= rj["count"]
    # the velocity update is the same code path as every other preconditioner
    for a in range(3):
        assert np.isfinite(g["velocity"][a]).all()
    gb = run_projection(sc, {"Residual": BAR_STOP["mg"], "Preconditioner": "mg"})
    rb = gb["result"]
    assert rb["status"] == 0 and gb["div_after_fp64"] / gb["div_before"] <= 1e-6
    truth = oracle_truth(oracle, sc, ("3d", str(kind), str(n), str(seed)))
    assert_pressure_bar(gb["system"]["x"], truth, f"mg kind={kind} n={n} at {BAR_STOP['mg']:g} ({rb['count']} its)")


def test_multigrid_reports_nan(scenes):
    import shiokaze_amd
    sc = dict(scenes(1, 16, 0))
    sc["u"] = sc["u"].copy()
    sc["u"][4, 2, 5] = np.nan
    ps, res = _solver_after_projection(sc, {"Residual": 1e-6, "Preconditioner": "mg"})
    ps.close()
    assert res["status"] == shiokaze_amd._lib.SHK_NAN


def test_multigrid_iteration_cap(scenes):
    import shiokaze_amd
    sc = scenes(3, 32, 2)
    ps, res = _solver_after_projection(sc, {"Residual": 1e-8, "Preconditioner": "mg", "MaxIterations": 3})
    ps.close()
    assert res["status"] == shiokaze_amd._lib.SHK_NOT_CONVERGED
    assert res["count"] == 3


def test_multigrid_is_deterministic_and_host_pointers_agree(scenes):
    import ctypes

    import shiokaze_amd
    sc = scenes(3, (24, 40, 16), 3)
    cfg = {"Residual": 1e-6, "Preconditioner": "mg"}
    a = run_projection(sc, cfg, want_system=False)
    b = run_projection(sc, cfg, want_system=False)
    np.testing.assert_array_equal(a["pressure"], b["pressure"])
    assert a["result"]["count"] == b["result"]["count"]
    nx, ny, nz = sc["shape"]
    host = {k: np.ascontiguousarray(sc[k]).copy() for k in ("fluid", "solid", "u", "v", "w", "us", "vs", "ws")}
    ps = shiokaze_amd.MacPressureSolver3(sc["shape"], sc["dx"], config=dict(cfg, HostPointers=1))
    res = shiokaze_amd.ShkResult()
    p = lambda arr: arr.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    assert ps.lib.shk_project(ps.ctx, 1.0, p(host["u"]), p(host["v"]), p(host["w"]), p(host["solid"]),
                              p(host["us"]), p(host["vs"]), p(host["ws"]), p(host["fluid"]), 0.0, ctypes.byref(res)) == 0
    pres = np.empty((nz, ny, nx), dtype=np.float32)
    assert ps.lib.shk_copy_pressure(ps.ctx, p(pres)) == 0
    ps.close()
    assert res.count == a["result"]["count"]
    np.testing.assert_array_equal(pres, a["pressure"])
