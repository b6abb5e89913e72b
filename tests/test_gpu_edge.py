"""Edge cases of shk_project / shk_solve_csr through the C-ABI on the GPU:
empty scenes, NaN input, iteration caps, host pointers, CSR parity."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from gpu_helpers import (BAR_STOP, PRESSURE_BAR, iterate_error, iteration_spread, log, run_projection, to_device,
                         within_spread)

pytestmark = pytest.mark.gpu


def test_no_liquid_leaves_velocity_untouched(scenes):
    """macpressuresolver3.cpp:170: with no unknowns nothing is solved or updated."""
    sc = dict(scenes(1, 16, 0))
    sc["fluid"] = np.abs(sc["fluid"]) + 0.01
    g = run_projection(sc, {"Residual": 1e-6})
    assert g["result"]["unknowns"] == 0
    assert g["result"]["status"] == 0
    for a, k in enumerate("uvw"):
        np.testing.assert_array_equal(g["velocity"][a], sc[k])


def test_nan_velocity_reports_nan(scenes):
    """pcg_solver.h:267-271: a NaN rho stops the solve with 0 iterations."""
    import shiokaze_amd
    sc = dict(scenes(1, 16, 0))
    sc["u"] = sc["u"].copy()
    sc["u"][4, 2, 5] = np.nan
    d = to_device(sc)
    with shiokaze_amd.MacPressureSolver3(sc["shape"], sc["dx"], config={"Residual": 1e-6}) as ps:
        res = ps.project(1.0, (d["u"], d["v"], d["w"]), d["solid"], (d["us"], d["vs"], d["ws"]), d["fluid"])
    log(f"[gpu] nan case: status={res['status_name']} count={res['count']}")
    assert res["status"] == shiokaze_amd._lib.SHK_NAN
    assert res["count"] == 0


def test_iteration_cap(scenes, oracle):
    """pcg_solver.h:292-294: the cap returns the iterations run and the residual."""
    sc = scenes(1, 24, 0)
    g = run_projection(sc, {"Residual": 1e-6, "MaxIterations": 7, "Preconditioner": "none"})
    o = oracle.project(sc, tol=1e-6, max_iterations=7, mode=oracle.SHK_OR_CG)
    r = g["result"]
    assert o["status"] == 1
    assert r["status"] == 1 and r["count"] == 7
    assert r["reresid"] == pytest.approx(o["reresid"], rel=1e-6)


def test_host_pointers_match_device_pointers(scenes):
    import shiokaze_amd
    sc = scenes(3, 24, 11)
    g = run_projection(sc, {"Residual": 1e-6})
    nx, ny, nz = sc["shape"]
    host = {k: np.ascontiguousarray(sc[k]).copy() for k in ("fluid", "solid", "u", "v", "w", "us", "vs", "ws")}
    ps = shiokaze_amd.MacPressureSolver3(sc["shape"], sc["dx"], config={"Residual": 1e-6, "HostPointers": 1})
    lib = ps.lib
    res = shiokaze_amd.ShkResult()
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    rc = lib.shk_project(ps.ctx, 1.0, p(host["u"]), p(host["v"]), p(host["w"]), p(host["solid"]), p(host["us"]),
                         p(host["vs"]), p(host["ws"]), p(host["fluid"]), 0.0, ctypes.byref(res))
    assert rc == 0
    pres = np.empty((nz, ny, nx), dtype=np.float32)
    assert lib.shk_copy_pressure(ps.ctx, p(pres)) == 0
    ps.close()
    assert res.count == g["result"]["count"]
    np.testing.assert_array_equal(pres, g["pressure"])
    for a, k in enumerate("uvw"):
        np.testing.assert_array_equal(host[k], g["velocity"][a])


def test_configure_rejects_bad_keys():
    import shiokaze_amd
    ps = shiokaze_amd.MacPressureSolver3((8, 8, 8), 1 / 8)
    for key, val in [("NoSuchKey", "1"), ("Residual", "abc"), ("WarmStart", "maybe"), ("Gain", "x"),
                     ("Precision", "fp16"), ("Preconditioner", "ilu")]:
        with pytest.raises(shiokaze_amd.ShkError):
            ps.configure(key, val)
    ps.configure("PCG.Residual", "1e-8")     # scoped names as the reference's cmdparser accepts
    ps.configure("MacUtility.EpsFluid", "1e-3")
    ps.close()


@pytest.mark.parametrize("precond", ["none", "jacobi"])
def test_csr_solve_parity(scenes, oracle, precond):
    """shk_solve_csr (the RCMatrix_solver_interface path) on the oracle-assembled matrix."""
    import torch

    import shiokaze_amd
    sc = scenes(3, 32, 5)
    mode = oracle.SHK_OR_CG if precond == "none" else oracle.SHK_OR_JACOBI
    o = oracle.project(sc, tol=1e-6, mode=mode)
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dtype=dt).cuda()  # noqa: E731
    solver = shiokaze_amd.PCGSolver(config={"Residual": 1e-6, "Preconditioner": precond})
    x, res = solver.solve(t(o["rowptr"], torch.int64), t(o["cols"], torch.int64), t(o["vals"], torch.float64),
                          t(o["rhs"], torch.float64))
    solver.close()
    log(f"[gpu] csr {precond}: iters={res['count']} oracle={o['iters']} true={res['true_relres_inf']:.3e}")
    spread = iteration_spread(oracle, o["rowptr"], o["cols"], o["vals"], o["rhs"], 1e-6, mode)
    log(f"[gpu] csr oracle renumbered spread={spread}")
    assert res["status"] == 0
    assert within_spread(res["count"], spread)
    assert res["true_relres_inf"] <= 1e-6
    # the oracle's iterate after the same number of iterations: equal up to rounding
    # order, which moves late CG iterates by a fraction of their distance from the solution
    xk, k, _, _ = oracle.pcg(o["rowptr"], o["cols"], o["vals"], o["rhs"], 1e-30, res["count"], mode)
    xg = x.cpu().numpy()
    rel = np.linalg.norm(xg - xk) / np.linalg.norm(xk)
    err = iterate_error(oracle, o["rowptr"], o["cols"], o["vals"], o["rhs"], xk)
    log(f"[gpu] csr {precond}: x vs the oracle's iterate {k}: {rel:.2e} (its distance from the solution {err:.2e})")
    assert rel <= 0.5 * err
    # and at the bar stop, the pressure bar against the oracle's FP64 solve to 1e-12
    solver = shiokaze_amd.PCGSolver(config={"Residual": BAR_STOP[precond], "Preconditioner": precond})
    xb, rb = solver.solve(t(o["rowptr"], torch.int64), t(o["cols"], torch.int64), t(o["vals"], torch.float64),
                          t(o["rhs"], torch.float64))
    solver.close()
    assert rb["status"] == 0
    xt, _, _, st, _ = oracle.pcg_threaded(o["rowptr"], o["cols"], o["vals"], o["rhs"], 1e-12, 100000,
                                         oracle.SHK_OR_JACOBI)
    relb = np.linalg.norm(xb.cpu().numpy() - xt) / np.linalg.norm(xt)
    log(f"[gpu] csr {precond} at {BAR_STOP[precond]:g}: {rb['count']} its, pressure rel-L2 vs 1e-12 {relb:.2e}")
    assert relb <= PRESSURE_BAR


@pytest.mark.parametrize("precond", ["unet", "mg"])
def test_csr_solve_refuses_grid_preconditioners(scenes, oracle, precond):
    """The learned and the brick multigrid preconditioners need the projection's grid;
    shk_solve_csr refuses them with SHK_BAD_ARG instead of running another method."""
    import torch

    import shiokaze_amd
    o = oracle.project(scenes(1, 16, 0), tol=1e-6, mode=oracle.SHK_OR_CG)
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dtype=dt).cuda()  # noqa: E731
    solver = shiokaze_amd.PCGSolver(config={"Residual": 1e-6, "Preconditioner": precond})
    with pytest.raises(shiokaze_amd.ShkError):
        solver.solve(t(o["rowptr"], torch.int64), t(o["cols"], torch.int64), t(o["vals"], torch.float64),
                     t(o["rhs"], torch.float64))
    solver.close()
