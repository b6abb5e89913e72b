"""GPU parity of the projection (shk_project through the C-ABI) against the CPU oracle.

Bit-exact: face fractions, unknown labels, stencil coefficients, rhs, and the
velocity update given the same pressure.  Within tolerance: the Krylov solution
(iteration counts inside the spread that rounding order alone gives the oracle,
see gpu_helpers.iteration_spread; solution relative L2; the FP64
post-projection divergence)."""
from __future__ import annotations

import numpy as np
import pytest

from gpu_helpers import (BAR_STOP, assert_pressure_bar, iterate_error, iteration_spread, log, oracle_dense_system,
                         oracle_truth, pressure_rel_l2, run_projection, within_spread)

pytestmark = pytest.mark.gpu

CASES = [(1, 32, 0), (3, 32, 7), (3, (24, 40, 16), 3), (1, (13, 21, 9), 0)]


def _ids(c):
    return f"kind{c[0]}-n{c[1]}-seed{c[2]}".replace(" ", "")


@pytest.fixture(scope="module")
def projected(scenes, oracle):
    cache = {}

    def get(kind, n, seed, precond="none", precision="fp64", tol=1e-6):
        key = (kind, str(n), seed, precond, precision, tol)
        if key not in cache:
            sc = scenes(kind, n, seed)
            log(f"[gpu] project kind={kind} n={n} seed={seed} precond={precond} precision={precision} tol={tol}")
            g = run_projection(sc, {"Residual": tol, "Preconditioner": precond, "Precision": precision})
            mode = {"none": oracle.SHK_OR_CG, "jacobi": oracle.SHK_OR_JACOBI}[precond]
            o = oracle.project(sc, tol=tol, mode=mode)
            cache[key] = (sc, g, o)
        return cache[key]

    return get


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_fractions_bitexact(projected, case):
    sc, g, o = projected(*case)
    for a in range(3):
        np.testing.assert_array_equal(g["areas"][a].astype(np.float64), o["areas"][a])
        np.testing.assert_array_equal(g["rhos"][a].astype(np.float64), o["rhos"][a])


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_system_bitexact(projected, oracle, case):
    sc, g, o = projected(*case)
    assert g["result"]["unknowns"] == o["n"]
    s = g["system"]
    np.testing.assert_array_equal(s["label"].astype(bool), o["index_map"] >= 0)
    ref = oracle_dense_system(oracle, o)
    lab = o["index_map"] >= 0
    np.testing.assert_array_equal(s["diag"][lab], ref["diag"][lab])
    np.testing.assert_array_equal(s["rhs"][lab], ref["rhs"][lab])
    # off-diagonals exist where both cells are unknowns
    nz, ny, nx = lab.shape
    both_x = lab[:, :, 1:] & lab[:, :, :-1]
    both_y = lab[:, 1:, :] & lab[:, :-1, :]
    both_z = lab[1:, :, :] & lab[:-1, :, :]
    np.testing.assert_array_equal(s["cx"][:, :, 1:][both_x], ref["cx"][:, :, 1:][both_x])
    np.testing.assert_array_equal(s["cy"][:, 1:, :][both_y], ref["cy"][:, 1:, :][both_y])
    np.testing.assert_array_equal(s["cz"][1:, :, :][both_z], ref["cz"][1:, :, :][both_z])


@pytest.mark.parametrize("precond", ["none", "jacobi"])
@pytest.mark.parametrize("case", CASES[:2], ids=_ids)
def test_solve_parity(projected, oracle, case, precond):
    sc, g, o = projected(*case, precond=precond)
    r = g["result"]
    mode = oracle.SHK_OR_CG if precond == "none" else oracle.SHK_OR_JACOBI
    spread = iteration_spread(oracle, o["rowptr"], o["cols"], o["vals"], o["rhs"], 1e-6, mode)
    log(f"[gpu] iters gpu={r['count']} oracle={o['iters']} oracle renumbered={spread} relres={r['reresid']:.3e} "
        f"true={r['true_relres_inf']:.3e}")
    assert r["status"] == 0
    assert within_spread(r["count"], spread)
    assert r["true_relres_inf"] <= 1e-6 * 1.0001
    # the same algorithm after the same number of iterations: the GPU's iterate and the
    # oracle's differ only by rounding order
    xk, k, _, _ = oracle.pcg(o["rowptr"], o["cols"], o["vals"], o["rhs"], 1e-30, r["count"], mode)
    assert k == r["count"]
    lab = o["index_map"] >= 0
    x = g["system"]["x"][lab]
    xo = xk[o["index_map"][lab]]
    rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    err = iterate_error(oracle, o["rowptr"], o["cols"], o["vals"], o["rhs"], xk)
    log(f"[gpu] x relL2 vs the oracle's iterate {k} = {rel:.3e} (its distance from the solution {err:.2e})")
    assert rel <= 0.5 * err


@pytest.mark.parametrize("precond,precision", [("none", "fp64"), ("jacobi", "fp64"), ("none", "mixed"),
                                               ("jacobi", "mixed")])
@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_pressure_bar(scenes, oracle, case, precond, precision):
    """Every FP64 and mixed-precision path meets the pressure bar (rel-L2 <= 1e-5 against
    the oracle's FP64 solve to 1e-12) at its bar stop, and the FP64 post-projection
    divergence bar (true relative residual <= 1e-6)."""
    sc = scenes(*case)
    stop = BAR_STOP[precond]
    g = run_projection(sc, {"Residual": stop, "Preconditioner": precond, "Precision": precision})
    r = g["result"]
    assert r["status"] == 0 and r["true_relres_inf"] <= stop * 1.0001
    assert g["div_after_fp64"] / g["div_before"] <= 1e-6
    truth = oracle_truth(oracle, sc, ("3d",) + tuple(map(str, case)))
    assert_pressure_bar(g["system"]["x"], truth, f"{_ids(case)} {precond}/{precision} stop {stop:g} "
                        f"({r['count']} its, {r['outer']} outer)")


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_velocity_update_bitexact(projected, oracle, case):
    """The update is checked bit for bit by feeding the GPU's pressure to the oracle."""
    sc, g, o = projected(*case)
    p = g["pressure"].astype(np.float64)
    vel = oracle.update_velocity(1.0, sc["dx"], p, o["areas"], o["rhos"], [sc["us"], sc["vs"], sc["ws"]],
                                 [sc["u"], sc["v"], sc["w"]], real_is_double=False)
    for a in range(3):
        np.testing.assert_array_equal(g["velocity"][a].astype(np.float64), vel[a])
    # and the pressure is the float rounding of the FP64 solution on unknowns, 0 elsewhere
    lab = o["index_map"] >= 0
    np.testing.assert_array_equal(g["pressure"][lab], g["system"]["x"][lab].astype(np.float32))
    assert np.all(g["pressure"][~lab] == 0)


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_post_divergence(projected, case):
    """FP64 post-projection divergence relative to the initial one equals the true
    relative residual (D(u_new) = -(b - A p)) and meets the 1e-6 bar."""
    sc, g, o = projected(*case)
    r = g["result"]
    rel = g["div_after_fp64"] / g["div_before"]
    log(f"[gpu] div before={g['div_before']:.3e} after(fp64)={g['div_after_fp64']:.3e} rel={rel:.3e} "
        f"after(f32 storage)={g['div_after_f32']:.3e}")
    assert g["div_before"] == pytest.approx(r["b_inf"], rel=1e-12)
    assert rel <= 1e-6 * 1.0001
    assert rel == pytest.approx(r["true_relres_inf"], rel=1e-6, abs=1e-12)


@pytest.mark.parametrize("case", CASES[:2], ids=_ids)
def test_mixed_precision(projected, oracle, scenes, case):
    """Mixed precision at the reference's stop (1e-6) meets the residual and the
    divergence bar; at the bar stop it meets the pressure bar too."""
    sc, g, o = projected(*case, precond="jacobi", precision="mixed")
    r = g["result"]
    truth = oracle_truth(oracle, sc, ("3d",) + tuple(map(str, case)))
    rel = pressure_rel_l2(g["system"]["x"], truth)
    log(f"[gpu] mixed at 1e-6: iters={r['count']} outer={r['outer']} true relres={r['true_relres_inf']:.3e} "
        f"p relL2 vs 1e-12 oracle={rel:.3e}")
    assert r["status"] == 0
    assert r["true_relres_inf"] <= 1e-6
    sc, g, o = projected(*case, precond="jacobi", precision="mixed", tol=BAR_STOP["jacobi"])
    assert g["result"]["status"] == 0
    assert_pressure_bar(g["system"]["x"], truth, f"mixed jacobi at {BAR_STOP['jacobi']:g}")


def test_pool_sphere_64_fp64_plain_cg_against_oracle(scenes, oracle):
    """The small plain-method workload: pool+sphere at 64^3, FP64 plain CG (the
    reference's effective iteration) to 1e-6 on the GPU against the oracle's CG on the
    CPU: iteration count within the renumbering spread, the iterate equal to the
    oracle's after the same count up to rounding order, the post-projection divergence
    bar, and the pressure bar at the bar stop."""
    sc = scenes(1, 64, 0)
    g = run_projection(sc, {"Residual": 1e-6, "Preconditioner": "none", "Precision": "fp64"})
    o = oracle.project(sc, tol=1e-6, mode=oracle.SHK_OR_CG)
    r = g["result"]
    assert r["status"] == 0 and r["unknowns"] == o["n"]
    spread = iteration_spread(oracle, o["rowptr"], o["cols"], o["vals"], o["rhs"], 1e-6, oracle.SHK_OR_CG, n_perm=4)
    log(f"[gpu] 64^3 plain CG: gpu {r['count']} its, oracle {o['iters']}, renumbered {spread}")
    assert within_spread(r["count"], spread)
    xk, k, _, _ = oracle.pcg(o["rowptr"], o["cols"], o["vals"], o["rhs"], 1e-30, r["count"], oracle.SHK_OR_CG)
    lab = o["index_map"] >= 0
    rel = np.linalg.norm(g["system"]["x"][lab] - xk[o["index_map"][lab]]) / np.linalg.norm(xk)
    err = iterate_error(oracle, o["rowptr"], o["cols"], o["vals"], o["rhs"], xk)
    log(f"[gpu] 64^3 plain CG: x vs the oracle's iterate {k}: {rel:.3e} (its distance from the solution {err:.2e})")
    assert rel <= 0.5 * err
    assert g["div_after_fp64"] / g["div_before"] <= 1e-6
    truth = oracle_truth(oracle, sc, ("3d", "1", "64", "0"))
    at_stop = pressure_rel_l2(g["system"]["x"], truth)
    log(f"[gpu] 64^3 plain CG at 1e-6: pressure rel-L2 vs 1e-12 = {at_stop:.3e}")
    gb = run_projection(sc, {"Residual": BAR_STOP["none"], "Preconditioner": "none", "Precision": "fp64"})
    assert gb["result"]["status"] == 0
    assert_pressure_bar(gb["system"]["x"], truth, f"64^3 plain CG at {BAR_STOP['none']:g}")
