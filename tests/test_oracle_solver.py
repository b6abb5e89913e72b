"""Pins the oracle's solver (pcg_solver.h / pcg.cpp semantics) and the whole
projection against direct solves, CG identities, MIC(0) properties, the
reference's status rules and the accuracytest3 known answer
(src/examples/accuracytest3-example.cpp: p = |x-c|^2 - r^2, second order)."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURES = sorted(GOLDEN.glob("*.npz"))


@pytest.fixture(params=FIXTURES, ids=lambda p: p.stem)
def fixture(request):
    return dict(np.load(request.param, allow_pickle=False))


def _scene_from(fx):
    sc = {k: fx[k] for k in ("fluid", "solid", "u", "v", "w", "us", "vs", "ws")}
    sc["dx"] = float(fx["dx"])
    nz, ny, nx = sc["fluid"].shape
    sc["shape"] = (nx, ny, nx) if False else (nx, ny, nz)
    return sc


def test_assembly_reproduces_fixture(oracle, fixture):
    o = oracle.project(_scene_from(fixture), tol=1e-12, mode=oracle.SHK_OR_JACOBI)
    assert np.array_equal(o["rowptr"], fixture["rowptr"])
    assert np.array_equal(o["cols"], fixture["cols"])
    assert np.array_equal(o["vals"], fixture["vals"])
    assert np.array_equal(o["rhs"], fixture["rhs"])


@pytest.mark.parametrize("mode", [0, 1, 2, 3], ids=["reference", "cg", "jacobi", "mic0"])
def test_pcg_matches_direct_solve(oracle, fixture, mode):
    x, it, rr, st = oracle.pcg(fixture["rowptr"], fixture["cols"], fixture["vals"], fixture["rhs"], tol=1e-13,
                               mode=mode)
    assert st == 0 and rr <= 1e-13
    xd = fixture["x_direct"]
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9


def test_reference_mode_iterates_are_plain_cg(oracle, fixture):
    """pcg_solver.h:379-384 overwrites the MIC(0) result with r: the reference's
    iterates are plain CG, bit for bit, whatever the discarded work."""
    a = oracle.pcg(fixture["rowptr"], fixture["cols"], fixture["vals"], fixture["rhs"], tol=1e-8,
                   mode=oracle.SHK_OR_REFERENCE)
    b = oracle.pcg(fixture["rowptr"], fixture["cols"], fixture["vals"], fixture["rhs"], tol=1e-8,
                   mode=oracle.SHK_OR_CG)
    assert a[1] == b[1] and a[2] == b[2]
    assert np.array_equal(a[0], b[0])


def test_stop_rule_and_outputs(oracle, fixture):
    A = oracle.to_scipy(fixture["rowptr"], fixture["cols"], fixture["vals"])
    b = fixture["rhs"]
    for tol in (1e-4, 1e-6):
        x, it, rr, st = oracle.pcg(fixture["rowptr"], fixture["cols"], fixture["vals"], b, tol=tol,
                                   mode=oracle.SHK_OR_CG)
        assert st == 0 and rr <= tol
        # the recursive residual tracks the true one
        true = np.abs(b - A @ x).max() / np.abs(b).max()
        assert true == pytest.approx(rr, rel=1e-3, abs=1e-12)
        # one iteration less would not have met the tolerance
        _, _, rr1, st1 = oracle.pcg(fixture["rowptr"], fixture["cols"], fixture["vals"], b, tol=tol,
                                    max_iterations=it - 1, mode=oracle.SHK_OR_CG)
        assert st1 == 1 and rr1 > tol


def test_status_rules(oracle, fixture):
    rp, c, v, b = fixture["rowptr"], fixture["cols"], fixture["vals"], fixture["rhs"]
    # zero rhs: 0 iterations, converged (pcg_solver.h:254-257)
    x, it, rr, st = oracle.pcg(rp, c, v, np.zeros_like(b), tol=1e-6)
    assert (it, st) == (0, 0) and np.all(x == 0)
    # NaN: 0 iterations (pcg_solver.h:267-271)
    bn = b.copy()
    bn[3] = np.nan
    _, it, _, st = oracle.pcg(rp, c, v, bn, tol=1e-6)
    assert (it, st) == (0, 2)
    # cap: the iterations run (pcg_solver.h:292-294)
    _, it, rr, st = oracle.pcg(rp, c, v, b, tol=1e-12, max_iterations=5)
    assert (it, st) == (5, 1) and rr > 1e-12


def test_mic0_factor_properties(oracle, fixture):
    """omega = 0: L L^T equals A on A's pattern (IC(0)).  omega = 1: L L^T has A's
    row sums (the modification moves every dropped fill onto the diagonal)."""
    import scipy.sparse as sp
    rp, c, v = fixture["rowptr"], fixture["cols"], fixture["vals"]
    A = oracle.to_scipy(rp, c, v).toarray()
    n = A.shape[0]
    for omega in (0.0, 1.0):
        cs, ri, val, invd = oracle.mic0(rp, c, v, omega=omega, min_diag_ratio=0.0)
        cols = np.repeat(np.arange(n), np.diff(cs))
        L = sp.csr_matrix((val, (ri, cols)), shape=(n, n)).toarray() + np.diag(1.0 / invd)
        LLt = L @ L.T
        if omega == 0.0:
            mask = A != 0
            assert np.allclose(LLt[mask], A[mask], rtol=1e-10, atol=1e-10 * np.abs(A).max())
        else:
            assert np.allclose(LLt.sum(axis=1), A.sum(axis=1), rtol=1e-9, atol=1e-9 * np.abs(A).max())


def test_preconditioners_reduce_iterations(oracle, fixture):
    its = {m: oracle.pcg(fixture["rowptr"], fixture["cols"], fixture["vals"], fixture["rhs"], tol=1e-8, mode=m)[1]
           for m in (1, 2, 3)}
    assert its[3] < its[2] < its[1]


def test_iteration_spread_under_renumbering(oracle, scenes):
    """CG's max-norm stop point moves with rounding order alone: random symmetric
    renumberings of one system stop at different counts (the band GPU parity uses)."""
    import sys
    sys.path.insert(0, str(Path(__file__).resolve().parent))
    from gpu_helpers import iteration_spread
    sc = scenes(3, 24, 7)
    o = oracle.project(sc, tol=1e-6, mode=oracle.SHK_OR_CG)
    lo, hi = iteration_spread(oracle, o["rowptr"], o["cols"], o["vals"], o["rhs"], 1e-6, oracle.SHK_OR_CG)
    assert lo <= o["iters"] <= hi
    assert hi - lo >= 1


# ---------------------------------------------------------------- accuracytest3

def _accuracy_scene(n, r):
    """accuracytest3-example.cpp:92-104: u = grad(|x-c|^2) on faces, phi = |x-c| - r
    at cell centres, no solid, dx = 1/n."""
    dx = 1.0 / n
    c = 0.5
    def faces(dim):
        shp = [n, n, n]
        shp[2 - dim] += 1
        k, j, i = np.meshgrid(*[np.arange(s) for s in shp], indexing="ij")
        pos = [i + 0.5 * (dim != 0), j + 0.5 * (dim != 1), k + 0.5 * (dim != 2)]
        return 2.0 * (pos[dim] * dx - c)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    d = np.sqrt(((i + 0.5) * dx - c) ** 2 + ((j + 0.5) * dx - c) ** 2 + ((k + 0.5) * dx - c) ** 2) - r
    sc = {"fluid": d, "solid": np.ones((n + 1, n + 1, n + 1)), "u": faces(0), "v": faces(1), "w": faces(2),
          "dx": dx, "shape": (n, n, n)}
    sc["us"], sc["vs"], sc["ws"] = (np.zeros_like(sc["u"]), np.zeros_like(sc["v"]), np.zeros_like(sc["w"]))
    exact = ((i + 0.5) * dx - c) ** 2 + ((j + 0.5) * dx - c) ** 2 + ((k + 0.5) * dx - c) ** 2 - r * r
    return sc, exact


def test_accuracytest3_second_order(oracle):
    """Max-norm pressure error against p = |x-c|^2 - r^2 over radii r = 0.4 + q*sqrt(3)dx/4
    (accuracytest3-example.cpp:68-139 with Real = double, EpsFluid = 1e-18)."""
    errs = {}
    for n in (16, 32, 64):
        dx = 1.0 / n
        worst = 0.0
        for q in (-4, -2, 0, 1, 3):
            r = 0.4 + np.sqrt(3.0) * dx / 4 * q
            sc, exact = _accuracy_scene(n, r)
            o = oracle.project(sc, dt=1.0, tol=1e-13, mode=oracle.SHK_OR_JACOBI, real_is_double=True,
                               eps_fluid=1e-18)
            lab = o["index_map"] >= 0
            worst = max(worst, np.abs(o["pressure"][lab] - exact[lab]).max())
        errs[n] = worst
    o1 = np.log2(errs[16] / errs[32])
    o2 = np.log2(errs[32] / errs[64])
    # 16^3 is pre-asymptotic (the sphere spans 13 cells); the finer pair is second order
    assert errs[64] < 2e-4
    assert o1 > 1.4 and o2 > 1.8, (errs, o1, o2)
    assert 0.5 * np.log2(errs[16] / errs[64]) > 1.6
