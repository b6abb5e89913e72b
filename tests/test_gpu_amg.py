"""shk_solve_csr with Preconditioner=mg: CG with the aggregation-multigrid V-cycle
built from the CSR matrix alone (the counterpart of the reference's amg LinSolver,
src/linsolver/amg.cpp:43-166), against the numpy restatement oracle/amg_oracle.py
(the same aggregation, Galerkin products and V-cycle): iteration counts, the iterate,
the pressure bar against the oracle's FP64 solve to 1e-12, and a matrix that is not a
grid (a random geometric graph Laplacian)."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from gpu_helpers import PRESSURE_BAR, log

pytestmark = pytest.mark.gpu

BAR_STOP_MG_CSR = 1e-7


def _solve(A, b, config, noalloc=False):
    import torch

    import shiokaze_amd
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dtype=dt).cuda()  # noqa: E731
    solver = shiokaze_amd.PCGSolver(config=config)
    x, res = solver.solve(t(A.indptr.astype(np.int64), torch.int64), t(A.indices.astype(np.int64), torch.int64),
                          t(A.data, torch.float64), t(b, torch.float64))
    solver.close()
    return x.cpu().numpy(), res


def _system(oracle, scenes, case):
    sy = oracle.system(scenes(*case))
    n = len(sy["rhs"])
    A = sp.csr_matrix((sy["vals"], sy["cols"], sy["rowptr"]), shape=(n, n))
    A.sort_indices()
    return A, sy["rhs"]


@pytest.mark.parametrize("case", [(1, 32, 0), (3, 32, 5), (3, 48, 2)])
def test_csr_multigrid_matches_oracle(oracle, scenes, case):
    from oracle import amg_oracle as G
    A, b = _system(oracle, scenes, case)
    x, res = _solve(A, b, {"Residual": 1e-6, "Preconditioner": "mg"})
    xo, it_o = G.pcg(A, b, tol=1e-6)
    jac = oracle.pcg(A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data, b, 1e-6, 100000,
                     oracle.SHK_OR_JACOBI)[1]
    log(f"[gpu] csr mg {case}: {A.shape[0]} rows, {res['count']} its (oracle {it_o}, Jacobi-PCG {jac}), "
        f"true relres {res['true_relres_inf']:.2e}, {res['ms_solve']:.2f} ms")
    assert res["status"] == 0 and res["true_relres_inf"] <= 1e-6
    assert abs(res["count"] - it_o) <= 2
    assert res["count"] * 5 <= jac
    # the same preconditioned CG: iterates agree up to rounding order
    if res["count"] == it_o:
        xt = oracle.pcg(A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data, b, 1e-13, 100000,
                        oracle.SHK_OR_JACOBI)[0]
        err = np.linalg.norm(xo - xt) / np.linalg.norm(xt)
        rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        log(f"[gpu]   x vs the oracle's iterate: {rel:.2e} (its distance from the solution {err:.2e})")
        assert rel <= 0.5 * err


@pytest.mark.parametrize("case", [(1, 32, 0), (3, 32, 5)])
def test_csr_multigrid_pressure_bar(oracle, scenes, case):
    A, b = _system(oracle, scenes, case)
    x, res = _solve(A, b, {"Residual": BAR_STOP_MG_CSR, "Preconditioner": "mg"})
    xt = oracle.pcg_threaded(A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data, b, 1e-12, 100000,
                             oracle.SHK_OR_JACOBI)[0]
    rel = np.linalg.norm(x - xt) / np.linalg.norm(xt)
    log(f"[gpu] csr mg {case} at {BAR_STOP_MG_CSR:g}: {res['count']} its, pressure rel-L2 vs 1e-12 {rel:.2e}")
    assert res["status"] == 0
    assert rel <= PRESSURE_BAR


def test_csr_multigrid_on_a_graph_laplacian():
    """A matrix with no grid behind it: the Laplacian of a random geometric graph in the
    unit cube (weights 1/distance) plus a small positive diagonal, an SPD M-matrix."""
    from oracle import amg_oracle as G
    rng = np.random.default_rng(3)
    n = 20000
    pts = rng.random((n, 3))
    from scipy.spatial import cKDTree
    pairs = cKDTree(pts).query_pairs(0.045, output_type="ndarray")
    w = 1.0 / np.linalg.norm(pts[pairs[:, 0]] - pts[pairs[:, 1]], axis=1)
    W = sp.coo_matrix((np.r_[w, w], (np.r_[pairs[:, 0], pairs[:, 1]], np.r_[pairs[:, 1], pairs[:, 0]])),
                      shape=(n, n)).tocsr()
    A = (sp.diags(np.asarray(W.sum(axis=1)).ravel() + 1e-3) - W).tocsr()
    A.sort_indices()
    b = rng.standard_normal(n)
    x, res = _solve(A, b, {"Residual": 1e-6, "Preconditioner": "mg"})
    _, it_o = G.pcg(A, b, tol=1e-6)
    _, res_j = _solve(A, b, {"Residual": 1e-6, "Preconditioner": "jacobi"})
    log(f"[gpu] csr mg graph Laplacian: {n} rows, {A.nnz} nnz, {res['count']} its (oracle {it_o}, "
        f"Jacobi-PCG {res_j['count']})")
    assert res["status"] == 0 and res["true_relres_inf"] <= 1e-6
    assert abs(res["count"] - it_o) <= 2
    assert res["count"] * 3 <= res_j["count"]


def test_csr_multigrid_reuses_the_hierarchy_of_the_same_matrix(oracle, scenes):
    """A second solve with a matrix of the same content (other buffers) reuses the
    hierarchy and returns the same result; a changed value builds a new one, whose result
    equals a fresh solver's."""
    import torch

    import shiokaze_amd
    A, b = _system(oracle, scenes, (3, 24, 8))
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dtype=dt).cuda()  # noqa: E731
    args = lambda M: (t(M.indptr.astype(np.int64), torch.int64), t(M.indices.astype(np.int64), torch.int64),  # noqa
                      t(M.data, torch.float64), t(b, torch.float64))
    s = shiokaze_amd.PCGSolver(config={"Residual": 1e-8, "Preconditioner": "mg"})
    x1, r1 = s.solve(*args(A))
    x2, r2 = s.solve(*args(A))
    assert r1["count"] == r2["count"] and torch.equal(x1, x2)
    B = A.copy()
    d5 = B.indptr[5] + np.flatnonzero(B.indices[B.indptr[5]:B.indptr[6]] == 5)[0]
    B.data[d5] *= 1.0 + 2.0 ** -20  # one diagonal entry changed
    x3, r3 = s.solve(*args(B))
    s.close()
    f = shiokaze_amd.PCGSolver(config={"Residual": 1e-8, "Preconditioner": "mg"})
    x4, r4 = f.solve(*args(B))
    f.close()
    log(f"[gpu] csr mg cache: {r1['count']} / {r2['count']} its (same matrix), changed matrix {r3['count']} "
        f"(fresh solver {r4['count']})")
    assert r3["count"] == r4["count"] and torch.equal(x3, x4)
