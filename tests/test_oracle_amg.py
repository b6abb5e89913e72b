"""Pins the numpy restatement of the CSR aggregation multigrid (oracle/amg_oracle.py),
the checker of shk_solve_csr with Preconditioner=mg, by routes independent of its
code: the coarse matrix of a 1D Laplacian in closed form, explicit P'AP products,
aggregate properties checked by brute force, symmetry and positivity of the V-cycle
operator (built column by column), and PCG convergence against a direct solve."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from oracle import amg_oracle as G


def laplacian_1d(n, dirichlet=1.0):
    """tridiag(-1, 2, -1) with an extra `dirichlet` on the first diagonal entry."""
    A = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tolil()
    A[0, 0] += dirichlet - 1.0
    return A.tocsr()


def projection_matrix(oracle, scenes, kind, n, seed):
    sy = oracle.system(scenes(kind, n, seed))
    N = len(sy["rhs"])
    return sp.csr_matrix((sy["vals"], sy["cols"], sy["rowptr"]), shape=(N, N)), sy["rhs"]


def test_pairwise_on_a_chain_pairs_neighbours_in_order():
    """Equal couplings: row 0 pairs with 1 (ties go to the lower index), 2 with 3, ..."""
    A = laplacian_1d(9, dirichlet=2.0)
    agg, nc = G.pairwise(A)
    assert nc == 5
    np.testing.assert_array_equal(agg, [0, 0, 1, 1, 2, 2, 3, 3, 4])


def test_galerkin_of_the_chain_in_closed_form():
    """Pairs of a Dirichlet 1D Laplacian: P'AP = tridiag(-1, 2, -1) with the boundary
    row keeping its Dirichlet weight (row sums of the pairs)."""
    A = laplacian_1d(8, dirichlet=2.0)
    agg, nc = G.pairwise(A)
    Ac = G.galerkin(A, agg, nc).toarray()
    # coarse row I = rows {2I, 2I+1}: a_2I,2I + a_2I+1,2I+1 + 2 a_2I,2I+1 on the diagonal
    # (3 + 2 - 2 for the first pair, whose row 0 carries the extra Dirichlet weight),
    # the coupling a_2I+1,2I+2 = -1 between neighbouring pairs
    want = np.diag(2.0 * np.ones(4)) - np.diag(np.ones(3), 1) - np.diag(np.ones(3), -1)
    want[0, 0] = 3.0
    np.testing.assert_allclose(Ac, want)


def test_galerkin_equals_explicit_product(oracle, scenes):
    A, _ = projection_matrix(oracle, scenes, 3, 12, 4)
    agg, nc = G.pairwise(A)
    P = np.zeros((A.shape[0], nc))
    P[np.arange(A.shape[0]), agg] = 1.0
    np.testing.assert_allclose(G.galerkin(A, agg, nc).toarray(), P.T @ A.toarray() @ P, rtol=0, atol=1e-12)


def test_aggregates_are_small_and_connected(oracle, scenes):
    """Two pairwise passes: every aggregate holds 1..4 rows and is connected through
    nonzero couplings of A (brute force over each aggregate's members)."""
    A, _ = projection_matrix(oracle, scenes, 3, 16, 1)
    levels = G.hierarchy(A)
    for Al, agg in levels[:-1]:
        counts = np.bincount(agg)
        assert counts.min() >= 1 and counts.max() <= 4
        Ad = Al.toarray()
        for I in range(len(counts)):
            mem = np.flatnonzero(agg == I)
            seen = {mem[0]}
            frontier = [mem[0]]
            while frontier:
                i = frontier.pop()
                for j in mem:
                    if j not in seen and Ad[i, j] != 0:
                        seen.add(j)
                        frontier.append(j)
            assert len(seen) == len(mem)
    sizes = [Al.shape[0] for Al, _ in levels]
    assert all(b < a for a, b in zip(sizes, sizes[1:]))
    assert sizes[-1] <= G.COARSE_MAX


def test_vcycle_is_symmetric_positive_definite(oracle, scenes):
    A, _ = projection_matrix(oracle, scenes, 1, 10, 0)
    levels = G.hierarchy(A)
    n = A.shape[0]
    M = np.column_stack([G.vcycle(levels, e) for e in np.eye(n)])
    np.testing.assert_allclose(M, M.T, rtol=0, atol=1e-12 * np.abs(M).max())
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0


@pytest.mark.parametrize("case", [(1, 24, 0), (3, 20, 6)])
def test_pcg_converges_to_the_direct_solution(oracle, scenes, case):
    A, b = projection_matrix(oracle, scenes, *case)
    x, it = G.pcg(A, b, tol=1e-10)
    xd = spl.spsolve(A.tocsc(), b)
    assert np.abs(A @ x - b).max() <= 1e-10 * np.abs(b).max()
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    jac = oracle.pcg(A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data, b, 1e-10, 100000,
                     oracle.SHK_OR_JACOBI)[1]
    assert it * 4 < jac, (it, jac)


def test_pcg_converges_on_a_graph_laplacian():
    """A matrix without a grid (the Laplacian of a random geometric graph plus a small
    diagonal), whose aggregation stalls after a few levels: the hierarchy stops where a
    level keeps more than STALL of its rows, the V-cycle stays SPD and PCG converges."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(5)
    n = 4000
    pts = rng.random((n, 3))
    pairs = cKDTree(pts).query_pairs(0.08, output_type="ndarray")
    w = 1.0 / np.linalg.norm(pts[pairs[:, 0]] - pts[pairs[:, 1]], axis=1)
    W = sp.coo_matrix((np.r_[w, w], (np.r_[pairs[:, 0], pairs[:, 1]], np.r_[pairs[:, 1], pairs[:, 0]])),
                      shape=(n, n)).tocsr()
    A = (sp.diags(np.asarray(W.sum(axis=1)).ravel() + 1e-3) - W).tocsr()
    A.sort_indices()
    levels = G.hierarchy(A)
    sizes = [Al.shape[0] for Al, _ in levels]
    assert all(b <= G.STALL * a for a, b in zip(sizes, sizes[1:]))
    b = rng.standard_normal(n)
    x, it = G.pcg(A, b, tol=1e-8)
    assert np.abs(A @ x - b).max() <= 1e-8 * np.abs(b).max()
    assert it < 60, (it, sizes)
