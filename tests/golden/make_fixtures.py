"""Writes the golden fixtures of tests/golden/: small scenes with their assembled
systems and a direct (sparse LU, scipy) solution.

The reference ships no golden vectors for this path, so the solver is pinned to
an independent algorithm: scipy.sparse.linalg.spsolve on the oracle-assembled
system.  Run from the repo root:  python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import scipy.sparse.linalg as sla

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

CASES = [("pool_sphere_12", 1, 12, 0), ("irregular_16_s1", 3, 16, 1), ("irregular_14x18x10_s2", 3, (14, 18, 10), 2)]


def main():
    from oracle import oracle as O
    from shiokaze_amd import scene

    for name, kind, n, seed in CASES:
        sc = scene.generate(kind, n, seed)
        o = O.project(sc, tol=1e-12, mode=O.SHK_OR_JACOBI)
        A = O.to_scipy(o["rowptr"], o["cols"], o["vals"]).tocsc()
        x = sla.spsolve(A, o["rhs"])
        np.savez_compressed(
            HERE / f"{name}.npz",
            **{k: sc[k] for k in ("fluid", "solid", "u", "v", "w", "us", "vs", "ws")},
            dx=sc["dx"], rowptr=o["rowptr"], cols=o["cols"], vals=o["vals"], rhs=o["rhs"],
            index_map=o["index_map"], x_direct=x)
        print(name, "N =", o["n"], "nnz =", len(o["vals"]))


if __name__ == "__main__":
    main()
