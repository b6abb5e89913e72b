"""shk_solve_csr_noalloc through the C-ABI: the allocation-free variant of
shk_solve_csr (workspace in the library's static per-device scratch array) returns
the same iterate, iteration count and residuals as shk_solve_csr, bit for bit (the
same kernels in the same order), on two calls in a row with the second system
smaller than the first, without changing the device's free memory; systems beyond
the array are refused."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from gpu_helpers import log

pytestmark = pytest.mark.gpu


def _csr(oracle, scenes, kind, n, seed):
    o = oracle.system(scenes(kind, n, seed))
    return o["rowptr"], o["cols"], o["vals"], o["rhs"]


def _dev(csr):
    import torch
    rowptr, cols, vals, rhs = csr
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dtype=dt).cuda()  # noqa: E731
    return t(rowptr, torch.int64), t(cols, torch.int64), t(vals, torch.float64), t(rhs, torch.float64)


def _solve(lib, ctx, dev, x, noalloc):
    """One CSR solve through the C-ABI into the caller's preallocated x."""
    import torch

    from shiokaze_amd import _lib
    res = _lib.ShkResult()
    fn = lib.shk_solve_csr_noalloc if noalloc else lib.shk_solve_csr
    torch.cuda.synchronize()
    rc = fn(ctx, dev[3].numel(), *[ctypes.c_void_p(t.data_ptr()) for t in (*dev, x)], ctypes.byref(res))
    torch.cuda.synchronize()
    assert rc == 0, lib.shk_last_error(ctx)
    return res.as_dict()


@pytest.mark.parametrize("precond", ["jacobi", "none"])
def test_noalloc_matches_solve_csr(scenes, oracle, precond):
    import torch

    import shiokaze_amd
    big = _dev(_csr(oracle, scenes, 3, 32, 5))
    small = _dev(_csr(oracle, scenes, 1, 16, 0))
    tiny = _dev(_csr(oracle, scenes, 1, 8, 0))
    cfg = {"Residual": 1e-6, "Preconditioner": precond}
    ref = shiokaze_amd.PCGSolver(config=cfg)
    na = shiokaze_amd.PCGSolver(config=cfg)
    lib = ref.lib
    xe = [torch.full_like(big[3], float("nan")), torch.full_like(small[3], float("nan"))]
    xg = [torch.full_like(big[3], float("nan")), torch.full_like(small[3], float("nan"))]
    xt = torch.empty_like(tiny[3])
    expect = [_solve(lib, ref.ctx, big, xe[0], False), _solve(lib, ref.ctx, small, xe[1], False)]
    _solve(lib, na.ctx, tiny, xt, True)  # the first launches on this context's stream
    free0 = torch.cuda.mem_get_info()[0]
    got = [_solve(lib, na.ctx, big, xg[0], True), _solve(lib, na.ctx, small, xg[1], True)]
    free1 = torch.cuda.mem_get_info()[0]
    for re, rg, a, b, name in zip(expect, got, xe, xg, ("first (larger)", "second (smaller)")):
        log(f"[gpu] noalloc {precond} {name}: {rg['count']} its (shk_solve_csr {re['count']}), "
            f"true relres {rg['true_relres_inf']:.3e}")
        assert rg["status"] == 0 and re["status"] == 0
        assert rg["count"] == re["count"]
        assert torch.equal(b, a)
        for k in ("reresid", "true_relres_inf", "true_relres_l2", "b_inf"):
            assert rg[k] == re[k], k
    # no device memory was taken by the two allocation-free calls
    assert free1 == free0, (free0, free1)
    ref.close()
    na.close()


def test_noalloc_refuses_oversize_systems():
    import torch

    import shiokaze_amd
    from shiokaze_amd import _lib
    lib = _lib.load()
    cap = lib.shk_noalloc_scratch_bytes()
    assert cap == 128 << 20
    rows = cap // (6 * 8) + 1  # six FP64 vectors alone exceed the array
    rowptr = torch.zeros(rows + 1, dtype=torch.int64, device="cuda")
    cols = torch.zeros(1, dtype=torch.int64, device="cuda")
    vals = torch.zeros(1, dtype=torch.float64, device="cuda")
    b = torch.zeros(rows, dtype=torch.float64, device="cuda")
    x = torch.zeros(rows, dtype=torch.float64, device="cuda")
    for precond in ("none", "jacobi"):
        s = shiokaze_amd.PCGSolver(config={"Preconditioner": precond})
        res = _lib.ShkResult()
        torch.cuda.synchronize()
        rc = lib.shk_solve_csr_noalloc(s.ctx, rows, *[ctypes.c_void_p(t.data_ptr()) for t in (rowptr, cols, vals,
                                                                                              b, x)],
                                       ctypes.byref(res))
        assert rc == _lib.SHK_BAD_ARG
        assert b"scratch" in lib.shk_last_error(s.ctx)
        s.close()
    for precond in ("mg", "unet"):
        s = shiokaze_amd.PCGSolver(config={"Preconditioner": precond})
        rc = lib.shk_solve_csr_noalloc(s.ctx, 4, *[ctypes.c_void_p(t.data_ptr()) for t in (rowptr, cols, vals, b,
                                                                                           x)],
                                       ctypes.byref(_lib.ShkResult()))
        assert rc == _lib.SHK_BAD_ARG
        s.close()
