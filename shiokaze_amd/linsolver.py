"""Host-side mirror of RCMatrix_solver_interface<size_t,double>::solve for the pcg
plugin (include/shiokaze/linsolver/RCMatrix_solver.h:77, src/linsolver/pcg.cpp:45-73)
over libshk_projection's GPU conjugate gradient on CSR.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


class PCGSolver:
    """solve(A, b) -> (x, result); A is CSR (rowptr, cols, vals) as GPU tensors."""

    def __init__(self, device: int = 0, config: dict | None = None):
        self.lib = _lib.load()
        self.device = int(device)
        ctx = ctypes.c_void_p()
        # the context's grid shape is unused by the CSR path
        _lib.check(self.lib.shk_create(1, 1, 1, 1.0, self.device, ctypes.byref(ctx)))
        self.ctx = ctx
        for k, v in (config or {}).items():
            self.configure(k, v)

    def configure(self, key: str, value) -> None:
        _lib.check(self.lib.shk_configure(self.ctx, key.encode(), str(value).encode()), self.ctx)

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.shk_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, rowptr: torch.Tensor, cols: torch.Tensor, vals: torch.Tensor, b: torch.Tensor,
              noalloc: bool = False):
        """noalloc=True runs shk_solve_csr_noalloc: the same solve with its device
        workspace in the library's static per-device scratch array (no hipMalloc)."""
        rows = b.numel()
        for t, dt, name in ((rowptr, torch.int64, "rowptr"), (cols, torch.int64, "cols"),
                            (vals, torch.float64, "vals"), (b, torch.float64, "b")):
            if not t.is_cuda or t.dtype != dt or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous {dt} GPU tensor")
        if rowptr.numel() != rows + 1:
            raise ValueError("rowptr must have rows + 1 entries")
        x = torch.empty_like(b)
        res = _lib.ShkResult()
        torch.cuda.synchronize(self.device)
        fn = self.lib.shk_solve_csr_noalloc if noalloc else self.lib.shk_solve_csr
        rc = fn(self.ctx, rows, *[ctypes.c_void_p(t.data_ptr()) for t in (rowptr, cols, vals, b, x)],
                                    ctypes.byref(res))
        _lib.check(rc, self.ctx, _lib.SOLVER_OUTCOMES)
        out = res.as_dict()
        out["status_name"] = _lib.STATUS_NAMES.get(res.status, str(res.status))
        return x, out
