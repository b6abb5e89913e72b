"""Host-side handle of the learned preconditioner (shk_unet_* in libshk_projection).

The reference's pcg plugin applies no preconditioner (pcg_solver.h:379-384); this
network is an addition of this library, used by MacPressureSolver3 when it is
configured with ``Preconditioner=unet`` and ``UnetWeights=<file>``.  Grids are
dense FP32 torch tensors on the GPU of shape (nz, ny, nx) with dims that are
multiples of 16, holding the Jacobi-scaled system that
:meth:`MacPressureSolver3.export_scaled` writes.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .projection import _ptr


class UNet:
    """A 4-level 3D U-Net (BF16 MFMA convolutions) and its Adam trainer on one GPU."""

    def __init__(self, seed: int = 0, device: int = 0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self.lib.shk_unet_create(int(seed), int(device), ctypes.byref(h)))
        self.h = h
        self.device = int(device)

    def close(self):
        if getattr(self, "h", None):
            self.lib.shk_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int) -> None:
        if rc != _lib.SHK_OK:
            msg = self.lib.shk_unet_last_error(self.h)
            raise _lib.ShkError(rc, msg.decode() if msg else "")

    @property
    def num_params(self) -> int:
        n = ctypes.c_int64(0)
        self._check(self.lib.shk_unet_num_params(self.h, ctypes.byref(n)))
        return int(n.value)

    def get_params(self, grads: bool = False) -> np.ndarray:
        out = np.empty(self.num_params, dtype=np.float32)
        self._check(self.lib.shk_unet_get_params(self.h, out.ctypes.data_as(ctypes.c_void_p), int(bool(grads))))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {p.size}")
        self._check(self.lib.shk_unet_set_params(self.h, p.ctypes.data_as(ctypes.c_void_p)))

    def save(self, path) -> None:
        self._check(self.lib.shk_unet_save(self.h, str(path).encode()))

    def load(self, path) -> None:
        self._check(self.lib.shk_unet_load(self.h, str(path).encode()))

    def set_smoothing(self, steps: int, chebyshev: float = 0.0) -> None:
        """Smoothing of the training loss's M and of apply(): steps steps with omega, or
        with Chebyshev weights for [2 / chebyshev, 2] when chebyshev > 1."""
        self._check(self.lib.shk_unet_set_smoothing(self.h, int(steps), float(chebyshev)))

    @staticmethod
    def _grid(t: torch.Tensor):
        if t.dim() != 3:
            raise ValueError("grids are (nz, ny, nx) tensors")
        nz, ny, nx = t.shape
        return int(nx), int(ny), int(nz), int(t.numel())

    def train_step(self, r, mask, cx, cy, cz, xstar, omega: float = 1.0, lr: float = 0.0) -> float:
        """A-norm loss of M(r) against xstar; gradients accumulate, lr > 0 takes an Adam step."""
        nx, ny, nz, n = self._grid(r)
        ptrs = [_ptr(t, n, nm) for t, nm in ((r, "r"), (mask, "mask"), (cx, "cx"), (cy, "cy"), (cz, "cz"),
                                            (xstar, "xstar"))]
        loss = ctypes.c_double(0.0)
        torch.cuda.synchronize(self.device)
        self._check(self.lib.shk_unet_train_step(self.h, nx, ny, nz, *ptrs, float(omega), float(lr),
                                                 ctypes.byref(loss)))
        return float(loss.value)

    def apply(self, r, mask, cx, cy, cz, omega: float = 1.0) -> torch.Tensor:
        """z = M(r)."""
        nx, ny, nz, n = self._grid(r)
        z = torch.empty_like(r)
        ptrs = [_ptr(t, n, nm) for t, nm in ((r, "r"), (mask, "mask"), (cx, "cx"), (cy, "cy"), (cz, "cz"))]
        torch.cuda.synchronize(self.device)
        self._check(self.lib.shk_unet_apply(self.h, nx, ny, nz, *ptrs, float(omega), _ptr(z, n, "z")))
        return z


# (cin, cout, level) of the network's 3x3x3 convolutions as inference computes them; level
# l runs on the fine grid / 2^(l+1).  The head has 16 outputs but only channel 0 is used,
# and inference computes that one alone (on the vector ALUs).
LAYERS = [(8, 16, 0), (16, 32, 1), (32, 64, 2), (64, 128, 3), (192, 64, 2), (96, 32, 1), (48, 16, 0), (16, 1, 0)]


def flops_per_apply(shape) -> float:
    """FLOPs of one network forward (inference) for a fine grid (nx, ny, nz), padded to
    multiples of 16 as the solver pads it: 2 * 27 * cin * cout per output voxel."""
    pad = [(int(d) + 15) // 16 * 16 for d in shape]
    total = 0.0
    for cin, cout, lvl in LAYERS:
        v = 1
        for d in pad:
            v *= d >> (lvl + 1)
        total += 2.0 * 27 * cin * cout * v
    return total


def residual(cx, cy, cz, x, b) -> torch.Tensor:
    """r = b - A^ x for the unit-diagonal operator with low-face couplings cx, cy, cz."""
    lib = _lib.load()
    nz, ny, nx = b.shape
    n = b.numel()
    r = torch.empty_like(b)
    torch.cuda.synchronize(b.device)
    _lib.check(lib.shk_unet_residual(nx, ny, nz, *[_ptr(t, n, nm) for t, nm in
                                                   ((cx, "cx"), (cy, "cy"), (cz, "cz"), (x, "x"), (b, "b"))],
                                     _ptr(r, n, "r")))
    return r


def conv3d(x: torch.Tensor, w: torch.Tensor, bias: torch.Tensor | None, relu: bool) -> torch.Tensor:
    """One 3x3x3 convolution layer of the network's kernels: x (nz, ny, nx, cin) bf16,
    w (27, cin, cout) fp32 with tap = (dz+1)*9 + (dy+1)*3 + (dx+1); returns bf16."""
    lib = _lib.load()
    nz, ny, nx, cin = x.shape
    cout = w.shape[2]
    y = torch.empty((nz, ny, nx, cout), dtype=torch.bfloat16, device=x.device)
    nv = nx * ny * nz
    bp = _ptr(bias, cout, "bias") if bias is not None else None
    torch.cuda.synchronize(x.device)
    _lib.check(lib.shk_nn_conv3d(nx, ny, nz, _ptr(x, nv * cin, "x", torch.bfloat16), cin,
                                 _ptr(w, 27 * cin * cout, "w"), bp, cout, int(bool(relu)),
                                 _ptr(y, nv * cout, "y", torch.bfloat16)))
    return y


def conv3d_cat(coarse: torch.Tensor | None, x: torch.Tensor, w: torch.Tensor, bias: torch.Tensor | None,
               relu: bool, kernel: int = 0) -> torch.Tensor:
    """conv3d of the concatenation [upsample(coarse), x] (coarse (nz/2, ny/2, nx/2, c1)
    bf16 or None) by kernel 0 (the network's inference choice for this grid), 1 (the
    tiled kernel), 2 (the row-reuse kernel, 16 or 32 outputs) or 3 (the direct-load
    kernel of the small grids)."""
    lib = _lib.load()
    nz, ny, nx, cx = x.shape
    c1 = 0 if coarse is None else coarse.shape[3]
    cin = c1 + cx
    cout = w.shape[2]
    if w.shape[1] != cin:
        raise ValueError(f"weights take {w.shape[1]} input channels, the input has {cin}")
    y = torch.empty((nz, ny, nx, cout), dtype=torch.bfloat16, device=x.device)
    nv = nx * ny * nz
    cp = None if coarse is None else _ptr(coarse, (nv // 8) * c1, "coarse", torch.bfloat16)
    bp = _ptr(bias, cout, "bias") if bias is not None else None
    torch.cuda.synchronize(x.device)
    _lib.check(lib.shk_nn_conv3d_cat(nx, ny, nz, cp, c1, _ptr(x, nv * cx, "x", torch.bfloat16), cin,
                                     _ptr(w, 27 * cin * cout, "w"), bp, cout, int(bool(relu)), int(kernel),
                                     _ptr(y, nv * cout, "y", torch.bfloat16)))
    return y


def conv3d_backward(x: torch.Tensor, w: torch.Tensor, dz: torch.Tensor, need_dx: bool = True):
    """(dx bf16 or None, dw fp32, db fp32) of conv3d for the output gradient dz."""
    lib = _lib.load()
    nz, ny, nx, cin = x.shape
    cout = w.shape[2]
    nv = nx * ny * nz
    dx = torch.empty_like(x) if need_dx else None
    dw = torch.zeros((27, cin, cout), dtype=torch.float32, device=x.device)
    db = torch.zeros(cout, dtype=torch.float32, device=x.device)
    torch.cuda.synchronize(x.device)
    _lib.check(lib.shk_nn_conv3d_backward(nx, ny, nz, _ptr(x, nv * cin, "x", torch.bfloat16), cin,
                                          _ptr(w, 27 * cin * cout, "w"), cout,
                                          _ptr(dz, nv * cout, "dz", torch.bfloat16),
                                          _ptr(dx, nv * cin, "dx", torch.bfloat16) if need_dx else None,
                                          _ptr(dw, dw.numel(), "dw"), _ptr(db, cout, "db")))
    return dx, dw, db
