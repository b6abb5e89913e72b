"""Host-side mirror of Shiokaze's macproject3_interface over libshk_projection.

Method names and argument meaning follow the reference:
``initialize(shape, dx)`` (macproject3_interface.h:104-112), ``configure``
(macpressuresolver3.cpp:301-307, pcg.cpp:39-44, macutility3.cpp:501-504),
``project(dt, velocity, solid, solid_velocity, fluid)`` (macproject3_interface.h:79-87)
and ``get_pressure()`` (:93).  Fields are torch tensors on the GPU (torch is used
for device memory only); the work runs in the library's HIP kernels.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


def _ptr(t: torch.Tensor, numel: int, name: str, dtype=torch.float32) -> ctypes.c_void_p:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must live on the GPU")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() != numel:
        raise ValueError(f"{name} has {t.numel()} elements, expected {numel}")
    return ctypes.c_void_p(t.data_ptr())


class MacPressureSolver3:
    """macpressuresolver3 on the GPU (src/projection/macpressuresolver3.cpp:40-372)."""

    def __init__(self, shape, dx: float, device: int = 0, config: dict | None = None):
        self.lib = _lib.load()
        nx, ny, nz = shape
        self.shape = (int(nx), int(ny), int(nz))
        self.dx = float(dx)
        self.device = int(device)
        ctx = ctypes.c_void_p()
        _lib.check(self.lib.shk_create(self.shape[0], self.shape[1], self.shape[2], self.dx, self.device,
                                       ctypes.byref(ctx)))
        self.ctx = ctx
        for k, v in (config or {}).items():
            self.configure(k, v)

    # --- lifecycle -----------------------------------------------------------------
    def close(self):
        if getattr(self, "ctx", None):
            self.lib.shk_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- configuration ---------------------------------------------------------------
    def configure(self, key: str, value) -> None:
        if isinstance(value, bool):
            value = "1" if value else "0"
        _lib.check(self.lib.shk_configure(self.ctx, key.encode(), str(value).encode()), self.ctx)

    # --- sizes -------------------------------------------------------------------------
    @property
    def cells(self) -> int:
        nx, ny, nz = self.shape
        return nx * ny * nz

    def face_count(self, dim: int) -> int:
        nx, ny, nz = self.shape
        return (nx + (dim == 0)) * (ny + (dim == 1)) * (nz + (dim == 2))

    @property
    def nodes(self) -> int:
        nx, ny, nz = self.shape
        return (nx + 1) * (ny + 1) * (nz + 1)

    # --- project() -----------------------------------------------------------------------
    def set_target_volume(self, current_volume: float, target_volume: float) -> None:
        """macproject3_interface::set_target_volume: volumes for the PI volume correction
        of the following projections (active with Gain != 0 and a nonzero target)."""
        _lib.check(self.lib.shk_set_target_volume(self.ctx, float(current_volume), float(target_volume)), self.ctx)

    def save_state(self) -> bytes:
        """macpressuresolver3::serialize (macpressuresolver3.cpp:332-344): shape, dx,
        volumes, the PI integral and the warm-start pressure, as bytes."""
        n = ctypes.c_size_t()
        _lib.check(self.lib.shk_state_bytes(self.ctx, ctypes.byref(n)), self.ctx)
        buf = ctypes.create_string_buffer(n.value)
        _lib.check(self.lib.shk_save_state(self.ctx, buf, n.value), self.ctx)
        return buf.raw

    def load_state(self, state: bytes) -> None:
        """macpressuresolver3::initialize(const filestream &) (macpressuresolver3.cpp:319-330)."""
        _lib.check(self.lib.shk_load_state(self.ctx, state, len(state)), self.ctx)

    def set_solid_density(self, density) -> None:
        """macproject3_interface::set_solid_density: (du, dv, dw) face tensors of the
        u, v, w shapes, or None to switch it off.  The tensors must stay alive while set."""
        if density is None:
            _lib.check(self.lib.shk_set_solid_density(self.ctx, None, None, None), self.ctx)
            self._density = None
            return
        du, dv, dw = density
        ptrs = [_ptr(du, self.face_count(0), "du"), _ptr(dv, self.face_count(1), "dv"),
                _ptr(dw, self.face_count(2), "dw")]
        _lib.check(self.lib.shk_set_solid_density(self.ctx, *ptrs), self.ctx)
        self._density = (du, dv, dw)  # keep the storage alive while the library points at it

    def project(self, dt: float, velocity, solid, solid_velocity, fluid, surface_tension: float = 0.0) -> dict:
        """Projects `velocity` (u, v, w) in place; returns the solver result as a dict."""
        u, v, w = velocity
        us, vs, ws = solid_velocity
        ptrs = [
            _ptr(u, self.face_count(0), "u"), _ptr(v, self.face_count(1), "v"), _ptr(w, self.face_count(2), "w"),
            _ptr(solid, self.nodes, "solid"),
            _ptr(us, self.face_count(0), "us"), _ptr(vs, self.face_count(1), "vs"), _ptr(ws, self.face_count(2), "ws"),
            _ptr(fluid, self.cells, "fluid"),
        ]
        torch.cuda.synchronize(self.device)
        res = _lib.ShkResult()
        rc = self.lib.shk_project(self.ctx, float(dt), *ptrs, float(surface_tension), ctypes.byref(res))
        _lib.check(rc, self.ctx, _lib.SOLVER_OUTCOMES)
        out = res.as_dict()
        out["status_name"] = _lib.STATUS_NAMES.get(res.status, str(res.status))
        return out

    def get_pressure(self) -> torch.Tensor:
        """Copy of the last pressure field, shape (nz, ny, nx), float32 on the GPU."""
        nx, ny, nz = self.shape
        p = torch.empty((nz, ny, nx), dtype=torch.float32, device=f"cuda:{self.device}")
        torch.cuda.synchronize(self.device)
        _lib.check(self.lib.shk_copy_pressure(self.ctx, ctypes.c_void_p(p.data_ptr())), self.ctx)
        return p

    # --- diagnostics -------------------------------------------------------------------------
    def export_fractions(self) -> dict:
        nx, ny, nz = self.shape
        dev = f"cuda:{self.device}"
        shp = [(nz, ny, nx + 1), (nz, ny + 1, nx), (nz + 1, ny, nx)]
        out = {f"area{a}": torch.empty(shp[a], dtype=torch.float32, device=dev) for a in range(3)}
        out.update({f"rho{a}": torch.empty(shp[a], dtype=torch.float32, device=dev) for a in range(3)})
        torch.cuda.synchronize(self.device)
        _lib.check(self.lib.shk_export_fractions(
            self.ctx, *[ctypes.c_void_p(out[k].data_ptr()) for k in
                        ("area0", "area1", "area2", "rho0", "rho1", "rho2")]), self.ctx)
        return out

    def export_system(self) -> dict:
        nx, ny, nz = self.shape
        dev = f"cuda:{self.device}"
        out = {"label": torch.empty((nz, ny, nx), dtype=torch.uint8, device=dev)}
        for k in ("diag", "cx", "cy", "cz", "rhs", "x"):
            out[k] = torch.empty((nz, ny, nx), dtype=torch.float64, device=dev)
        torch.cuda.synchronize(self.device)
        _lib.check(self.lib.shk_export_system(
            self.ctx, *[ctypes.c_void_p(out[k].data_ptr()) for k in ("label", "diag", "cx", "cy", "cz", "rhs", "x")]),
            self.ctx)
        return out

    def export_scaled(self) -> dict:
        """Dense FP32 (nz, ny, nx) grids of the last system scaled by D^-1/2:
        rhat = D^-1/2 b, mask (1 on unknowns), low-face couplings hx/hy/hz of the
        unit-diagonal operator and xhat = D^1/2 x (the last solution)."""
        nx, ny, nz = self.shape
        out = {k: torch.empty((nz, ny, nx), dtype=torch.float32, device=f"cuda:{self.device}")
               for k in ("rhat", "mask", "hx", "hy", "hz", "xhat")}
        torch.cuda.synchronize(self.device)
        _lib.check(self.lib.shk_export_scaled(
            self.ctx, *[ctypes.c_void_p(out[k].data_ptr()) for k in ("rhat", "mask", "hx", "hy", "hz", "xhat")]),
            self.ctx)
        return out

    def mg_apply(self, r: torch.Tensor) -> torch.Tensor:
        """z = M r with the multigrid V-cycle of the last projection (Preconditioner=mg),
        on dense (nz, ny, nx) FP32 grids of the Jacobi-scaled unknowns."""
        z = torch.empty_like(r)
        torch.cuda.synchronize(self.device)
        _lib.check(self.lib.shk_mg_apply(self.ctx, _ptr(r, self.cells, "r"), _ptr(z, self.cells, "z")), self.ctx)
        return z

    def unet_precondition(self, r: torch.Tensor) -> torch.Tensor:
        """z = M r with the learned preconditioner of the last projection
        (Preconditioner=unet), on dense (nz, ny, nx) FP32 grids of the scaled unknowns."""
        z = torch.empty_like(r)
        torch.cuda.synchronize(self.device)
        _lib.check(self.lib.shk_unet_precondition(self.ctx, _ptr(r, self.cells, "r"), _ptr(z, self.cells, "z")),
                   self.ctx)
        return z

    def divergence(self, dt: float, velocity, solid_velocity, fp64_update: bool) -> float:
        u, v, w = velocity
        us, vs, ws = solid_velocity
        ptrs = [_ptr(u, self.face_count(0), "u"), _ptr(v, self.face_count(1), "v"), _ptr(w, self.face_count(2), "w"),
                _ptr(us, self.face_count(0), "us"), _ptr(vs, self.face_count(1), "vs"),
                _ptr(ws, self.face_count(2), "ws")]
        out = ctypes.c_double(0.0)
        torch.cuda.synchronize(self.device)
        _lib.check(self.lib.shk_divergence(self.ctx, float(dt), *ptrs, int(bool(fp64_update)), ctypes.byref(out)),
                   self.ctx)
        return float(out.value)

    def kernel_stats(self, kernel: str) -> tuple[int, float]:
        n = ctypes.c_int64(0)
        ms = ctypes.c_double(0.0)
        _lib.check(self.lib.shk_kernel_stats(self.ctx, kernel.encode(), ctypes.byref(n), ctypes.byref(ms)), self.ctx)
        return int(n.value), float(ms.value)


class MacPressureSolver2(MacPressureSolver3):
    """macpressuresolver2 on the GPU (src/projection/macpressuresolver2.cpp:40-300):
    ``project(dt, (u, v), solid, (us, vs), fluid, surface_tension)`` on an nx x ny grid.

    Arrays are (ny, nx) cells, (ny, nx+1) u faces, (ny+1, nx) v faces and (ny+1, nx+1)
    solid nodes.  The library holds the grid as nx x 1 x ny, so exported systems come
    back as (ny, nx) arrays whose `cz` is the 2D y coupling (and `cy` is zero)."""

    def __init__(self, shape, dx: float, device: int = 0, config: dict | None = None):
        self.lib = _lib.load()
        nx, ny = shape
        self.shape2 = (int(nx), int(ny))
        self.shape = (int(nx), 1, int(ny))
        self.dx = float(dx)
        self.device = int(device)
        ctx = ctypes.c_void_p()
        _lib.check(self.lib.shk_create2d(self.shape2[0], self.shape2[1], self.dx, self.device, ctypes.byref(ctx)))
        self.ctx = ctx
        for k, v in (config or {}).items():
            self.configure(k, v)

    def face_count2(self, dim: int) -> int:
        nx, ny = self.shape2
        return (nx + (dim == 0)) * (ny + (dim == 1))

    @property
    def nodes(self) -> int:
        nx, ny = self.shape2
        return (nx + 1) * (ny + 1)

    def set_solid_density(self, density) -> None:
        if density is None:
            _lib.check(self.lib.shk_set_solid_density(self.ctx, None, None, None), self.ctx)
            self._density = None
            return
        du, dv = density
        _lib.check(self.lib.shk_set_solid_density(self.ctx, _ptr(du, self.face_count2(0), "du"),
                                                  _ptr(dv, self.face_count2(1), "dv"), None), self.ctx)
        self._density = (du, dv)

    def project(self, dt: float, velocity, solid, solid_velocity, fluid, surface_tension: float = 0.0) -> dict:
        """Projects `velocity` (u, v) in place; returns the solver result as a dict."""
        u, v = velocity
        us, vs = solid_velocity
        ptrs = [_ptr(u, self.face_count2(0), "u"), _ptr(v, self.face_count2(1), "v"), _ptr(solid, self.nodes, "solid"),
                _ptr(us, self.face_count2(0), "us"), _ptr(vs, self.face_count2(1), "vs"),
                _ptr(fluid, self.cells, "fluid")]
        torch.cuda.synchronize(self.device)
        res = _lib.ShkResult()
        rc = self.lib.shk_project2d(self.ctx, float(dt), *ptrs, float(surface_tension), ctypes.byref(res))
        _lib.check(rc, self.ctx, _lib.SOLVER_OUTCOMES)
        out = res.as_dict()
        out["status_name"] = _lib.STATUS_NAMES.get(res.status, str(res.status))
        return out

    def get_pressure(self) -> torch.Tensor:
        """Copy of the last pressure field, shape (ny, nx)."""
        nx, ny = self.shape2
        return super().get_pressure().reshape(ny, nx)

    def export_fractions(self) -> dict:
        nx, ny = self.shape2
        dev = f"cuda:{self.device}"
        shp = [(ny, nx + 1), (ny + 1, nx)]
        out = {f"area{a}": torch.empty(shp[a], dtype=torch.float32, device=dev) for a in range(2)}
        out.update({f"rho{a}": torch.empty(shp[a], dtype=torch.float32, device=dev) for a in range(2)})
        torch.cuda.synchronize(self.device)
        p = lambda k: ctypes.c_void_p(out[k].data_ptr())  # noqa: E731
        _lib.check(self.lib.shk_export_fractions(self.ctx, p("area0"), p("area1"), None, p("rho0"), p("rho1"), None),
                   self.ctx)
        return out

    def export_system(self) -> dict:
        nx, ny = self.shape2
        return {k: v.reshape(ny, nx) for k, v in super().export_system().items()}

    def divergence(self, dt: float, velocity, solid_velocity, fp64_update: bool) -> float:
        u, v = velocity
        us, vs = solid_velocity
        ptrs = [_ptr(u, self.face_count2(0), "u"), _ptr(v, self.face_count2(1), "v"), None,
                _ptr(us, self.face_count2(0), "us"), _ptr(vs, self.face_count2(1), "vs"), None]
        out = ctypes.c_double(0.0)
        torch.cuda.synchronize(self.device)
        _lib.check(self.lib.shk_divergence(self.ctx, float(dt), *ptrs, int(bool(fp64_update)), ctypes.byref(out)),
                   self.ctx)
        return float(out.value)

    def export_scaled(self) -> dict:
        raise NotImplementedError("export_scaled is defined for 3D contexts")

