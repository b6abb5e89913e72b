"""Host-side mirror of the velocity half of macliquid3's step over libshk_projection.

``MacLiquidVelocity`` wraps the calls of include/shk_liquid.h on the context of a
``MacPressureSolver3`` and keeps macliquid3::idle's order (src/liquid/macliquid3.cpp:
342-391): the time step from max |u| (``TimeStepper``, timestepper.cpp:48-126), the
velocity extension of ``extend_both`` (:330-340), self-advection (:369-370), gravity
(:207-208) and the projection (:382), with the face activity the reference's sparse
arrays carry held in uint8 masks.  The level set's advection on its narrow band
(``advect_levelset``) is here too; its redistancing (maclevelsetsurfacetracker3.cpp:
54-62) is not, so ``step`` takes the level set as given.
Fields are torch tensors on the GPU (device memory only); the work runs in HIP kernels.
"""
from __future__ import annotations

import ctypes
import math

import torch

from . import _lib
from .projection import MacPressureSolver3, _ptr


class TimeStepper:
    """timestepper::advance without a fixed step and without frame synchronisation
    (timestepper.cpp:48-126, defaults FPS 120, CFL 2, no minimum step)."""

    def __init__(self, fps: float = 120.0, cfl: float = 2.0, min_dt: float = 0.0):
        self.fps, self.cfl, self.min_dt = float(fps), float(cfl), float(min_dt)
        self.time = 0.0
        self.step = 0
        self.frame = 0
        self.accumulated_time = 0.0
        self.current_cfl = 0.0

    def advance(self, max_velocity: float, dx: float) -> float:
        max_unit_u = max_velocity / dx
        max_dt = max(self.min_dt, min(1.0 / self.fps, self.cfl * dx))
        if max_unit_u:
            dt = max(self.min_dt, min(max_dt, self.cfl / max_unit_u))
        else:
            dt = self.min_dt
        if self.accumulated_time + dt >= 1.0 / self.fps:
            while self.accumulated_time + dt >= 1.0 / self.fps:
                self.accumulated_time -= 1.0 / self.fps
                self.frame += 1
        self.time += dt
        self.accumulated_time = math.fmod(self.time, 1.0 / self.fps)
        self.current_cfl = dt * max_velocity / dx
        self.step += 1
        return dt


class MacLiquidVelocity:
    """The velocity half of macliquid3's step on the GPU, on `solver`'s context."""

    def __init__(self, solver: MacPressureSolver3, gravity=(0.0, -9.8, 0.0), extend_width: int = 2):
        self.solver = solver
        self.lib = solver.lib
        self.gravity = tuple(float(g) for g in gravity)   # macliquid3.h:144
        self.extend_width = int(extend_width)              # macliquid3.h:166 (extend_both(int w=2))
        self.timestepper = TimeStepper()

    # --- arguments -----------------------------------------------------------------------
    def _faces(self, velocity, masks):
        s = self.solver
        ptrs = [_ptr(t, s.face_count(d), "uvw"[d]) for d, t in enumerate(velocity)]
        mptrs = [_ptr(m, s.face_count(d), "m" + "uvw"[d], torch.uint8) for d, m in enumerate(masks)]
        return ptrs, mptrs

    def _check(self, code):
        _lib.check(code, self.solver.ctx)

    # --- the calls of include/shk_liquid.h ---------------------------------------------------
    def extrapolate(self, velocity, masks, width: int) -> int:
        """macarray_extrapolator3::extrapolate(velocity, width); returns the faces activated."""
        p, m = self._faces(velocity, masks)
        n = ctypes.c_int64(0)
        self._check(self.lib.shk_extrapolate_velocity(self.solver.ctx, *p, *m, int(width), ctypes.byref(n)))
        return int(n.value)

    def constrain(self, solid, velocity, masks) -> None:
        """macutility3::constrain_velocity(solid, velocity); solid may be None."""
        p, m = self._faces(velocity, masks)
        sp = _ptr(solid, self.solver.nodes, "solid") if solid is not None else None
        self._check(self.lib.shk_constrain_velocity(self.solver.ctx, sp, *p, *m))

    def compute_max_u(self, velocity, masks) -> float:
        """macutility3::compute_max_u(velocity)."""
        p, m = self._faces(velocity, masks)
        out = ctypes.c_double(0.0)
        self._check(self.lib.shk_compute_max_u(self.solver.ctx, *p, *m, ctypes.byref(out)))
        return float(out.value)

    def advect(self, dt: float, velocity, masks, fluid) -> None:
        """macadvection3::advect_vector(velocity, velocity, fluid, dt) (MacCormack)."""
        p, m = self._faces(velocity, masks)
        fp = _ptr(fluid, self.solver.cells, "fluid")
        self._check(self.lib.shk_advect_velocity(self.solver.ctx, float(dt), *p, *m, fp))

    def advect_levelset(self, dt: float, phi, band, velocity, masks, background: float) -> None:
        """macadvection3::advect_scalar(fluid, velocity, fluid_save, dt) on the narrow
        band `band` (uint8 cell mask) of the level set `phi`, in place; `background` is
        the level set's background value (its half bandwidth, 3 dx by default)."""
        p, m = self._faces(velocity, masks)
        self._check(self.lib.shk_advect_levelset(self.solver.ctx, float(dt), _ptr(phi, self.solver.cells, "phi"),
                                                 _ptr(band, self.solver.cells, "band", torch.uint8), *p, *m,
                                                 float(background)))

    def redistance(self, phi, band, width: int = 3) -> int:
        """redistancer3_interface::redistance(phi, width) (the pderedistancer3 plugin) on
        the level set `phi` and its narrow band `band` (uint8 cell mask), in place;
        returns the final band's cell count.  Cells off the band are set to -/+ width*dx
        by the sign of their value."""
        active = ctypes.c_int64()
        self._check(self.lib.shk_redistance(self.solver.ctx, _ptr(phi, self.solver.cells, "phi"),
                                            _ptr(band, self.solver.cells, "band", torch.uint8), int(width),
                                            ctypes.byref(active)))
        return int(active.value)

    def extrapolate_levelset(self, solid, phi, band, background: float, solid_band=None) -> int:
        """gridutility3::extrapolate_levelset(solid, fluid) on the level set `phi` and its
        band (uint8 cell mask), in place, with the nodal solid and its node activity
        (None: every node); `background` is the level set's half bandwidth.  Returns the
        final band's cell count."""
        nn = (self.solver.shape[0] + 1) * (self.solver.shape[1] + 1) * (self.solver.shape[2] + 1)
        active = ctypes.c_int64()
        sb = None if solid_band is None else _ptr(solid_band, nn, "solid_band", torch.uint8)
        self._check(self.lib.shk_extrapolate_levelset(self.solver.ctx, _ptr(solid, nn, "solid"), sb,
                                                      _ptr(phi, self.solver.cells, "phi"),
                                                      _ptr(band, self.solver.cells, "band", torch.uint8),
                                                      float(background), ctypes.byref(active)))
        return int(active.value)

    def add_gravity(self, dt: float, velocity, masks) -> None:
        """`m_velocity += dt*gravity` (activates every face)."""
        p, m = self._faces(velocity, masks)
        self._check(self.lib.shk_add_gravity(self.solver.ctx, float(dt), *self.gravity, *p, *m))

    def project_activity(self, masks) -> None:
        """The faces the last projection switched off become inactive."""
        mptrs = [_ptr(mm, self.solver.face_count(d), "m" + "uvw"[d], torch.uint8) for d, mm in enumerate(masks)]
        self._check(self.lib.shk_project_activity(self.solver.ctx, *mptrs))

    # --- one step -------------------------------------------------------------------------------
    def step(self, velocity, masks, solid, solid_velocity, fluid, dt: float | None = None) -> dict:
        """macliquid3::idle's velocity half and projection for the level set `fluid`:
        dt from max |u| (unless given), extend (width = extend_width + CFL, truncated),
        constrain, self-advection, gravity, projection and the activity it leaves.
        Returns the projection's result with dt, CFL and the extension count added."""
        max_u = self.compute_max_u(velocity, masks)
        if dt is None:
            dt = self.timestepper.advance(max_u, self.solver.dx)
            cfl = self.timestepper.current_cfl
        else:
            cfl = dt * max_u / self.solver.dx
        width = int(self.extend_width + cfl)   # `unsigned width = w + CFL` (macliquid3.cpp:335)
        grown = self.extrapolate(velocity, masks, width)
        self.constrain(solid, velocity, masks)
        self.advect(dt, velocity, masks, fluid)
        self.add_gravity(dt, velocity, masks)
        res = self.solver.project(dt, velocity, solid, solid_velocity, fluid)
        self.project_activity(masks)
        res.update({"dt": dt, "cfl": cfl, "max_u": max_u, "extend_width": width, "extended": grown})
        return res
