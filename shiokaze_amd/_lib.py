"""ctypes binding of libshk_projection.so (declarations: include/shk_projection.h).

The library is the product path: there is no Python or CPU fallback.  Loading
fails loudly when the shared object has not been built.
"""
from __future__ import annotations

import ctypes
import os
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libshk_projection.so"
HEADER = PKG.parent / "include" / "shk_projection.h"

SHK_OK = 0
SHK_NOT_CONVERGED = 1
SHK_NAN = 2
SHK_BAD_ARG = 3
SHK_HIP_ERROR = 4
SHK_NO_DEVICE = 5
SHK_BREAKDOWN = 6
SHK_OUT_OF_MEMORY = 7
STATUS_NAMES = {
    0: "OK", 1: "NOT_CONVERGED", 2: "NAN", 3: "BAD_ARG", 4: "HIP_ERROR",
    5: "NO_DEVICE", 6: "BREAKDOWN", 7: "OUT_OF_MEMORY",
}


class ShkResult(ctypes.Structure):
    """shk_result (include/shk_projection.h)."""

    _fields_ = [
        ("count", ctypes.c_int64),
        ("reresid", ctypes.c_double),
        ("true_relres_inf", ctypes.c_double),
        ("true_relres_l2", ctypes.c_double),
        ("b_inf", ctypes.c_double),
        ("unknowns", ctypes.c_int64),
        ("active_bricks", ctypes.c_int64),
        ("outer", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("fallback", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("ms_setup", ctypes.c_double),
        ("ms_solve", ctypes.c_double),
        ("ms_update", ctypes.c_double),
        ("volume_correct_rhs", ctypes.c_double),
        ("report_nnz", ctypes.c_int64),
        ("report_min_row", ctypes.c_int32),
        ("report_max_row", ctypes.c_int32),
        ("report_min_diag", ctypes.c_double),
        ("report_max_diag", ctypes.c_double),
        ("report_max_offdiag_ratio", ctypes.c_double),
        ("report_symmetry_error", ctypes.c_double),
        ("report_has_nan", ctypes.c_int32),
        ("report_valid", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ShkError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


_vp = ctypes.c_void_p
_SIGNATURES = {
    "shk_abi_version": (ctypes.c_int, []),
    "shk_build_info": (ctypes.c_char_p, []),
    "shk_last_error": (ctypes.c_char_p, [_vp]),
    "shk_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                  ctypes.POINTER(_vp)]),
    "shk_destroy": (None, [_vp]),
    "shk_configure": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "shk_project": (ctypes.c_int, [_vp, ctypes.c_double] + [_vp] * 8 + [ctypes.c_double, ctypes.POINTER(ShkResult)]),
    "shk_create2d": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.POINTER(_vp)]),
    "shk_project2d": (ctypes.c_int, [_vp, ctypes.c_double] + [_vp] * 6 + [ctypes.c_double, ctypes.POINTER(ShkResult)]),
    "shk_set_target_volume": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double]),
    "shk_set_solid_density": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "shk_state_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    "shk_save_state": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "shk_load_state": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "shk_pressure": (_vp, [_vp]),
    "shk_copy_pressure": (ctypes.c_int, [_vp, _vp]),
    "shk_export_fractions": (ctypes.c_int, [_vp] * 7),
    "shk_export_system": (ctypes.c_int, [_vp] * 8),
    "shk_divergence": (ctypes.c_int, [_vp, ctypes.c_double] + [_vp] * 6 + [ctypes.c_int,
                                                                          ctypes.POINTER(ctypes.c_double)]),
    "shk_solve_csr": (ctypes.c_int, [_vp, ctypes.c_int64] + [_vp] * 5 + [ctypes.POINTER(ShkResult)]),
    "shk_solve_csr_noalloc": (ctypes.c_int, [_vp, ctypes.c_int64] + [_vp] * 5 + [ctypes.POINTER(ShkResult)]),
    "shk_noalloc_scratch_bytes": (ctypes.c_size_t, []),
    "shk_kernel_stats": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_double)]),
    "shk_scene_generate": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_uint64] + [_vp] * 8 + [ctypes.POINTER(ctypes.c_int64)]),
    "shk_unet_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(_vp)]),
    "shk_unet_destroy": (None, [_vp]),
    "shk_unet_last_error": (ctypes.c_char_p, [_vp]),
    "shk_unet_num_params": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "shk_unet_get_params": (ctypes.c_int, [_vp, _vp, ctypes.c_int]),
    "shk_unet_set_params": (ctypes.c_int, [_vp, _vp]),
    "shk_unet_set_smoothing": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double]),
    "shk_unet_save": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "shk_unet_load": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "shk_unet_train_step": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [_vp] * 6
                            + [ctypes.c_double, ctypes.c_float, ctypes.POINTER(ctypes.c_double)]),
    "shk_unet_apply": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [_vp] * 5
                       + [ctypes.c_double, _vp]),
    "shk_export_scaled": (ctypes.c_int, [_vp] * 7),
    "shk_unet_residual": (ctypes.c_int, [ctypes.c_int] * 3 + [_vp] * 6),
    "shk_mg_apply": (ctypes.c_int, [_vp, _vp, _vp]),
    "shk_unet_precondition": (ctypes.c_int, [_vp, _vp, _vp]),
    "shk_nn_conv3d": (ctypes.c_int, [ctypes.c_int] * 3 + [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                                          _vp]),
    "shk_nn_conv3d_cat": (ctypes.c_int, [ctypes.c_int] * 3 + [_vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp,
                                                              ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]),
    "shk_nn_conv3d_backward": (ctypes.c_int, [ctypes.c_int] * 3 + [_vp, ctypes.c_int, _vp, ctypes.c_int] + [_vp] * 4),
    # include/shk_liquid.h
    "shk_extrapolate_velocity": (ctypes.c_int, [_vp] * 7 + [ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]),
    "shk_constrain_velocity": (ctypes.c_int, [_vp] * 8),
    "shk_extrapolate_levelset": (ctypes.c_int, [_vp] * 5 + [ctypes.c_float, ctypes.POINTER(ctypes.c_int64)]),
    "shk_compute_max_u": (ctypes.c_int, [_vp] * 7 + [ctypes.POINTER(ctypes.c_double)]),
    "shk_advect_velocity": (ctypes.c_int, [_vp, ctypes.c_double] + [_vp] * 7),
    "shk_add_gravity": (ctypes.c_int, [_vp] + [ctypes.c_double] * 4 + [_vp] * 6),
    "shk_advect_levelset": (ctypes.c_int, [_vp, ctypes.c_double] + [_vp] * 8 + [ctypes.c_float]),
    "shk_project_activity": (ctypes.c_int, [_vp] * 4),
    "shk_redistance": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]),
}

_lib = None


def header_symbols() -> list[str]:
    """Function names declared in the public headers include/*.h."""
    text = "".join(h.read_text() for h in sorted(HEADER.parent.glob("*.h")))
    return sorted(set(re.findall(r"\b(shk_[a-z_0-9]+)\s*\(", text)))


def load() -> ctypes.CDLL:
    """Loads the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("SHK_LIBRARY", LIB_PATH))
    if not path.exists():
        raise ImportError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(or python -m shiokaze_amd.build); there is no fallback path")
    lib = ctypes.CDLL(str(path))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


SOLVER_OUTCOMES = (SHK_OK, SHK_NOT_CONVERGED, SHK_NAN, SHK_BREAKDOWN)


def check(code: int, ctx=None, outcomes=(SHK_OK,)) -> int:
    """Raises ShkError unless `code` is one of `outcomes`.  Solver calls pass
    SOLVER_OUTCOMES: like the reference (pcg_solver.h:263-294) a capped, NaN or
    broken-down solve returns its status instead of failing."""
    if code in outcomes:
        return code
    lib = load()
    msg = lib.shk_last_error(ctx).decode() if lib.shk_last_error(ctx) else ""
    raise ShkError(code, msg)
