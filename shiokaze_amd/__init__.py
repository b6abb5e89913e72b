"""shiokaze_amd -- MI355X-native MAC liquid pressure projection for Shiokaze.

The product is libshk_projection.so (C-ABI in include/shk_projection.h, HIP
kernels for gfx950 in csrc/).  This package holds its in-tree build and the
host-side Python mirror of the two reference plugin interfaces it replaces:

* :class:`MacPressureSolver3` -- macproject3_interface / macpressuresolver3
  (include/shiokaze/projection/macproject3_interface.h:39-116,
  src/projection/macpressuresolver3.cpp:40-372);
* :class:`PCGSolver` -- RCMatrix_solver_interface::solve as the pcg plugin
  implements it (include/shiokaze/linsolver/RCMatrix_solver.h:77,
  src/linsolver/pcg.cpp:32-82);
* :class:`MacLiquidVelocity` -- the velocity half of macliquid3's step around the
  projection (src/liquid/macliquid3.cpp:342-391; C-ABI in include/shk_liquid.h).
"""
from ._lib import ShkError, ShkResult, load  # noqa: F401
from .projection import MacPressureSolver2, MacPressureSolver3  # noqa: F401
from .linsolver import PCGSolver  # noqa: F401
from .learned import UNet  # noqa: F401
from .liquid import MacLiquidVelocity, TimeStepper  # noqa: F401
from . import scene  # noqa: F401

__all__ = ["MacLiquidVelocity", "MacPressureSolver2", "MacPressureSolver3", "PCGSolver", "ShkError", "ShkResult",
           "TimeStepper", "UNet", "load", "scene"]
