"""classic_amd -- MI355X (gfx950) wave-propagation step for Clawpack Classic.

The hot path is the HIP code under ``csrc/`` behind the C-ABI of
``include/clawgpu.h`` (``libclawgpu.so``, built in-tree by
``python -m classic_amd.build``).  Python front ends:

* ``gpu`` (2D grids: step2, dimsp2, claw2, user routines), ``gpu3`` (3D),
  ``gpu1`` (1D), ``batch`` (many 2D problems in one persistent kernel),
  ``learned`` (the learned flux limiter);
* ``rundata`` (``ClawRunData`` / claw.data), ``output`` (fort.q/t/a/info
  files), ``driver`` (the claw1ez / claw2ez / claw3ez drivers);
* ``problems`` (initial data of the benchmark and test problems).

There is no CPU fallback: loading fails loudly when the library is missing.
"""
from ._lib import (  # noqa: F401
    SOLVER_ACOUSTICS,
    SOLVER_ADVECTION,
    SOLVER_EULER,
    SOLVER_VC_ACOUSTICS,
    SOLVER_VC_ADVECTION,
    ClawGpuError,
)

__all__ = ["SOLVER_ACOUSTICS", "SOLVER_ADVECTION", "SOLVER_EULER", "SOLVER_VC_ACOUSTICS",
           "SOLVER_VC_ADVECTION", "ClawGpuError"]
