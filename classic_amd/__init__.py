"""classic_amd -- MI355X (gfx950) wave-propagation step for Clawpack Classic.

The hot path is the fused HIP kernel in ``csrc/step2d.hip`` behind the C-ABI
of ``include/clawgpu.h`` (``libclawgpu.so``, built in-tree by
``python -m classic_amd.build``); ``gpu`` is its Python front end and
``problems`` holds the initial data of the benchmark problems.
"""
from ._lib import SOLVER_ACOUSTICS, SOLVER_EULER, ClawGpuError  # noqa: F401

__all__ = ["SOLVER_ACOUSTICS", "SOLVER_EULER", "ClawGpuError"]
