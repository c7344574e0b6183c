"""pmmc_amd -- MI355X-native (gfx950) simulated-annealing hot path of pMMC.

The compute path is libpmmc_hip.so (hand-written HIP kernels behind the C ABI
declared in include/pmmc_hip.h).  This package only binds it.
"""
from . import _abi  # noqa: F401
from .api import (Arcs, Context, Heatmap, PmmcError, Smooth, schedule_preset,  # noqa: F401
                  settings)
from ._abi import (LEVEL_ARCS, LEVEL_HEATMAP, LEVEL_SMOOTH, PRESET_CPU,  # noqa: F401
                   PRESET_GPU)

__all__ = ["Arcs", "Context", "Heatmap", "PmmcError", "Smooth", "schedule_preset", "settings",
           "LEVEL_ARCS", "LEVEL_HEATMAP", "LEVEL_SMOOTH", "PRESET_CPU", "PRESET_GPU"]
