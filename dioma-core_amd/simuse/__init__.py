"""simuse — MI355X (gfx950) engine for the SiMuse per-timestep network step.

The compute path is the HIP library ``libsimuse.so`` (C ABI in
include/simuse.h); this package is its host-side mirror of the reference's
``pillars_cuda::CompiledNetwork`` API plus workload builders.
"""
from ._lib import (FORM_BINNED, FORM_CHAINED, FORM_ONE_KERNEL, FORM_SPLIT, MODEL_ALPHA_LIF, MODEL_GPEA, MODEL_HH,
                   MODEL_IZHIKEVICH, MODEL_LIF,
                   SCATTER_AUTO, SCATTER_TILE, SCATTER_WAVE, SimuseError, device_count, lib)
from .models import alpha_lif_pop, hh_pop, izh_pop, lif_pop, make_pop
from .network import CompiledNetwork, NetworkPop, stride_for

__all__ = [
    "CompiledNetwork", "NetworkPop", "SimuseError", "device_count", "lib", "stride_for",
    "lif_pop", "izh_pop", "hh_pop", "alpha_lif_pop", "make_pop",
    "MODEL_LIF", "MODEL_IZHIKEVICH", "MODEL_HH", "MODEL_GPEA", "MODEL_ALPHA_LIF", "SCATTER_AUTO", "SCATTER_TILE", "SCATTER_WAVE",
    "FORM_ONE_KERNEL", "FORM_BINNED", "FORM_CHAINED", "FORM_SPLIT",
]
