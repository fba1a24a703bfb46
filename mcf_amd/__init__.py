"""MI355X-native SDPD pair-interaction step (density, forces, integration).

A drop-in for the per-step hot path of MCF (nmorenoc/MCF): cell list,
density + equation of state, pair forces with thermal noise and the
velocity-Verlet update, as hand-written HIP kernels for gfx950 behind the
C-ABI in include/sdpd.h.
"""
from .params import PRECISION_FP64, PRECISION_MIXED, SDPDParams  # noqa: F401

__all__ = ["SDPDParams", "PRECISION_FP64", "PRECISION_MIXED"]
