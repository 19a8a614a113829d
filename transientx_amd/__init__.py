"""MI355X (gfx950) implementation of TransientX's single-pulse search hot path.

Two-stage sub-band dedispersion over a ddplan row followed by the boxcar
matched filter with S/N normalisation, as HIP kernels behind the C ABI of
include/txgpu.h.  `search` is the Python host over that ABI; the C++ host
(transientx_amd/host) provides the drop-in SubbandDedispersion / Boxcar /
SinglePulse classes and the transientx_fil command line.
"""
from .search import Beam, Plan, width_ladder  # noqa: F401

__all__ = ["Beam", "Plan", "width_ladder"]
