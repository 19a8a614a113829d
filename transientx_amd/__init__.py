"""MI355X (gfx950) implementation of TransientX's single-pulse search hot path.

Two-stage sub-band dedispersion over a ddplan row followed by the boxcar
matched filter with S/N normalisation, as HIP kernels behind the C ABI of
include/txgpu.h.  `search` is the Python host over that ABI and `host` its
host-side rows (ddplan/JSON, clustering, .cands); include/transientx_gpu.hpp
is the header-only C++ drop-in for SubbandDedispersion and Boxcar.
"""
from .search import Beam, Dump, Plan, width_ladder  # noqa: F401

__all__ = ["Beam", "Dump", "Plan", "width_ladder"]
