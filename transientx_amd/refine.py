"""The candidate refinement grid on the GPU (include/txgpu.h tx_refine_*):
Candidate::dedisperse, Candidate::optimize with the dedispersion tree and
Candidate::matched_filter (src/singlepulse/candidate.cpp:108-140,899-1087)
for a batch of cutouts resident on the device."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class Refiner:
    """Refinement grid for cutouts of one shape [nchan][nbin]."""

    def __init__(self, freqs, nbin: int, tbin: float, snrloss: float = 0.1,
                 dedisperse: bool = True, isdedispersed: bool = True, max_batch: int = 256,
                 device: int = 0):
        self._f = np.ascontiguousarray(freqs, np.float64)
        self.nchan, self.nbin, self.tbin = self._f.size, int(nbin), float(tbin)
        d = _lib.RefineDesc(self.nchan, self.nbin, self._f.ctypes.data_as(C.POINTER(C.c_double)),
                            self.tbin, snrloss, int(dedisperse), int(isdedispersed), device)
        h = C.c_void_p()
        _lib.check(_lib.lib().tx_refine_create(C.byref(d), max_batch, C.byref(h)))
        self.h = h
        nrow, nbox, npass = C.c_int(), C.c_int(), C.c_int()
        _lib.check(_lib.lib().tx_refine_info(self.h, C.byref(nrow), C.byref(nbox), C.byref(npass)))
        self.ndm, self.nbox, self.npass = nrow.value, nbox.value, npass.value

    def stream(self) -> int:
        s = C.c_void_p()
        _lib.check(_lib.lib().tx_refine_stream(self.h, C.byref(s)))
        return s.value or 0

    def run(self, cutouts, dm, maps: bool = False):
        """cutouts: torch CUDA float32 tensor [ncand][nchan][nbin]; dm: the
        candidates' DMs.  Returns the results (REFINE_RESULT_DTYPE) and, with
        maps, the (dmtmap, snrmap, widmap) device tensors."""
        import torch

        assert cutouts.is_cuda and cutouts.dtype == torch.float32 and cutouts.is_contiguous()
        ncand = cutouts.shape[0]
        assert tuple(cutouts.shape[1:]) == (self.nchan, self.nbin)
        dmv = np.ascontiguousarray(dm, np.float64)
        assert dmv.size == ncand
        res = np.zeros(ncand, _lib.REFINE_RESULT_DTYPE)
        out = (None, None, None)
        if maps:
            shp = (ncand, self.ndm, self.nbin)
            out = (torch.empty(shp, dtype=torch.float32, device=cutouts.device),
                   torch.empty(shp, dtype=torch.float32, device=cutouts.device),
                   torch.empty(shp, dtype=torch.int32, device=cutouts.device))
        torch.cuda.synchronize(cutouts.device)
        ptrs = [o.data_ptr() if o is not None else None for o in out]
        _lib.check(_lib.lib().tx_refine_run(self.h, ncand, cutouts.data_ptr(), _lib.ptr(dmv),
                                            _lib.ptr(res), *ptrs))
        return (res, out) if maps else res

    def close(self):
        if self.h:
            _lib.lib().tx_refine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
