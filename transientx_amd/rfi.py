"""RFI cleaning of channel x time frames on the device (tx_rfi_zdot,
tx_rfi_kadanef): Candidate::zdot and Candidate::kadaneF
(src/singlepulse/candidate.cpp:710-897), as replot_fil applies them to
candidate cutouts (src/singlepulse/replot_fil.cpp:409-447) and as the front
end's "zdot" / "kadaneF" items apply them to search blocks
(src/singlepulse/singlepulse.cpp:305-317).

Frames are float32 torch tensors on the device, cleaned in place on torch's
current stream:
  layout "channel": [nframes][nchan][nsamp] (or [nchan][nsamp]), the Candidate layout
  layout "time":    [nframes][nsamp][nchan] (or [nsamp][nchan]), the DataBuffer layout
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _frames(x, layout: str):
    import torch

    if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float32 and x.is_contiguous()):
        raise ValueError("frames must be a contiguous float32 CUDA tensor")
    shp = tuple(x.shape)
    if len(shp) == 2:
        shp = (1,) + shp
    if len(shp) != 3:
        raise ValueError("frames must be 2-D or 3-D")
    nf = shp[0]
    if layout == "channel":
        nchan, nsamp = shp[1], shp[2]
        cs, ts = nsamp, 1
    elif layout == "time":
        nsamp, nchan = shp[1], shp[2]
        cs, ts = 1, nchan
    else:
        raise ValueError(f"unknown layout {layout!r}")
    return nf, nchan * nsamp, nchan, nsamp, cs, ts


def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def zdot(x, layout: str = "channel"):
    """Candidate::zdot of every frame, in place: each channel normalised to
    zero mean and unit variance, then its least-squares fit to the channel
    mean removed.  Returns x."""
    import torch

    nf, fs, nchan, nsamp, cs, ts = _frames(x, layout)
    nb = L.lib().tx_rfi_work_bytes(L.TX_RFI_ZDOT, nf, nchan, nsamp, 1, 1)
    work = torch.empty(max(nb, 1), dtype=torch.uint8, device=x.device)
    L.check(L.lib().tx_rfi_zdot(_stream(), x.data_ptr(), nf, fs, nchan, nsamp, cs, ts, work.data_ptr(), nb))
    return x


def kadane_f(x, td: int, fd: int, threshold: float, wnlimit, layout: str = "channel"):
    """Candidate::kadaneF of every frame, in place (the frames are expected to
    be normalised).  wnlimit: one int per frame (or one for all), the
    reference's int(nwidth*width/tbin/td).  Returns x."""
    import torch

    nf, fs, nchan, nsamp, cs, ts = _frames(x, layout)
    wl = torch.as_tensor(np.broadcast_to(np.asarray(wnlimit, np.int32), (nf,)).copy(), device=x.device)
    nb = L.lib().tx_rfi_work_bytes(L.TX_RFI_KADANEF, nf, nchan, nsamp, td, fd)
    work = torch.empty(max(nb, 1), dtype=torch.uint8, device=x.device)
    L.check(L.lib().tx_rfi_kadanef(_stream(), x.data_ptr(), nf, fs, nchan, nsamp, cs, ts, td, fd, threshold,
                                   wl.data_ptr(), work.data_ptr(), nb))
    return x


def wnlimit(nwidth: int, width: float, tbin: float, td: int) -> int:
    """Candidate::kadaneF's run-length limit int(nwidth*width/tbin/td): the
    product nwidth*width in float (width is a float member), the rest in
    double, truncated towards zero."""
    return int(float(np.float32(np.float32(nwidth) * np.float32(width))) / tbin / td)
