"""The front end's Downsample and Equalize on the device for raw 8-bit
blocks (tx_frontend_u8, tx_frontend_stats, tx_beam_set_frontend): what
SinglePulse::run applies before dedispersion
(src/singlepulse/singlepulse.cpp:289-291, prepared at :208-216, with the
--td/--fd options of src/singlepulse/dedisperse_fil.cpp:54-55).  XLibs'
Downsample and Equalize are not in the TransientX tree, so the rule is
declared: output channel c < nchans/fd, sample t < nsamp/td takes y, the sum
of its td x fd neighbouring 8-bit samples (docs/pipeline/preprocess.rst:5-9,
exact), then (y - offset[c]) * gain[c] in float32; masked channels are 0.
Equalize's statistics: offset = mean, gain = 1/sd (1 for a constant
channel) from exact integer sums over a block.

Blocks are uint8 CUDA tensors, time-major [nsamp][nchans] (the filterbank
and DataBuffer layout); work runs on torch's current stream.
"""
from __future__ import annotations

import ctypes as C

from . import _lib as L


def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _raw(x):
    import torch

    if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.uint8 and x.is_contiguous()
            and x.dim() == 2):
        raise ValueError("a raw block must be a contiguous 2-D uint8 CUDA tensor [nsamp][nchans]")
    return x.shape[0], x.shape[1]


def _f32(v, n, dev):
    import torch

    t = torch.as_tensor(v, dtype=torch.float32, device=dev).contiguous()
    if t.numel() != n:
        raise ValueError(f"need {n} values, got {t.numel()}")
    return t


def stats(x, td: int = 1, fd: int = 1):
    """Equalize's (offset, gain) of one raw block: float32 CUDA tensors of
    nchans/fd values."""
    import torch

    ns, nch = _raw(x)
    ncho = nch // fd if fd >= 1 else 0
    off = torch.empty(max(ncho, 1), dtype=torch.float32, device=x.device)
    gain = torch.empty_like(off)
    nb = L.lib().tx_frontend_stats_work_bytes(nch, fd)
    work = torch.empty(max(nb, 8), dtype=torch.uint8, device=x.device)
    L.check(L.lib().tx_frontend_stats(_stream(), x.data_ptr(), ns, nch, td, fd, off.data_ptr(),
                                      gain.data_ptr(), work.data_ptr(), nb))
    return off[:ncho], gain[:ncho]


def apply(x, td: int, fd: int, offset, gain, chan_mask=None):
    """The front end's output for one raw block: float32 CUDA tensor
    [nsamp/td][nchans/fd] (time-major, the DataBuffer layout)."""
    import torch

    ns, nch = _raw(x)
    if td < 1 or fd < 1 or nch < fd:
        raise ValueError("td, fd >= 1 and nchans >= fd")
    ncho = nch // fd
    off = _f32(offset, ncho, x.device)
    g = _f32(gain, ncho, x.device)
    m = None
    if chan_mask is not None:
        m = torch.as_tensor(chan_mask, dtype=torch.uint8, device=x.device).contiguous()
        if m.numel() != ncho:
            raise ValueError("one mask byte per output channel")
    out = torch.empty((ns // td, ncho), dtype=torch.float32, device=x.device)
    L.check(L.lib().tx_frontend_u8(_stream(), x.data_ptr(), ns, nch, td, fd, off.data_ptr(), g.data_ptr(),
                                   None if m is None else m.data_ptr(), out.data_ptr()))
    return out
