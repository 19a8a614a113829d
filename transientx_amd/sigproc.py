"""SIGPROC filterbank input for the search (tx_fil_* in include/txgpu.h):
the reader transientx_fil opens and reads block by block
(src/singlepulse/dedisperse_fil.cpp:167-195 open and header, :214-217
nchans/tsamp/nsamples, :278-283 --jump, :324 read_data), and a writer of
files in the same format for synthetic inputs.  The reader is XLibs' and
not in the TransientX tree; the format is the public SIGPROC one (8-bit
unsigned or 32-bit float samples, one IF).  Host code: blocks read here go
to Beam.push like any host block.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _lib as L

_FIELDS = [f for f, _ in L.FilHeader._fields_]


def _to_dict(h: L.FilHeader) -> dict:
    d = {f: getattr(h, f) for f in _FIELDS}
    d["source_name"] = h.source_name.decode()
    d["rawdatafile"] = h.rawdatafile.decode()
    return d


def read_header(path) -> dict:
    """The header of one file as a dict (nsamples from the file's size,
    header_bytes where the samples start)."""
    h = L.FilHeader()
    L.check(L.lib().tx_fil_read_header(str(path).encode(), C.byref(h)))
    return _to_dict(h)


def write(path, data: np.ndarray, *, tsamp: float, fch1: float, foff: float, tstart: float = 60000.0,
          source_name: str = "", telescope_id: int = 0, machine_id: int = 0, ibeam: int = 0, nbeams: int = 1,
          src_raj: float = 0.0, src_dej: float = 0.0, rawdatafile: str = ""):
    """Writes data [nsamples][nchans] (uint8 or float32) as a SIGPROC
    filterbank file: tx_fil_write_header, then the samples in order."""
    x = np.ascontiguousarray(data)
    if x.ndim != 2 or x.dtype not in (np.uint8, np.float32):
        raise ValueError("data must be [nsamples][nchans] uint8 or float32")
    h = L.FilHeader()
    h.source_name = source_name.encode()
    h.rawdatafile = rawdatafile.encode()
    h.telescope_id, h.machine_id, h.data_type = telescope_id, machine_id, 1
    h.nbits = 8 if x.dtype == np.uint8 else 32
    h.nchans, h.nifs, h.nbeams, h.ibeam = x.shape[1], 1, nbeams, ibeam
    h.tstart, h.tsamp, h.fch1, h.foff = tstart, tsamp, fch1, foff
    h.src_raj, h.src_dej = src_raj, src_dej
    L.check(L.lib().tx_fil_write_header(str(path).encode(), C.byref(h)))
    with open(path, "ab") as f:
        x.tofile(f)


class FilReader:
    """Files read in order as one stream (tx_fil_open).  check_continuity:
    each file must start where the previous one ends (transientx_fil checks
    this unless --cont).  jump: (seconds at the start, seconds at the end)
    left unread, as --jump sets skip_start = (long)(j0/tsamp) and skip_end =
    ntotal - (long)(ntotal - j1/tsamp) (dedisperse_fil.cpp:278-283)."""

    def __init__(self, paths, check_continuity: bool = True, jump=(0.0, 0.0)):
        if isinstance(paths, (str, Path)):
            paths = [paths]
        enc = [str(p).encode() for p in paths]
        arr = (C.c_char_p * len(enc))(*enc)
        h = L.FilHeader()
        self.h = C.c_void_p()
        L.check(L.lib().tx_fil_open(arr, len(enc), int(bool(check_continuity)), C.byref(self.h), C.byref(h)))
        self.header = _to_dict(h)
        self.nchans = h.nchans
        self.nbits = h.nbits
        self.tsamp = h.tsamp
        self.nsamples = h.nsamples
        self.dtype = np.uint8 if h.nbits == 8 else np.float32
        self.freqs = np.zeros(self.nchans, np.float64)
        L.check(L.lib().tx_fil_frequencies(self.h, L.ptr(self.freqs)))
        ntotal = self.nsamples
        nstart = int(jump[0] / self.tsamp)
        nend = int(ntotal - jump[1] / self.tsamp)
        self.skip_start, self.skip_end = nstart, ntotal - nend
        L.check(L.lib().tx_fil_skip(self.h, self.skip_start, self.skip_end))

    def read(self, n: int, out: np.ndarray | None = None) -> np.ndarray:
        """Up to n samples [k][nchans] (k < n only at the end of the stream),
        into out (a contiguous host array of at least n rows) when given."""
        if out is None:
            out = np.empty((n, self.nchans), self.dtype)
        if out.dtype != self.dtype or out.ndim != 2 or out.shape[1] != self.nchans or out.shape[0] < n \
                or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous [>= n][nchans] array of the file's sample type")
        k = L.lib().tx_fil_read(self.h, out.ctypes.data_as(C.c_void_p), n)
        L.check(int(k))
        return out[:k]

    def blocks(self, ndump: int, pinned: bool = False):
        """Whole blocks of ndump samples, as transientx_fil reads them: a last
        block shorter than ndump ends the loop unread (dedisperse_fil.cpp:324).
        pinned: each block is a fresh page-locked torch tensor, for
        Beam.push_async; otherwise one numpy buffer is reused, so a block is
        valid until the next one is read."""
        buf = None
        while True:
            if pinned:
                import torch

                t = torch.empty((ndump, self.nchans), dtype=torch.uint8 if self.nbits == 8 else torch.float32,
                                pin_memory=True)
                got = self.read(ndump, t.numpy())
                if len(got) != ndump:
                    return
                yield t
            else:
                if buf is None:
                    buf = np.empty((ndump, self.nchans), self.dtype)
                got = self.read(ndump, buf)
                if len(got) != ndump:
                    return
                yield buf

    def close(self):
        if self.h:
            L.lib().tx_fil_close(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
