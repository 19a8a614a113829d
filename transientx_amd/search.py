"""Python host of the GPU search path, over the C ABI (include/txgpu.h).

Mirrors the pieces of TransientX's SinglePulse that touch the hot path:
  Beam  -- the block stream transientx_fil feeds (src/singlepulse/dedisperse_fil.cpp:319-338)
  Plan  -- SubbandDedispersion + Boxcar of one ddplan row
           (src/singlepulse/singlepulse.cpp:230-254, 328-330)
  width_ladder -- SinglePulse::prepare's boxcar widths (src/singlepulse/singlepulse.cpp:242-254)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def width_ladder(tsamp: float, minw: float, maxw: float, snrloss: float) -> np.ndarray:
    """Boxcar widths in samples, src/singlepulse/singlepulse.cpp:242-254.

    minw, maxw and snrloss are float members of SinglePulse; tsamp is the
    double Boxcar::tsamp, so the arithmetic mixes float32 and float64 exactly
    as the C++ does.
    """
    f32 = np.float32
    minw, maxw, snrloss = f32(minw), f32(maxw), f32(snrloss)
    if float(minw) < tsamp:
        minw = f32(tsamp)
    one_minus = 1.0 - float(snrloss)
    wfactor = f32(1.0 / (one_minus * one_minus))
    v = [int(np.round(float(minw) / tsamp))]
    while True:
        w1 = v[-1]
        w2 = int(f32(f32(w1) * wfactor))
        w2 = w1 + 1 if w2 < w1 + 1 else w2
        if w2 * tsamp > float(maxw):
            break
        v.append(w2)
    return np.asarray(v, np.int32)


def _dev_ptr(block):
    """(pointer, on_device, keepalive) for a numpy array or a torch tensor."""
    if isinstance(block, np.ndarray):
        a = np.ascontiguousarray(block)
        return a.ctypes.data_as(C.c_void_p), 0, a
    data_ptr = getattr(block, "data_ptr", None)
    if data_ptr is not None:
        if not block.is_contiguous():
            raise ValueError("block must be contiguous")
        return C.c_void_p(block.data_ptr()), int(block.is_cuda), block
    raise TypeError("block must be a numpy array or a torch tensor")


class Beam:
    def __init__(self, nchans: int, ndump: int, dtype: str = "u8", zero_point: int = 128,
                 chan_mask=None, device: int = 0):
        self.nchans = nchans
        self.ndump = ndump
        self.dtype = dtype
        d = L.BeamDesc()
        d.nchans = nchans
        d.dtype = L.TX_DTYPE_U8 if dtype == "u8" else L.TX_DTYPE_F32
        d.zero_point = zero_point
        self._mask = None
        if chan_mask is not None:
            self._mask = np.ascontiguousarray(chan_mask, np.uint8)
            d.chan_mask = self._mask.ctypes.data_as(C.POINTER(C.c_uint8))
        d.ndump = ndump
        d.device = device
        h = C.c_void_p()
        L.check(L.lib().tx_beam_create(C.byref(d), C.byref(h)))
        self.h = h
        self.plans = []
        self.frontend = None

    def push(self, block):
        p, on_dev, keep = _dev_ptr(block)
        ns = block.shape[0]
        L.check(L.lib().tx_beam_push(self.h, p, ns, on_dev))
        del keep

    def set_frontend(self, nchans_in: int, td: int, fd: int, offset, gain, in_load=None):
        """Take raw 8-bit blocks [ndump*td][nchans_in] from now on: the front
        end's td x fd neighbour sums, then (y - offset) * gain per channel,
        formed on the device straight into this f32 beam's ring
        (tx_beam_set_frontend).  nchans_in == 0 switches it off.  in_load
        (before the first push): True keeps raw 8-bit samples in the ring
        and applies the front end in stage 1's load when td = fd = 1
        (tx_beam_set_frontend_in_load); same results, a quarter of the
        ring's memory, more stage-1 time."""
        if in_load is not None:
            L.check(L.lib().tx_beam_set_frontend_in_load(self.h, int(bool(in_load))))
        if nchans_in == 0:
            L.check(L.lib().tx_beam_set_frontend(self.h, 0, 1, 1, None, None))
            self.frontend = None
            return
        off = np.ascontiguousarray(offset, np.float32)
        g = np.ascontiguousarray(gain, np.float32)
        if off.size != self.nchans or g.size != self.nchans:
            raise ValueError("offset and gain need one value per beam channel")
        L.check(L.lib().tx_beam_set_frontend(self.h, nchans_in, td, fd, L.ptr(off), L.ptr(g)))
        self.frontend = (nchans_in, td, fd)

    def set_rfi(self, ops):
        """The front end's RFI filters on every pushed block, in order
        (tx_beam_set_rfi): ops is a list of ("zdot",) and ("kadaneF", td,
        fd, threshold, wnlimit) items; [] removes them."""
        n = len(ops)
        code = np.zeros(max(n, 1), np.int32)
        td = np.ones(max(n, 1), np.int32)
        fd = np.ones(max(n, 1), np.int32)
        thr = np.zeros(max(n, 1), np.float32)
        wnl = np.zeros(max(n, 1), np.int32)
        for k, o in enumerate(ops):
            if o[0] == "zdot":
                code[k] = L.TX_RFI_ZDOT
            elif o[0] == "kadaneF":
                code[k] = L.TX_RFI_KADANEF
                td[k], fd[k], thr[k], wnl[k] = o[1], o[2], o[3], o[4]
            else:
                raise ValueError(f"unknown RFI filter {o[0]!r}")
        L.check(L.lib().tx_beam_set_rfi(self.h, n, L.ptr(code), L.ptr(td), L.ptr(fd), L.ptr(thr), L.ptr(wnl)))

    def push_async(self, block):
        """Append a block held in pinned host memory (a torch CPU tensor from
        pin_memory) without waiting for its copy; the tensor must stay
        unchanged until the next sync()."""
        if getattr(block, "is_cuda", True) or not block.is_pinned() or not block.is_contiguous():
            raise ValueError("push_async takes a contiguous pinned host tensor")
        L.check(L.lib().tx_beam_push(self.h, C.c_void_p(block.data_ptr()), block.shape[0],
                                     L.TX_PUSH_HOST_ASYNC))

    def build_graph(self, thre: float):
        """Capture one block of every plan (all in session mode) as a
        hipGraph (tx_beam_graph); graph_push then runs a block through it."""
        L.check(L.lib().tx_beam_graph(self.h, thre))

    def graph_push(self, block, host_async: bool = False):
        """Push + dedisperse + boxcar of every plan for one block, one graph
        launch (tx_beam_graph_push).  host_async: a pinned host tensor the
        caller leaves unchanged until the next sync()."""
        if host_async:
            if getattr(block, "is_cuda", True) or not block.is_pinned() or not block.is_contiguous():
                raise ValueError("host_async takes a contiguous pinned host tensor")
            L.check(L.lib().tx_beam_graph_push(self.h, C.c_void_p(block.data_ptr()), block.shape[0],
                                               L.TX_PUSH_HOST_ASYNC))
            return
        p, on_dev, keep = _dev_ptr(block)
        L.check(L.lib().tx_beam_graph_push(self.h, p, block.shape[0], on_dev))
        del keep

    def sync(self):
        L.check(L.lib().tx_beam_sync(self.h))

    def reset(self):
        """Start a fresh stream (counter 0, rings back to the zero signal)."""
        L.check(L.lib().tx_beam_reset(self.h))

    @property
    def stream(self) -> int:
        s = C.c_void_p()
        L.check(L.lib().tx_beam_stream(self.h, C.byref(s)))
        return s.value or 0

    def set_timing(self, enable: bool):
        L.check(L.lib().tx_beam_set_timing(self.h, int(enable)))

    def timing(self):
        """{kernel: (total_ms, launches)} since the last call."""
        ms = np.zeros(4)
        n = np.zeros(4, np.int64)
        L.check(L.lib().tx_beam_timing(self.h, L.ptr(ms), L.ptr(n)))
        names = ("transpose", "stage1", "stage2", "boxcar")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(names)}

    def close(self):
        for p in self.plans:
            p.close()
        self.plans = []
        self.frontend = None
        if self.h:
            L.lib().tx_beam_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    def __init__(self, beam: Beam, freqs, tsamp: float, dms: float, ddm: float, ndm: int, vwn,
                 overlap: float = 0.0, nsub: int = 0, group: int = 0, iqr: bool = False,
                 store_maps: bool = True, max_points: int = 0, exact_order: bool = False,
                 shapes: dict | None = None):
        self.beam = beam
        self._freqs = np.ascontiguousarray(freqs, np.float64)
        self._vwn = np.ascontiguousarray(vwn, np.int32)
        d = L.PlanDesc()
        d.frequencies = self._freqs.ctypes.data_as(C.POINTER(C.c_double))
        d.tsamp = tsamp
        d.dms = dms
        d.ddm = ddm
        d.ndm = ndm
        d.overlap = overlap
        d.nsub = nsub
        d.group = group
        d.vwn = self._vwn.ctypes.data_as(C.POINTER(C.c_int))
        d.nbox = self._vwn.size
        d.iqr = int(iqr)
        d.store_maps = int(store_maps)
        d.max_points = max_points
        d.exact_order = int(exact_order)
        # launch shapes (an autotuner's choice): s1_ncc, s1_q, s2_shape, s2_lds
        for k, v in (shapes or {}).items():
            setattr(d, k, int(v))
        h = C.c_void_p()
        L.check(L.lib().tx_plan_create(beam.h, C.byref(d), C.byref(h)))
        self.h = h
        self.tsamp = tsamp
        self.dms = dms
        self.ddm = ddm
        beam.plans.append(self)

    def close(self):
        if self.h:
            L.lib().tx_plan_destroy(self.h)
            self.h = None

    @property
    def info(self) -> L.Info:
        o = L.Info()
        L.check(L.lib().tx_get_info(self.h, C.byref(o)))
        return o

    def tables(self):
        o = self.info
        d1 = np.zeros((o.ndmc, self.beam.nchans), np.int32)
        d2 = np.zeros((o.ndm, o.nsub), np.int32)
        dmc = np.zeros(o.ndmc)
        fref = np.zeros(o.nsub)
        fsub = np.zeros(o.nsub)
        fcnt = np.zeros(o.nsub)
        L.check(L.lib().tx_get_tables(self.h, L.ptr(d1), L.ptr(d2), L.ptr(dmc), L.ptr(fref),
                                      L.ptr(fsub), L.ptr(fcnt)))
        return dict(d1=d1, d2=d2, dmc=dmc, fref=fref, fsub=fsub, fcnt=fcnt)

    def dedisp_run(self):
        L.check(L.lib().tx_dedisp_run(self.h))

    def boxcar_run(self, thre: float) -> bool:
        return L.check(L.lib().tx_boxcar_run(self.h, thre)) == L.TX_OK

    def points(self) -> np.ndarray:
        n = C.c_size_t()
        rc = L.lib().tx_boxcar_points(self.h, None, 0, C.byref(n))
        if rc < 0 and rc != L.TX_EOVERFLOW:
            L.check(rc)
        out = np.zeros(n.value, L.POINT_DTYPE)
        L.check(L.lib().tx_boxcar_points(self.h, out.ctypes.data_as(C.c_void_p), out.size,
                                         C.byref(n)))
        return out

    def timdata(self, idm: int, with_overlap: bool = True) -> np.ndarray:
        o = self.info
        out = np.zeros(o.nsamples if with_overlap else o.ndump, np.float32)
        L.check(L.lib().tx_get_timdata(self.h, idm, int(with_overlap), L.ptr(out)))
        return out

    def subdata(self, idm: int) -> np.ndarray:
        o = self.info
        out = np.zeros((o.nsub, o.nsamples), np.float32)
        L.check(L.lib().tx_get_subdata(self.h, idm, L.ptr(out)))
        return out

    def rows(self, dm0: int, dm1: int):
        o = self.info
        S = np.zeros((dm1 - dm0, o.nsamples), np.float32)
        w = np.zeros((dm1 - dm0, o.nsamples), np.int32)
        L.check(L.lib().tx_boxcar_rows(self.h, dm0, dm1, L.ptr(S), L.ptr(w)))
        return S, w

    def cand_spectra(self, idm, isamp, wn) -> np.ndarray:
        idm = np.ascontiguousarray(idm, np.int32)
        isamp = np.ascontiguousarray(isamp, np.int32)
        wn = np.ascontiguousarray(wn, np.int32)
        out = np.zeros((idm.size, self.info.nsub), np.float32)
        L.check(L.lib().tx_cand_spectra(self.h, idm.size, L.ptr(idm), L.ptr(isamp), L.ptr(wn),
                                        L.ptr(out)))
        return out

    def session_begin(self, cap: int, host=None):
        """Session mode (tx_session_begin).  host: a pinned torch CPU tensor
        of at least cap*24 bytes that receives the points directly as the
        kernels find them (tx_session_begin_host); keep it alive and unread
        until session_end / session_end_into."""
        if host is None:
            L.check(L.lib().tx_session_begin(self.h, cap))
            self._shost = None
        else:
            if getattr(host, "is_cuda", True) or not host.is_pinned() or not host.is_contiguous() \
                    or host.numel() * host.element_size() < cap * L.POINT_EX_DTYPE.itemsize:
                raise ValueError("host must be a contiguous pinned CPU tensor of at least cap points")
            L.check(L.lib().tx_session_begin_host(self.h, C.c_void_p(host.data_ptr()), cap))
            self._shost = host
        self._scap = cap

    def session_end(self) -> np.ndarray:
        """End the session: every point since session_begin, device order."""
        return self.session_end_points(self._scap)

    def session_end_into(self, host_ptr: int, cap: int) -> int:
        """End the session, copying up to cap tx_point_ex records to host_ptr
        (pinned memory copies fastest); returns the number of points found."""
        n = C.c_size_t()
        self._shost = None
        L.check(L.lib().tx_session_end(self.h, C.c_void_p(host_ptr), cap, C.byref(n)))
        return n.value

    def session_end_points(self, cap: int) -> np.ndarray:
        n = C.c_size_t()
        if getattr(self, "_shost", None) is not None:
            # the points are already in the host list
            L.check(L.lib().tx_session_end(self.h, None, 0, C.byref(n)))
            h = self._shost.numpy().view(np.uint8)[: self._scap * L.POINT_EX_DTYPE.itemsize]
            self._shost = None
            return h.view(L.POINT_EX_DTYPE)[: min(n.value, cap)].copy()
        out = np.zeros(cap, L.POINT_EX_DTYPE)
        L.check(L.lib().tx_session_end(self.h, out.ctypes.data_as(C.c_void_p), cap, C.byref(n)))
        return out[: min(n.value, cap)]

    def device_buffers(self):
        t, s, w = C.c_void_p(), C.c_void_p(), C.c_void_p()
        L.check(L.lib().tx_device_buffers(self.h, C.byref(t), C.byref(s), C.byref(w)))
        return t.value, s.value, w.value

    def dump_series(self, mean: float, std: float, nbits: int):
        """The current run's --savetim values (declared rule of
        include/txgpu.h): (tau0, [ndm][n] int8 | int16 | float32), n = 0
        while every new sample precedes the observation's start."""
        o = self.info
        dt = {8: np.int8, 16: np.int16, 32: np.float32}[nbits]
        out = np.zeros((o.ndm, o.ndump), dt)
        tau0, n = C.c_long(), C.c_long()
        L.check(L.lib().tx_dump_series(self.h, mean, std, nbits, L.ptr(out), out.nbytes, C.byref(tau0),
                                       C.byref(n)))
        return tau0.value, out[:, : n.value]


_DUMP_FORMATS = {"pulsarx": L.TX_DUMP_PULSARX, "sigproc": L.TX_DUMP_SIGPROC, "presto": L.TX_DUMP_PRESTO}


class Dump:
    """--savetim files of one ddplan row (SubbandDedispersion::preparedump,
    rundump, makeinf / modifynblock; src/singlepulse/singlepulse.cpp:238,340,
    src/singlepulse/dedisperse_fil.cpp:356-369), written by the library's
    tx_dump_* calls in the declared formats of include/txgpu.h."""

    def __init__(self, plan: Plan, rootname: str, fmt: str = "pulsarx", nbits: int = 8, tstart: float = 0.0,
                 fch1: float = 0.0, foff: float = 0.0, nchans: int = 0, ibeam: int = 0,
                 source_name: str = "J0000-00"):
        if fmt == "presto":
            nbits = 32  # src/singlepulse/singlepulse.cpp:228
        d = L.DumpDesc(rootname=rootname.encode(), source_name=source_name.encode(), format=_DUMP_FORMATS[fmt],
                       nbits=nbits, tstart=tstart, fch1=fch1, foff=foff, nchans=nchans, ibeam=ibeam)
        h = C.c_void_p()
        L.check(L.lib().tx_dump_open(plan.h, C.byref(d), C.byref(h)))
        self.h = h
        self.nbits = nbits
        self.fmt = fmt

    def block(self, mean: float = 0.0, std: float = 3.0):
        L.check(L.lib().tx_dump_block(self.h, mean, std))

    def finish(self):
        L.check(L.lib().tx_dump_finish(self.h))

    def close(self):
        if self.h:
            L.lib().tx_dump_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
