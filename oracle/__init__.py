"""CPU oracle of the TransientX search hot path.  TEST INFRASTRUCTURE ONLY.

numpy wrappers around oracle/liboracle.so (txoracle.c) and, when built,
oracle/_ref/libhough_ref.so (TransientX's include/hough_transform.h compiled
from the reference tree).  Only tests/, the smoke entry point and the
benchmark's CPU-baseline leg import this package; the product never does.

Parity status of each function is listed in txoracle.c's header.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_lib = None
_ref = None


class PlanIn(C.Structure):
    _fields_ = [
        ("nchans", C.c_int),
        ("freqs", C.POINTER(C.c_double)),
        ("tsamp", C.c_double),
        ("ndump", C.c_long),
        ("overlap", C.c_double),
        ("dms", C.c_double),
        ("ddm", C.c_double),
        ("ndm", C.c_long),
        ("nsub", C.c_int),
        ("group", C.c_int),
    ]


class PlanInfo(C.Structure):
    _fields_ = [
        ("noverlap", C.c_long),
        ("nsamples", C.c_long),
        ("nsub", C.c_int),
        ("cps", C.c_int),
        ("group", C.c_int),
        ("ndmc", C.c_int),
        ("D1", C.c_long),
        ("D2", C.c_long),
        ("maxdelay", C.c_long),
        ("offset", C.c_long),
        ("fmax", C.c_double),
        ("fmin", C.c_double),
    ]


def lib():
    global _lib
    if _lib is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError("oracle/liboracle.so missing: run `python oracle/build.py`")
        L = C.CDLL(str(path))
        vp = C.c_void_p
        L.txo_dmdelay.restype = C.c_double
        L.txo_dmdelay.argtypes = [C.c_double] * 3
        L.txo_ladder.argtypes = [C.c_double, C.c_float, C.c_float, C.c_float, vp, C.c_int]
        L.txo_plan_info_get.argtypes = [C.POINTER(PlanIn), C.POINTER(PlanInfo)]
        L.txo_plan_tables.argtypes = [C.POINTER(PlanIn), vp, vp, vp, vp]
        L.txo_dedisp_window.argtypes = [C.POINTER(PlanIn), vp, C.c_long, C.c_long, vp, vp, C.c_int]
        L.txo_dedisp_window_f64.argtypes = [C.POINTER(PlanIn), vp, C.c_long, C.c_long, vp, C.c_int]
        L.txo_boxcar_f64.argtypes = [vp, C.c_long, C.c_long, vp, C.c_int, vp, C.c_int]
        L.txo_dedisp_direct_f64.argtypes = [C.c_int, vp, C.c_double, C.c_double, C.c_double,
                                            C.c_long, vp, C.c_long, C.c_long, C.c_long, vp]
        L.txo_boxcar_row.argtypes = [vp, C.c_long, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp,
                                     vp, vp]
        L.txo_boxcar.argtypes = [vp, C.c_long, C.c_long, vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                 vp, vp, C.c_int]
        L.txo_boxcar_runnerup.argtypes = [vp, C.c_long, C.c_long, vp, C.c_int, C.c_int, C.c_long, vp,
                                          vp, vp, vp]
        L.txo_dump.restype = None
        L.txo_dump.argtypes = [vp, C.c_long, C.c_long, C.c_long, C.c_double, C.c_float, C.c_float, C.c_int, vp]
        L.txo_threshold.restype = C.c_long
        L.txo_threshold.argtypes = [vp, vp, C.c_long, C.c_long, C.c_float, vp, vp, vp, vp,
                                    C.c_long]
        L.txo_cluster.argtypes = [C.c_long, vp, vp, vp, vp, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                  C.c_int, vp, vp, vp]
        L.txo_kadane.restype = None
        L.txo_kadane.argtypes = [vp, C.c_long, vp, vp]
        L.txo_cand_flfh.restype = None
        L.txo_cand_flfh.argtypes = [vp, vp, vp, C.c_int, C.c_long, C.c_long, C.c_int, vp, vp]
        L.txo_stream_create.restype = vp
        L.txo_stream_create.argtypes = [C.POINTER(PlanIn), vp, C.c_int, C.c_int, C.c_int]
        L.txo_stream_destroy.restype = None
        L.txo_stream_destroy.argtypes = [vp]
        L.txo_stream_run_u8.restype = C.c_long
        L.txo_stream_run_u8.argtypes = [vp, vp, C.c_float]
        L.txo_stream_tim.restype = vp
        L.txo_stream_tim.argtypes = [vp]
        L.txo_stream_mxS.restype = vp
        L.txo_stream_mxS.argtypes = [vp]
        L.txo_stream_mxwn.restype = vp
        L.txo_stream_mxwn.argtypes = [vp]
        L.txo_cand_dedisperse.restype = None
        L.txo_cand_dedisperse.argtypes = [vp, C.c_int, C.c_long, vp, C.c_double, C.c_double]
        L.txo_hough.argtypes = [vp, vp, C.c_int, C.c_int, vp, C.c_double, C.c_double, C.c_double,
                                C.c_double]
        L.txo_refine_optimize.argtypes = [vp, C.c_int, C.c_int, vp, C.c_double, C.c_double, C.c_int,
                                          vp, C.c_int, vp, vp, vp]
        L.txo_mean_var.restype = None
        L.txo_mean_var.argtypes = [vp, C.c_long, vp, vp]
        L.txo_mf_ladder.argtypes = [C.c_long, C.c_float, vp, C.c_int]
        L.txo_zdot.restype = None
        L.txo_zdot.argtypes = [vp, C.c_long, C.c_long]
        L.txo_kadanef.restype = None
        L.txo_kadanef.argtypes = [vp, C.c_long, C.c_long, C.c_int, C.c_int, C.c_float, C.c_int]
        L.txo_refine_mf.argtypes = [vp, C.c_int, C.c_long, C.c_float, vp, vp, vp, vp, vp]
        L.txo_frontend_u8.restype = None
        L.txo_frontend_u8.argtypes = [vp, C.c_long, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]
        L.txo_frontend_stats.restype = None
        L.txo_frontend_stats.argtypes = [vp, C.c_long, C.c_int, C.c_int, C.c_int, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _chk(rc):
    if rc < 0:
        raise RuntimeError(f"oracle error {rc}")
    return rc


def dmdelay(dm, fh, fl):
    return lib().txo_dmdelay(dm, fh, fl)


def ladder(tsamp, minw, maxw, snrloss):
    out = np.zeros(4096, np.int32)
    n = _chk(lib().txo_ladder(tsamp, minw, maxw, snrloss, _p(out), out.size))
    return out[:n].copy()


@dataclass
class Plan:
    freqs: np.ndarray
    tsamp: float
    ndump: int
    dms: float
    ddm: float
    ndm: int
    overlap: float = 0.0
    nsub: int = 0
    group: int = 0

    def _in(self):
        self._f = np.ascontiguousarray(self.freqs, np.float64)
        return PlanIn(len(self._f), self._f.ctypes.data_as(C.POINTER(C.c_double)), self.tsamp,
                      self.ndump, self.overlap, self.dms, self.ddm, self.ndm, self.nsub,
                      self.group)

    def info(self) -> PlanInfo:
        o = PlanInfo()
        pi = self._in()
        _chk(lib().txo_plan_info_get(C.byref(pi), C.byref(o)))
        return o

    def tables(self):
        o = self.info()
        C_ = len(self.freqs)
        d1 = np.zeros((o.ndmc, C_), np.int32)
        d2 = np.zeros((self.ndm, o.nsub), np.int32)
        dmc = np.zeros(o.ndmc)
        fref = np.zeros(o.nsub)
        pi = self._in()
        _chk(lib().txo_plan_tables(C.byref(pi), _p(d1), _p(d2), _p(dmc), _p(fref)))
        return d1, d2, dmc, fref

    def window_start(self, block: int) -> int:
        """w0 of 1-based block `block` (counter = block*ndump)."""
        o = self.info()
        return block * self.ndump - o.maxdelay - o.nsamples

    def dedisperse(self, x: np.ndarray, w0: int, with_sub: bool = False, nthreads: int = 8):
        """Two-stage FP32 series [ndm][N] of window w0; x is time-major [ntot][nchans] f32."""
        o = self.info()
        x = np.ascontiguousarray(x, np.float32)
        tim = np.zeros((self.ndm, o.nsamples), np.float32)
        sub = np.zeros((self.ndm, o.nsub, o.nsamples), np.float32) if with_sub else None
        pi = self._in()
        _chk(lib().txo_dedisp_window(C.byref(pi), _p(x), x.shape[0], w0, _p(tim),
                                     _p(sub) if with_sub else None, nthreads))
        return (tim, sub) if with_sub else tim

    def dedisperse_f64(self, x: np.ndarray, w0: int, nthreads: int = 8):
        """The same two-stage sums accumulated in FP64: [ndm][N] float64."""
        o = self.info()
        x = np.ascontiguousarray(x, np.float32)
        tim = np.zeros((self.ndm, o.nsamples), np.float64)
        pi = self._in()
        _chk(lib().txo_dedisp_window_f64(C.byref(pi), _p(x), x.shape[0], w0, _p(tim), nthreads))
        return tim


def boxcar_f64(tim, vwn, nthreads=8):
    """Boxcar::match's S/N maximum per output entirely in FP64 (default
    statistics, AVX2 output ranges): [ndm][N] float64."""
    tim = np.ascontiguousarray(np.atleast_2d(tim), np.float64)
    ndm, N = tim.shape
    w = np.ascontiguousarray(vwn, np.int32)
    out = np.zeros((ndm, N), np.float64)
    _chk(lib().txo_boxcar_f64(_p(tim), ndm, N, _p(w), len(w), _p(out), nthreads))
    return out


def direct_f64(freqs, tsamp, dms, ddm, ndm, x, w0, N):
    f = np.ascontiguousarray(freqs, np.float64)
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros((ndm, N), np.float64)
    _chk(lib().txo_dedisp_direct_f64(len(f), _p(f), tsamp, dms, ddm, ndm, _p(x), x.shape[0], w0, N,
                                     _p(out)))
    return out


def boxcar(tim, vwn, iqr=False, variant=0, use_fma=True, nthreads=8):
    """Boxcar::match over every row: (mxS, mxwn).  variant 0 = AVX2, 1 = scalar."""
    tim = np.ascontiguousarray(np.atleast_2d(tim), np.float32)
    ndm, N = tim.shape
    w = np.ascontiguousarray(vwn, np.int32)
    mxS = np.zeros((ndm, N), np.float32)
    mxwn = np.zeros((ndm, N), np.int32)
    _chk(lib().txo_boxcar(_p(tim), ndm, N, _p(w), len(w), int(iqr), variant, int(use_fma),
                          _p(mxS), _p(mxwn), nthreads))
    return mxS, mxwn


def boxcar_row_stats(row, vwn, iqr=False, variant=0, use_fma=True):
    row = np.ascontiguousarray(row, np.float32)
    w = np.ascontiguousarray(vwn, np.int32)
    vS = np.zeros(row.size, np.float32)
    vw = np.zeros(row.size, np.int32)
    m = np.zeros(1, np.float32)
    v = np.zeros(1, np.float32)
    _chk(lib().txo_boxcar_row(_p(row), row.size, _p(w), len(w), int(iqr), variant, int(use_fma),
                              _p(vS), _p(vw), _p(m), _p(v)))
    return vS, vw, float(m[0]), float(v[0])


def boxcar_runnerup(tim, vwn, idm, isamp, use_fma=True):
    """Per point (idm[p], isamp[p]): (best, second) = the S/N Boxcar::match
    keeps there and the largest S/N of any other width there, in the
    sequential FP32 arithmetic of boxcar() (AVX2 output ranges)."""
    tim = np.ascontiguousarray(np.atleast_2d(tim), np.float32)
    ndm, N = tim.shape
    w = np.ascontiguousarray(vwn, np.int32)
    idm = np.ascontiguousarray(idm, np.int32)
    isamp = np.ascontiguousarray(isamp, np.int32)
    best = np.zeros(idm.size, np.float32)
    second = np.zeros(idm.size, np.float32)
    _chk(lib().txo_boxcar_runnerup(_p(tim), ndm, N, _p(w), len(w), int(use_fma), idm.size, _p(idm),
                                   _p(isamp), _p(best), _p(second)))
    return best, second


def dump(tim, noverlap, nlive, mean, std, nbits):
    """--savetim values of one run's window tim [ndm][N] (declared rule):
    [ndm][N - noverlap] int8 | int16 | float32."""
    tim = np.ascontiguousarray(np.atleast_2d(tim), np.float32)
    ndm, N = tim.shape
    out = np.zeros((ndm, N - noverlap), {8: np.int8, 16: np.int16, 32: np.float32}[nbits])
    lib().txo_dump(_p(tim), ndm, N, noverlap, float(nlive), mean, std, nbits, _p(out))
    return out


def threshold(mxS, mxwn, thre):
    mxS = np.ascontiguousarray(mxS, np.float32)
    mxwn = np.ascontiguousarray(mxwn, np.int32)
    ndm, N = mxS.shape
    cap = int((mxS > thre).sum())
    idm = np.zeros(cap, np.int32)
    isamp = np.zeros(cap, np.int32)
    S = np.zeros(cap, np.float32)
    wn = np.zeros(cap, np.int32)
    n = lib().txo_threshold(_p(mxS), _p(mxwn), ndm, N, thre, _p(idm), _p(isamp), _p(S), _p(wn), cap)
    assert n == cap
    return idm, isamp, S, wn


def cluster(idm, isamp, S, wn, dms, ddm, fmax, fmin, tsamp, radius, maxncand=100, minpts=0,
            drop=False, maxwn=0):
    idm = np.ascontiguousarray(idm, np.int32)
    isamp = np.ascontiguousarray(isamp, np.int32)
    S = np.ascontiguousarray(S, np.float32)
    wn = np.ascontiguousarray(wn, np.int32)
    n = idm.size
    cand = np.zeros(max(n, 1), np.int64)
    labels = np.zeros(max(n, 1), np.int32)
    nc = np.zeros(1, np.int32)
    _chk(lib().txo_cluster(n, _p(idm), _p(isamp), _p(S), _p(wn), dms, ddm, fmax, fmin, tsamp,
                           radius, maxncand, minpts, int(drop), maxwn, _p(cand), _p(nc),
                           _p(labels)))
    return cand[: nc[0]].copy(), labels[:n].copy()


def kadane(a):
    a = np.ascontiguousarray(a, np.float32)
    s = np.zeros(1, np.int64)
    e = np.zeros(1, np.int64)
    lib().txo_kadane(_p(a), a.size, _p(s), _p(e))
    return int(s[0]), int(e[0])


def cand_flfh(sub, fcnt, fsub, isamp, wn):
    sub = np.ascontiguousarray(sub, np.float32)
    nsub, N = sub.shape
    fcnt = np.ascontiguousarray(fcnt, np.float64)
    fsub = np.ascontiguousarray(fsub, np.float64)
    fl = np.zeros(1)
    fh = np.zeros(1)
    lib().txo_cand_flfh(_p(sub), _p(fcnt), _p(fsub), nsub, N, isamp, wn, _p(fl), _p(fh))
    return float(fl[0]), float(fh[0])


# ------------------------------------------------------- refinement grid
def cand_dedisperse(data, freqs, dm, tbin):
    """Candidate::dedisperse (incoherent): rows rotated by their delay."""
    d = np.array(data, np.float32, order="C", copy=True)
    f = np.ascontiguousarray(freqs, np.float64)
    lib().txo_cand_dedisperse(_p(d), d.shape[0], d.shape[1], _p(f), dm, tbin)
    return d


def hough(x, fP, dms, ddm, tbin, dm0=0.0):
    """The dedispersion tree of include/hough_transform.h; x [nrow][ncol]."""
    x = np.ascontiguousarray(x, np.float32)
    fP = np.ascontiguousarray(fP, np.float64)
    out = np.zeros_like(x)
    _chk(lib().txo_hough(_p(x), _p(out), x.shape[0], x.shape[1], _p(fP), dms, ddm, tbin, dm0))
    return out


def refine_optimize(data, freqs, tbin, dm, isdedispersed=True):
    """Candidate::optimize: (dmtmap [nrow][nbin], dms, ddm, ndm)."""
    data = np.ascontiguousarray(data, np.float32)
    f = np.ascontiguousarray(freqs, np.float64)
    nchan, nbin = data.shape
    nrow = 1 << int(np.ceil(np.log2(nchan)))
    out = np.zeros((nrow, nbin), np.float32)
    dms, ddm, ndm = np.zeros(1), np.zeros(1), np.zeros(1, np.int32)
    _chk(lib().txo_refine_optimize(_p(data), nchan, nbin, _p(f), tbin, dm, int(isdedispersed),
                                   _p(out), nrow, _p(dms), _p(ddm), _p(ndm)))
    return out, float(dms[0]), float(ddm[0]), int(ndm[0])


def mean_var(x):
    x = np.ascontiguousarray(x, np.float32)
    m, v = np.zeros(1), np.zeros(1)
    lib().txo_mean_var(_p(x), x.size, _p(m), _p(v))
    return float(m[0]), float(v[0])


def mf_ladder(nbin, snrloss):
    v = np.zeros(64, np.int32)
    n = _chk(lib().txo_mf_ladder(nbin, snrloss, _p(v), 64))
    return v[:n].copy()


def refine_mf(dmtmap, snrloss):
    """Candidate::matched_filter: (snrmap, widmap, maxdmid, maxtid, snr)."""
    m = np.ascontiguousarray(dmtmap, np.float32)
    ndm, nbin = m.shape
    S = np.zeros_like(m)
    W = np.zeros(m.shape, np.int32)
    j, i, s = np.zeros(1, np.int32), np.zeros(1, np.int64), np.zeros(1, np.float32)
    _chk(lib().txo_refine_mf(_p(m), ndm, nbin, snrloss, _p(S), _p(W), _p(j), _p(i), _p(s)))
    return S, W, int(j[0]), int(i[0]), float(s[0])


def ref_lib():
    """TransientX's own hough_transform.h built from the reference tree, or None."""
    global _ref
    if _ref is None:
        path = HERE / "_ref" / "libhough_ref.so"
        if not path.exists():
            return None
        L = C.CDLL(str(path))
        L.ref_hough.restype = None
        vp = C.c_void_p
        L.ref_hough.argtypes = [vp, vp, C.c_int, C.c_int, vp, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_int]
        _ref = L
    return _ref


class Stream:
    """Block-by-block CPU search of one ddplan row (8-bit input)."""

    def __init__(self, plan: Plan, vwn, zero_point=128, nthreads=8):
        self.plan = plan
        self._pi = plan._in()
        self._vwn = np.ascontiguousarray(vwn, np.int32)
        self.info = plan.info()
        self.h = lib().txo_stream_create(C.byref(self._pi), _p(self._vwn), self._vwn.size,
                                         zero_point, nthreads)
        if not self.h:
            raise RuntimeError("txo_stream_create failed")

    def run(self, block_u8: np.ndarray, thre: float) -> int:
        b = np.ascontiguousarray(block_u8, np.uint8)
        return _chk(lib().txo_stream_run_u8(self.h, _p(b), thre))

    def tim(self):
        n = self.plan.ndm * self.info.nsamples
        buf = (C.c_float * n).from_address(lib().txo_stream_tim(self.h))
        return np.frombuffer(buf, np.float32).reshape(self.plan.ndm, -1).copy()

    def mxS(self):
        n = self.plan.ndm * self.info.nsamples
        buf = (C.c_float * n).from_address(lib().txo_stream_mxS(self.h))
        return np.frombuffer(buf, np.float32).reshape(self.plan.ndm, -1).copy()

    def mxwn(self):
        n = self.plan.ndm * self.info.nsamples
        buf = (C.c_int * n).from_address(lib().txo_stream_mxwn(self.h))
        return np.frombuffer(buf, np.int32).reshape(self.plan.ndm, -1).copy()

    def close(self):
        if self.h:
            lib().txo_stream_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def zdot(frame):
    """Candidate::zdot of one frame [nchan][nbin] (a float32 copy is returned)."""
    x = np.array(frame, dtype=np.float32, order="C", copy=True)
    lib().txo_zdot(_p(x), x.shape[0], x.shape[1])
    return x


def kadanef(frame, td, fd, threshold, wnlimit):
    """Candidate::kadaneF of one frame [nchan][nbin] (a float32 copy is returned)."""
    x = np.array(frame, dtype=np.float32, order="C", copy=True)
    lib().txo_kadanef(_p(x), x.shape[0], x.shape[1], td, fd, threshold, wnlimit)
    return x


def kadane_wnlimit(nwidth, width, tbin, td):
    """Candidate::kadaneF's int(nwidth*width/tbin/td): nwidth*width in float
    (width is a float member), the rest in double, truncated."""
    return int(float(np.float32(np.float32(nwidth) * np.float32(width))) / tbin / td)


def frontend_u8(block, td, fd, off, gain, mask=None):
    """The declared Downsample + Equalize of one 8-bit block [nsamp][nchan]:
    td x fd neighbour sums, then (y - off) * gain in float32; masked output
    channels 0.  Returns float32 [nsamp/td][nchan/fd]."""
    x = np.ascontiguousarray(block, np.uint8)
    nsi, nchi = x.shape
    ncho = nchi // fd
    o = np.ascontiguousarray(off, np.float32)
    g = np.ascontiguousarray(gain, np.float32)
    assert o.size == ncho and g.size == ncho
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    out = np.empty((nsi // td, ncho), np.float32)
    lib().txo_frontend_u8(_p(x), nsi, nchi, td, fd, _p(o), _p(g), None if m is None else _p(m), _p(out))
    return out


def frontend_stats(block, td, fd):
    """Equalize's declared statistics of one 8-bit block: (off, gain), one
    float32 per output channel (mean, and 1/sd or 1 for a constant channel)."""
    x = np.ascontiguousarray(block, np.uint8)
    nsi, nchi = x.shape
    ncho = nchi // fd
    off = np.empty(ncho, np.float32)
    gain = np.empty(ncho, np.float32)
    lib().txo_frontend_stats(_p(x), nsi, nchi, td, fd, _p(off), _p(gain))
    return off, gain
