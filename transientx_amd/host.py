"""Host side of the search around the kernels, through the C ABI.

* :func:`parse_ddplan` / :func:`parse_config`: ddplan text and JSON config to
  search rows (``parse`` / ``parse_json`` of src/singlepulse/singlepulse.cpp:343-545)
* :func:`ndump`: the block length transientx_fil uses
  (src/singlepulse/dedisperse_fil.cpp:224-244)
* :func:`cluster`: Cluster<T>::run over compacted points
  (src/singlepulse/cluster.cpp:63-161, include/clustering.h:26-100)
* :func:`kadane`, :func:`cand_flfh`, :func:`cands_lines`: fl/fh and the
  ``.cands`` lines of CandPlot::plot (src/singlepulse/candplot.cpp:32-110,141-231,420-433)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib

# command-line defaults of transientx_fil (src/singlepulse/dedisperse_fil.cpp:48-106)
CLI_DEFAULTS = dict(td=1, fd=1, dms=0.0, ddm=1.0, ndm=200, overlap=0.0, nsub=0, group=0,
                    minw=1e-4, snrloss=0.1, maxw=0.02, iqr=0, thre=7.0, radius=1e-3,
                    neighbors=2, maxncand=100, minpts=5, drop=0, bswidth=0.1, bandlimit=10.0,
                    bandlimitKT=10.0, widthlimit=50e-3, threMask=10.0, threKadaneF=7.0,
                    threKadaneT=7.0, rfi="", zap="", id=0)


@dataclass
class Row:
    """One search row (the SinglePulse fields the ddplan sets)."""

    td: int
    fd: int
    dms: float
    ddm: float
    ndm: int
    overlap: float
    nsub: int
    group: int
    minw: float
    snrloss: float
    maxw: float
    iqr: bool
    thre: float
    radius: float
    neighbors: int
    maxncand: int
    minpts: int
    drop: bool
    bswidth: float
    bandlimit: float
    bandlimitKT: float
    widthlimit: float
    threMask: float
    threKadaneF: float
    threKadaneT: float
    rfi: list = field(default_factory=list)
    zap: list = field(default_factory=list)
    id: int = 0


def _to_c(d: dict) -> _lib.DdRow:
    r = _lib.DdRow()
    for name, _ in _lib.DdRow._fields_:
        v = d[name]
        if name in ("rfi", "zap"):
            if isinstance(v, (list, tuple)):
                v = ";".join(" ".join(str(x) for x in it) if isinstance(it, (list, tuple)) else str(it)
                             for it in v)
            v = v.encode()
        setattr(r, name, v)
    return r


def _from_c(r: _lib.DdRow) -> Row:
    kw = {}
    for name, _ in _lib.DdRow._fields_:
        v = getattr(r, name)
        if name == "rfi":
            v = [it.split() for it in v.decode().split(";") if it]
        elif name == "zap":
            v = [tuple(float(x) for x in it.split()) for it in v.decode().split(";") if it]
        elif name in ("iqr", "drop"):
            v = bool(v)
        kw[name] = v
    return Row(**kw)


def _parse(text: str, is_json: bool, defaults: dict | None, cap: int = 256) -> list[Row]:
    d = dict(CLI_DEFAULTS)
    if defaults:
        d.update(defaults)
    L = _lib.lib()
    rows = (_lib.DdRow * cap)()
    n = C.c_int(0)
    dd = _to_c(d)
    _lib.check(L.tx_ddplan_parse(text.encode(), int(is_json), C.byref(dd), rows, cap, C.byref(n)))
    return [_from_c(rows[i]) for i in range(n.value)]


def parse_ddplan(text: str, defaults: dict | None = None) -> list[Row]:
    """Rows of a ddplan file (``td fd dms ddm ndm snrloss maxw [rfi...]``
    per line); the other fields come from ``defaults`` (command line)."""
    return _parse(text, False, defaults)


def parse_config(text: str, defaults: dict | None = None) -> list[Row]:
    """Rows of a JSON config's ``ddplan`` array."""
    return _parse(text, True, defaults)


def ndump(seglen: float, tsamp: float, tds) -> int:
    """Samples per block: (int)(seglen/tsamp) rounded down to lcm(tds)."""
    tds = np.ascontiguousarray(tds, np.int32)
    v = _lib.lib().tx_ndump(C.c_float(seglen), tsamp, _lib.ptr(tds), tds.size)
    if v < 0:
        _lib.check(int(v))
    return int(v)


def cluster(points: np.ndarray, *, dms: float, ddm: float, fmax: float, fmin: float, tsamp: float,
            radius: float, maxncand: int = 100, minpts: int = 5, drop: bool = False,
            maxwn: int = 0, labels: bool = False):
    """Cluster::run over thresholded points (POINT_DTYPE, row-major order).
    Returns the candlist (CAND_DTYPE) and, if asked, the label per point."""
    pts = np.ascontiguousarray(points, _lib.POINT_DTYPE)
    n = pts.size
    desc = _lib.ClusterDesc(dms, ddm, fmax, fmin, tsamp, radius, maxncand, minpts, int(drop), maxwn)
    cap = max(1, min(n, max(maxncand, 0)))
    out = np.zeros(cap, _lib.CAND_DTYPE)
    lab = np.zeros(max(n, 1), np.int32)
    nc = C.c_size_t(0)
    _lib.check(_lib.lib().tx_cluster(_lib.ptr(pts), n, C.byref(desc), _lib.ptr(out), cap,
                                     C.byref(nc), _lib.ptr(lab)))
    cands = out[: nc.value].copy()
    return (cands, lab[:n].copy()) if labels else cands


def kadane(a) -> tuple[int, int]:
    a = np.ascontiguousarray(a, np.float32)
    s, e = C.c_long(0), C.c_long(0)
    _lib.lib().tx_kadane(_lib.ptr(a), a.size, C.byref(s), C.byref(e))
    return s.value, e.value


def cand_flfh(plan, cands) -> tuple[np.ndarray, np.ndarray]:
    """fl, fh (MHz, float32) of candidates at the plan's current block."""
    c = np.ascontiguousarray(cands)
    n = c.size
    idm = np.ascontiguousarray(c["idm"], np.int32)
    isamp = np.ascontiguousarray(c["isamp"], np.int32)
    wn = np.ascontiguousarray(c["wn"], np.int32)
    fl = np.zeros(max(n, 1), np.float32)
    fh = np.zeros(max(n, 1), np.float32)
    _lib.check(_lib.lib().tx_cand_flfh(plan.h, n, _lib.ptr(idm), _lib.ptr(isamp), _lib.ptr(wn),
                                       _lib.ptr(fl), _lib.ptr(fh)))
    return fl[:n], fh[:n]


class CandCounter:
    """CandPlot::num_cand: the running candidate number shared by all rows."""

    def __init__(self, start: int = 0):
        self.value = C.c_long(start)


def cands_lines(plan, cands, *, tstart: float, rootname: str, beam: str, plan_id: int,
                fname: str, counter: CandCounter | None = None) -> str:
    """The ``.cands`` lines CandPlot::plot appends for the plan's current block."""
    c = np.ascontiguousarray(cands, _lib.CAND_DTYPE)
    counter = counter or CandCounter()
    ctx = _lib.CandsCtx(tstart, rootname.encode(), beam.encode(), plan_id, fname.encode(),
                        C.pointer(counter.value))
    cap = 8192 + 1024 * max(1, c.size)
    buf = C.create_string_buffer(cap)
    n = C.c_size_t(0)
    _lib.check(_lib.lib().tx_cands_format(plan.h, _lib.ptr(c), c.size, C.byref(ctx), buf, cap,
                                          C.byref(n)))
    return buf.value.decode()


def cands_filename(rootname: str, tstart: float, beam: str) -> str:
    """rootname_<tstart, 10 decimals>_<beam>.cands (src/singlepulse/candplot.cpp:79-81,425)."""
    return f"{rootname}_{tstart:.10f}_{beam}.cands"


def beam_name(ibeam: int, incoherent: bool = False) -> str:
    """obsinfo["Beam"]: cfbf/ifbf + 5-digit beam number (src/singlepulse/singlepulse.cpp:271-277)."""
    return f"{'ifbf' if incoherent else 'cfbf'}{ibeam:05d}"
