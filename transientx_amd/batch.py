"""Many beams searched together.

Every beam owns its HIP stream (tx_beam), so a round that hands one block
to each live beam in turn issues independent work the device overlaps:
beams whose rows are too small to fill 256 CUs (1024-channel beams, short
ddplan rows) run beside each other instead of one after another.  Blocks
of uneven beams simply stop arriving when a beam ends.  The kernels, their
order within a beam and their arithmetic are those of the plain path, so
the candidate lists are identical to beam-by-beam runs.

With a BoxcarBatch, a round's boxcars (Boxcar::run of every row of every
beam that took a block) are one persistent launch over a device work queue
of (row, segment) items instead of one launch per row on each beam's stream.
"""
from __future__ import annotations

import threading
from dataclasses import dataclass, field

import numpy as np

import ctypes as C

from . import _lib as L
from .search import Beam, Plan


class BoxcarBatch:
    """tx_batch: Boxcar::run of many plans (of any beams on one device) as
    one persistent launch over a device work queue."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        L.check(L.lib().tx_batch_create(device, C.byref(h)))
        self.h = h

    def boxcar(self, plans, thre: float) -> np.ndarray:
        """tx_boxcar_run of every plan; per plan 1 if it ran, 0 if not ready."""
        n = len(plans)
        arr = (C.c_void_p * max(n, 1))(*[p.h for p in plans])
        ready = np.zeros(max(n, 1), np.int32)
        L.check(L.lib().tx_batch_boxcar(self.h, arr, n, thre, ready.ctypes.data_as(C.c_void_p)))
        return ready[:n]

    def close(self):
        if self.h:
            L.lib().tx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class BeamSpec:
    """One beam: channel frequencies, sampling, ddplan rows (dms, ddm, ndm,
    nsub), block length and number of blocks."""

    freqs: np.ndarray
    tsamp: float
    ndump: int
    nblocks: int
    rows: list
    vwn: np.ndarray
    overlap: float = 0.0
    zero_point: int = 128
    blocks: list = field(default_factory=list)  # device u8 tensors [ndump][nchans], one per block
    host_blocks: list = field(default_factory=list)  # the same blocks in pinned host memory (optional)
    shapes: list | None = None  # per row: launch-shape overrides (autotune.tune_row), or None


_tls = threading.local()


def _drain_buffer(nbytes: int):
    """A pinned host buffer per host thread that the thread's drains copy
    through (a drain takes one row at a time and copies its points out
    before the next), grown to the largest size asked for."""
    import torch

    buf = getattr(_tls, "pinned", None)
    if buf is None or buf.numel() < nbytes:
        buf = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
        _tls.pinned = buf
    return buf


class BeamSearch:
    """A beam and its rows, searched block by block on the beam's stream;
    thresholded points accumulate in session lists on the device."""

    def __init__(self, spec: BeamSpec, cap: int = 1 << 22, graph: bool = False):
        """cap: points a row's device list holds between drains (the whole
        run without drains)."""
        self.spec = spec
        self.graph = graph
        self._got = None
        self.beam = Beam(len(spec.freqs), spec.ndump, dtype="u8", zero_point=spec.zero_point)
        shapes = spec.shapes or [None] * len(spec.rows)
        self.plans = [Plan(self.beam, freqs=spec.freqs, tsamp=spec.tsamp, dms=dms, ddm=ddm, ndm=ndm,
                           vwn=spec.vwn, overlap=spec.overlap, nsub=nsub, store_maps=False, shapes=sh)
                      for (dms, ddm, ndm, nsub), sh in zip(spec.rows, shapes)]
        self.cap = cap
        self.next = 0

    def begin(self, thre: float = 7.0):
        self.beam.reset()
        for p in self.plans:
            p.session_begin(self.cap)
        if self.graph:
            self.beam.build_graph(thre)
        self.next = 0
        self._got = [[] for _ in self.plans]

    def drain(self):
        """Move the points the rows found since the last drain to the host
        and restart the rows' device lists (same lists, so a graph stays
        valid): device memory then bounds the points of a round, not of a
        whole beam.  Waits for the beam's queued work."""
        pin = _drain_buffer(self.cap * L.POINT_EX_DTYPE.itemsize)
        for r, p in enumerate(self.plans):
            n = p.session_end_into(pin.data_ptr(), self.cap)
            if n > self.cap:
                raise RuntimeError(f"row {r} found {n} points in one drain, more than cap {self.cap}")
            pts = np.frombuffer(pin.numpy().data, L.POINT_EX_DTYPE, count=n).copy()
            self._got[r].append(pts)
            p.session_begin(self.cap)

    def step(self, thre: float, source: str = "device", boxcar: bool = True) -> bool:
        """Search the next block; False once the beam has ended.  source:
        "device" (blocks resident in HBM), "host" (pinned host blocks copied
        before the push returns, as the drop-in's per-block run does) or
        "host_async" (pinned host blocks copied on the beam's stream).
        boxcar=False leaves the rows' Boxcar::run to the caller (a
        BoxcarBatch)."""
        if self.next >= self.spec.nblocks:
            return False
        if self.graph:
            # one graph launch: push, dedispersion and boxcar of every row
            if source == "device":
                self.beam.graph_push(self.spec.blocks[self.next])
            elif source == "host":
                self.beam.graph_push(self.spec.host_blocks[self.next])
            elif source == "host_async":
                self.beam.graph_push(self.spec.host_blocks[self.next], host_async=True)
            else:
                raise ValueError(f"unknown block source {source!r}")
            self.next += 1
            return True
        if source == "device":
            self.beam.push(self.spec.blocks[self.next])
        elif source == "host":
            self.beam.push(self.spec.host_blocks[self.next])
        elif source == "host_async":
            self.beam.push_async(self.spec.host_blocks[self.next])
        else:
            raise ValueError(f"unknown block source {source!r}")
        for p in self.plans:
            p.dedisp_run()
            if boxcar:
                p.boxcar_run(thre)
        self.next += 1
        return True

    def end(self, digest: bool = False):
        """Points of every row since begin(), drained ones included (sorted
        by block, DM, sample); digest: per row (count, points_digest) instead."""
        out = []
        for r, p in enumerate(self.plans):
            pts = p.session_end()
            if self._got and self._got[r]:
                pts = np.concatenate(self._got[r] + [pts])
                self._got[r] = []
            if digest:
                out.append(points_digest(pts))
                continue
            order = np.lexsort((pts["isamp"], pts["idm"], pts["block"]))
            out.append(pts[order])
        self._got = None
        return out

    def units(self) -> int:
        """DM-trial*samples searched: ndm*ndump per block per row."""
        return sum(p.info.ndm for p in self.plans) * self.spec.ndump * self.spec.nblocks

    def close(self):
        self.beam.close()


def run_plain(searches, thre: float, source: str = "device", drain: bool = False):
    """Beam after beam, each to its end before the next starts (as separate
    runs would be).  drain: the points of every block come to the host
    before the next block is pushed (transientx_fil clusters each block's
    candidates on the host, src/singlepulse/singlepulse.cpp:330-336)."""
    for s in searches:
        s.begin(thre)
        while s.step(thre, source):
            if drain:
                s.drain()
        s.beam.sync()


def _rounds(live, thre: float, source: str, drain: bool, drain_every: int):
    """Rounds of one block per live beam until every beam has ended."""
    while live:
        nxt = []
        for s in live:
            if drain and s.next > 0 and s.next % drain_every == 0:
                s.drain()
            if s.step(thre, source):
                nxt.append(s)
        live = nxt


def run_batched(searches, thre: float, source: str = "device", queue: BoxcarBatch | None = None,
                drain: bool = False, drain_every: int = 1, workers: int = 1):
    """Round after round, one block of every live beam per round.  With a
    queue, each round's boxcars are one persistent launch (graph mode off).
    drain: each beam's points come to the host just before its next block
    is queued, every drain_every rounds (a round of slack, so the device
    keeps the other beams' work; the session lists must hold drain_every
    rounds of points).  workers > 1: the beams are dealt to that many host
    threads, each running the rounds of its share (a beam and its plans are
    only ever driven by one thread; the library's calls on distinct beams
    are independent), so host work -- launches, drains -- overlaps itself
    as well as the device."""
    if queue is not None and any(s.graph for s in searches):
        raise ValueError("a boxcar queue needs direct launches (graph=False)")
    if queue is not None and (drain or workers > 1):
        raise ValueError("drains and worker threads drive beams one by one; the boxcar queue batches a round")
    for s in searches:
        s.begin(thre)
    if queue is None:
        if workers > 1:
            from concurrent.futures import ThreadPoolExecutor

            groups = [list(searches[i::workers]) for i in range(workers)]
            with ThreadPoolExecutor(workers) as ex:
                for f in [ex.submit(_rounds, g, thre, source, drain, drain_every) for g in groups if g]:
                    f.result()
        else:
            _rounds(list(searches), thre, source, drain, drain_every)
    else:
        live = list(searches)
        while live:
            live = [s for s in live if s.step(thre, source, boxcar=False)]
            plans = [p for s in live for p in s.plans]
            if plans:
                queue.boxcar(plans, thre)
    for s in searches:
        s.beam.sync()


def collect(searches, digest: bool = False):
    """Every beam's points per row, sorted by (block, DM, sample); digest:
    (count, points_digest) per row instead of the lists."""
    return [s.end(digest) for s in searches]


def _splitmix64(x):
    x = x + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def points_digest(pts: np.ndarray):
    """(count, digest) of a point list, independent of its order: the sum
    mod 2^64 of a splitmix64 chain over each 24-byte record's three 64-bit
    words.  Equal lists give equal digests; two runs whose lists differ give
    equal digests with probability about 2^-64."""
    n = int(len(pts))
    if n == 0:
        return (0, 0)
    w = np.ascontiguousarray(pts).view(np.uint64).reshape(n, 3)
    with np.errstate(over="ignore"):
        h = _splitmix64(w[:, 2])
        h = _splitmix64(w[:, 1] ^ h)
        h = _splitmix64(w[:, 0] ^ h)
        return (n, int(np.sum(h, dtype=np.uint64)))
