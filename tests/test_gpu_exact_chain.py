"""Exact-order boxcar on rows long enough for the chained statistics kernel
(k_boxcar_chain streams kXRows=16 rows per workgroup through LDS in
1056-sample chunks and keeps the running sum at every 33rd sample;
k_boxcar_xseg rebuilds each segment's window of the running sum from those
checkpoints).

The S/N and width maps must be the oracle's restatement of
src/singlepulse/boxcar.cpp:201-285 bit for bit on the device's own series:
* N = 4505: four whole chunks and a ragged one, the last segment partial,
  37 rows (the last workgroup holds 5);
* N = 4224: whole chunks only (N a multiple of 1056 and of 33), 16 rows;
* N = 44000: a row longer than one workgroup's LDS (the width loop runs
  over 4096-output segments, each rebuilding its window of the running
  sum from the checkpoints), 20 rows;
* N = 9011 at tsamp 0.1 ms with the ladder to 0.1 s (widths to 1000
  samples, as in production): windows reaching 500 samples either side
  of each segment;
* N = 88000 at tsamp 0.1 ms with the ladder to 8 s: a reach too wide for
  a segment's window, so the whole-row kernel rebuilds from the same
  checkpoints in its global scratch rows.
Also: rows recomputed for a points-only plan (tx_boxcar_rows with a DM
offset), the thresholded points, and the same search pushed through a
hipGraph pipeline giving the same session points."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

THRE = 6.0


def say(*a):
    print(*a, flush=True)


def sorted_points(p):
    return p[np.lexsort((p["isamp"], p["idm"], p["block"]))]


@pytest.mark.parametrize("ndump,ndm,nch,tsamp,maxw", [(4096, 37, 64, 1e-3, 0.05), (3840, 16, 64, 1e-3, 0.05),
                                                       (40000, 20, 16, 1e-3, 0.05), (8192, 20, 64, 1e-4, 0.1),
                                                       (80000, 8, 16, 1e-4, 8.0)])
def test_exact_order_long_rows(ndump, ndm, nch, tsamp, maxw):
    import torch

    from transientx_amd import Beam, Plan, width_ladder
    from transientx_amd.synth import Observation, Pulse, generate

    nb = 3
    obs = Observation(nchans=nch, fbot=1200.0, ftop=1500.0, tsamp=tsamp, nsamp=ndump * nb, sd=10, seed=71 + ndm,
                      impulses=1, pulses=[Pulse(dm=30.0, t0=ndump * tsamp * 1.4, fwhm=8e-3, snr=40),
                                          Pulse(dm=12.0, t0=ndump * tsamp * 2.3, fwhm=3e-3, snr=25)])
    x = torch.from_numpy(generate(obs)).cuda()
    blocks = [x[b * ndump:(b + 1) * ndump] for b in range(nb)]
    vwn = width_ladder(tsamp, tsamp, maxw, 0.1)
    kw = dict(freqs=obs.freqs, tsamp=tsamp, dms=0.0, ddm=2.0, ndm=ndm, vwn=vwn, overlap=0.1, nsub=8,
              exact_order=True)
    d0, d1 = 5, ndm - 3
    say("case exact chain ndump", ndump, "ndm", ndm, "nch", nch, "tsamp", tsamp, "widths to", int(vwn[-1]))
    beam = Beam(nch, ndump)
    checked = 0
    try:
        maps = Plan(beam, store_maps=True, **kw)
        pts_only = Plan(beam, store_maps=False, **kw)
        N = maps.info.nsamples
        say("N", N, "chunks", N / 1056.0, "segments", N / 33.0)
        for b, blk in enumerate(blocks):
            beam.push(blk)
            ready = []
            for p in (maps, pts_only):
                p.dedisp_run()
                ready.append(p.boxcar_run(THRE))
            assert ready[0] == ready[1]
            if not ready[0]:
                continue
            tim = np.stack([maps.timdata(i) for i in range(ndm)])
            rS, rW = oracle.boxcar(tim, vwn, variant=0, use_fma=True)
            S, W = maps.rows(0, ndm)
            assert np.array_equal(S.view(np.uint32), rS.view(np.uint32)), f"mxS block {b}"
            assert np.array_equal(W, rW), f"mxwn block {b}"
            S2, W2 = pts_only.rows(d0, d1)
            assert np.array_equal(S2.view(np.uint32), rS[d0:d1].view(np.uint32)), f"recomputed rows block {b}"
            assert np.array_equal(W2, rW[d0:d1])
            idm, isamp, s, wn = oracle.threshold(rS, rW, THRE)
            for p in (maps, pts_only):
                pts = p.points()
                assert np.array_equal(pts["idm"], idm) and np.array_equal(pts["isamp"], isamp)
                assert np.array_equal(pts["S"], s) and np.array_equal(pts["wn"], wn)
            checked += 1
            say("block", b, "points", idm.size)
    finally:
        beam.close()
    assert checked >= 1


def test_exact_order_graph_pipeline_equals_direct():
    import torch

    from transientx_amd import Beam, Plan, width_ladder
    from transientx_amd.synth import Observation, Pulse, generate

    nch, ndump, nb, tsamp = 64, 4096, 4, 1e-3
    obs = Observation(nchans=nch, fbot=1200.0, ftop=1500.0, tsamp=tsamp, nsamp=ndump * nb, sd=10, seed=83,
                      impulses=2, pulses=[Pulse(dm=40.0, t0=6.0, fwhm=5e-3, snr=30),
                                          Pulse(dm=90.0, t0=11.0, fwhm=9e-3, snr=25)])
    x = torch.from_numpy(generate(obs)).cuda()
    blocks = [x[b * ndump:(b + 1) * ndump] for b in range(nb)]
    vwn = width_ladder(tsamp, tsamp, 0.05, 0.1)
    rows = [(0.0, 2.0, 50, 8), (100.0, 4.0, 21, 16)]
    say("case exact chain graph")

    def run(graph):
        beam = Beam(nch, ndump)
        try:
            plans = [Plan(beam, freqs=obs.freqs, tsamp=tsamp, dms=d, ddm=dd, ndm=n, vwn=vwn, overlap=0.1, nsub=ns,
                          store_maps=False, exact_order=True) for d, dd, n, ns in rows]
            for p in plans:
                p.session_begin(1 << 16)
            if graph:
                beam.build_graph(THRE)
            for blk in blocks:
                if graph:
                    beam.graph_push(blk)
                else:
                    beam.push(blk)
                    for p in plans:
                        p.dedisp_run()
                        p.boxcar_run(THRE)
            return [sorted_points(p.session_end()) for p in plans]
        finally:
            beam.close()

    ref = run(False)
    got = run(True)
    assert sum(len(a) for a in ref) > 0
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    say("points", [len(a) for a in ref])
