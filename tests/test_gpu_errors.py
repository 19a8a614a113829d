"""Error paths and degenerate data on the device path.

* Plans the declared rule refuses (no DMs, a sub-band count that does not
  divide the channels, a negative DM step, a non-positive tsamp or channel
  frequency), a plan made after the first push, a block of the wrong
  length, DM rows out of range: each is an error code, and the beam still
  searches afterwards.
* Rows the boxcar must leave empty (src/singlepulse/boxcar.cpp:242 gates the
  width loop on var > 0): a constant series (var = 0), and series holding a
  NaN or an infinity (var is NaN).  An f32 beam with one NaN and one +Inf
  sample contaminates some DM rows of a window and not others; the maps are
  bit-identical to the oracle's in exact order, and in the default mode the
  contaminated rows are empty and the clean rows meet the tuple contract."""
import numpy as np
import pytest

import oracle

from test_gpu_parity import assert_tuple_contract

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_refused_plans_and_blocks_leave_the_beam_usable():
    import torch

    from transientx_amd import Beam, Plan, width_ladder
    from transientx_amd._lib import TxError
    from transientx_amd.synth import Observation, Pulse, generate

    nch, ndump = 64, 256
    obs = Observation(nchans=nch, fbot=1200.0, ftop=1500.0, tsamp=1e-3, nsamp=ndump * 6, sd=10, seed=51,
                      pulses=[Pulse(dm=40.0, t0=0.9, fwhm=5e-3, snr=40)])
    x = torch.from_numpy(generate(obs)).cuda()
    vwn = width_ladder(1e-3, 1e-3, 0.03, 0.1)
    good = dict(freqs=obs.freqs, tsamp=1e-3, dms=0.0, ddm=2.0, ndm=40, vwn=vwn, overlap=0.1, nsub=8)
    bad = [dict(good, ndm=0), dict(good, nsub=7), dict(good, ddm=-1.0), dict(good, tsamp=0.0),
           dict(good, freqs=np.linspace(300.0, -10.0, nch))]
    say("case refused plans")
    beam = Beam(nch, ndump)
    try:
        for kw in bad:
            with pytest.raises(TxError):
                Plan(beam, **kw)
        plan = Plan(beam, **good)
        with pytest.raises(TxError):
            beam.push(x[: ndump - 1])  # a block must hold ndump samples
        pts = 0
        for b in range(6):
            beam.push(x[b * ndump:(b + 1) * ndump])
            plan.dedisp_run()
            if plan.boxcar_run(7.0):
                pts += len(plan.points())
        assert pts > 0
        with pytest.raises(TxError):
            Plan(beam, **good)  # plans come before the first push
        with pytest.raises(TxError):
            plan.timdata(40)  # DM rows 0..39
        with pytest.raises(TxError):
            plan.rows(30, 41)
        beam.sync()
    finally:
        beam.close()


@pytest.mark.parametrize("case", ["constant", "nan_inf"])
def test_degenerate_rows_through_the_boxcar(case):
    import torch

    from transientx_amd import Beam, Plan, width_ladder
    from transientx_amd.synth import Observation, Pulse, generate

    nch, ndump, nb = 64, 512, 6
    tsamp = 1e-3
    obs = Observation(nchans=nch, fbot=1200.0, ftop=1500.0, tsamp=tsamp, nsamp=ndump * nb, sd=10, seed=52,
                      pulses=[Pulse(dm=60.0, t0=1.9, fwhm=6e-3, snr=40)])
    if case == "constant":
        x = np.full((ndump * nb, nch), 3.0, np.float32)  # integer sums: every series exactly constant
    else:
        x = generate(obs).astype(np.float32) - 128.0
        x[1700, 5] = np.nan
        x[2300, 40] = np.inf
    vwn = width_ladder(tsamp, tsamp, 0.05, 0.1)
    kw = dict(freqs=obs.freqs, tsamp=tsamp, dms=0.0, ddm=4.0, ndm=60, vwn=vwn, overlap=0.1, nsub=8)
    say("case degenerate rows", case)
    results = {}
    for exact in (True, False):
        beam = Beam(nch, ndump, dtype="f32")
        try:
            plan = Plan(beam, store_maps=True, exact_order=exact, **kw)
            got = []
            for b in range(nb):
                beam.push(torch.from_numpy(x[b * ndump:(b + 1) * ndump]).cuda())
                plan.dedisp_run()
                if plan.boxcar_run(7.0):
                    o = plan.info
                    tim = np.stack([plan.timdata(i) for i in range(o.ndm)])
                    S, W = plan.rows(0, o.ndm)
                    got.append((tim, S, W))
            results[exact] = got
        finally:
            beam.close()
    assert len(results[True]) == len(results[False]) > 0
    nbad_rows = 0
    for (tim, S, W), (tim2, S2, W2) in zip(results[True], results[False]):
        assert np.array_equal(tim.view(np.uint32), tim2.view(np.uint32))
        rS, rW = oracle.boxcar(tim, vwn)
        # exact order: the reference's maps bit for bit, NaN rows included
        assert np.array_equal(S.view(np.uint32), rS.view(np.uint32))
        assert np.array_equal(W, rW)
        if case == "constant":
            # windows that reach before the stream start hold zeros there;
            # every later window is constant: var = 0, empty rows
            bad = (tim == tim[:, :1]).all(axis=1)
        else:
            bad = ~np.isfinite(tim).all(axis=1)
        nbad_rows += int(bad.sum())
        # constant rows and rows holding a NaN or an infinity are empty in both modes
        assert not rS[bad].any() and not rW[bad].any()
        assert not S2[bad].any() and not W2[bad].any()
        good = ~bad
        if good.any():
            assert_tuple_contract(S2[good], W2[good], rS[good], rW[good], tim[good], vwn, tag=f"{case} clean rows")
    total = sum(t.shape[0] for t, _, _ in results[True])
    say("empty rows", nbad_rows, "of", total)
    if case == "nan_inf":
        assert 0 < nbad_rows < total  # some rows of a window are hit, others not
    else:
        assert nbad_rows > total // 2
