"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars:
  * dedispersed series and sub-band series: bit-exact (8-bit input sums are
    exact integers in FP32; f32 input is summed in the oracle's order);
  * S/N maps in exact_order mode: bit-identical to the oracle's restatement
    of the AVX2 Boxcar::match (src/singlepulse/boxcar.cpp:180-296);
  * S/N maps in the default (parallel FP64 statistics) mode: |dS| <= 2e-3
    relative, and thresholded tuples (idm, isamp, wn) identical away from a
    0.1 sigma margin around the threshold: the set of points, and the width
    of every point except where the oracle's best width beats every other
    width by less than TIE_EPS (a width near-tie, counted and reported).
"""
import numpy as np
import pytest

import oracle
from transientx_amd import Beam, Plan, width_ladder
from transientx_amd.synth import Observation, Pulse, channel_freqs, generate

pytestmark = pytest.mark.gpu

THRE = 7.0
# Width near-tie margin (S/N units): the default-mode boxcar sums in FP64
# from per-segment running sums, the reference in sequential FP32 over the
# whole row, so one width's S/N may differ by ~1e-4 between them; a point
# may take another width only where the oracle's best width beats every
# other width by less than 2x that.  TIE_EPS leaves a 10x allowance.
TIE_EPS = 2e-3
MARGIN = 0.1


def say(*a):
    print(*a, flush=True)


def assert_tuple_contract(S, W, rS, rW, tim, vwn, thre=THRE, tag=""):
    """Thresholded tuples (idm, isamp, wn) of the device maps (S, W) against
    the oracle's (rS, rW) over the same series tim: identical points
    outside |rS - thre| < MARGIN, identical widths at every such point
    except oracle width near-ties (best - second < TIE_EPS).  Returns
    (points, near-tie points, width mismatches)."""
    near = np.abs(rS - thre) < MARGIN
    g = (S > thre) & ~near
    r = (rS > thre) & ~near
    assert np.array_equal(g, r), f"{tag}: thresholded points differ outside the margin"
    jj, ii = np.nonzero(r)
    best, second = oracle.boxcar_runnerup(tim, vwn, jj, ii)
    assert np.array_equal(best.view(np.uint32), rS[jj, ii].view(np.uint32))
    tie = (best - second) < TIE_EPS
    bad = W[jj, ii] != rW[jj, ii]
    assert not (bad & ~tie).any(), \
        f"{tag}: {int((bad & ~tie).sum())} widths differ at points with a clear best width"
    say(tag, "points", jj.size, "width near-ties", int(tie.sum()), "width mismatches", int(bad.sum()),
        "eps", TIE_EPS, "max |dS|", float(np.abs(S[jj, ii] - rS[jj, ii]).max()) if jj.size else 0.0)
    return jj.size, int(tie.sum()), int(bad.sum())


def stream(obs_x, nchans, ndump, nblocks, plan_kw, beam_kw=None, thre=THRE, check=None):
    """Push nblocks blocks; after each run call check(block, plan)."""
    beam = Beam(nchans, ndump, **(beam_kw or {}))
    plan = Plan(beam, **plan_kw)
    out = []
    try:
        for b in range(1, nblocks + 1):
            blk = obs_x[(b - 1) * ndump:b * ndump]
            beam.push(blk)
            plan.dedisp_run()
            ready = plan.boxcar_run(thre)
            out.append(check(b, plan, ready) if check else ready)
    finally:
        beam.close()
    return out


def all_tim(plan):
    o = plan.info
    return np.stack([plan.timdata(i) for i in range(o.ndm)])


def small_case(seed=1, C=64, ndump=512, nblocks=6, dtype="u8"):
    obs = Observation(nchans=C, fbot=1200, ftop=1500, tsamp=1e-3, nsamp=ndump * nblocks,
                      sd=10, seed=seed,
                      pulses=[Pulse(dm=37.0, t0=0.9, fwhm=6e-3, snr=30),
                              Pulse(dm=61.0, t0=2.1, fwhm=2e-3, snr=15)])
    x = generate(obs)
    vwn = width_ladder(1e-3, 1e-3, 0.05, 0.1)
    plan_kw = dict(freqs=obs.freqs, tsamp=obs.tsamp, dms=0.0, ddm=2.0, ndm=40, vwn=vwn,
                   overlap=0.25, nsub=8)
    return obs, x, vwn, plan_kw


def oracle_plan(obs, ndump, plan_kw):
    return oracle.Plan(freqs=obs.freqs, tsamp=obs.tsamp, ndump=ndump, dms=plan_kw["dms"],
                       ddm=plan_kw["ddm"], ndm=plan_kw["ndm"], overlap=plan_kw.get("overlap", 0),
                       nsub=plan_kw.get("nsub", 0), group=plan_kw.get("group", 0))


def test_delay_tables_identical():
    obs, x, vwn, kw = small_case()
    beam = Beam(obs.nchans, 512)
    try:
        plan = Plan(beam, **kw)
        t = plan.tables()
        d1, d2, dmc, fref = oracle_plan(obs, 512, kw).tables()
        assert np.array_equal(t["d1"], d1) and np.array_equal(t["d2"], d2)
        assert np.array_equal(t["dmc"], dmc) and np.array_equal(t["fref"], fref)
        o = plan.info
        oi = oracle_plan(obs, 512, kw).info()
        assert (o.offset, o.noverlap, o.nsamples, o.group, o.ndmc) == \
            (oi.offset, oi.noverlap, oi.nsamples, oi.group, oi.ndmc)
    finally:
        beam.close()


@pytest.mark.parametrize("dtype", ["u8", "f32"])
def test_series_bitexact_and_exact_order_maps(dtype):
    ndump, nblocks = 512, 6
    obs, x, vwn, kw = small_case(dtype=dtype)
    if dtype == "f32":
        rng = np.random.default_rng(3)
        xin = (x.astype(np.float32) - 128) * np.float32(0.37) + \
            rng.normal(0, 0.01, x.shape).astype(np.float32)
        xo = xin
    else:
        xin = x
        xo = x.astype(np.float32) - 128
    op = oracle_plan(obs, ndump, kw)
    say("case", dtype, "C", obs.nchans, "ndump", ndump)

    def check(b, plan, ready):
        tim = all_tim(plan)
        w0 = op.window_start(b)
        ref = op.dedisperse(xo, w0)
        assert np.array_equal(tim, ref), f"block {b}"
        if not ready:
            return False
        S, W = plan.rows(0, plan.info.ndm)
        rS, rW = oracle.boxcar(ref, vwn, variant=0, use_fma=True)
        assert np.array_equal(S.view(np.uint32), rS.view(np.uint32)), f"mxS block {b}"
        assert np.array_equal(W, rW), f"mxwn block {b}"
        pts = plan.points()
        idm, isamp, s, wn = oracle.threshold(rS, rW, THRE)
        assert np.array_equal(pts["idm"], idm) and np.array_equal(pts["isamp"], isamp)
        assert np.array_equal(pts["S"], s) and np.array_equal(pts["wn"], wn)
        return True

    flags = stream(xin, obs.nchans, ndump, nblocks, dict(kw, exact_order=True),
                   dict(dtype=dtype), check=check)
    assert flags[-1] and not flags[0]


def test_fast_mode_tuples_match_away_from_margin():
    ndump, nblocks = 512, 6
    obs, x, vwn, kw = small_case(seed=4)
    xo = x.astype(np.float32) - 128
    op = oracle_plan(obs, ndump, kw)

    def check(b, plan, ready):
        if not ready:
            return 0
        tim = all_tim(plan)
        S, W = plan.rows(0, plan.info.ndm)
        rS, rW = oracle.boxcar(tim, vwn, variant=0, use_fma=True)
        np.testing.assert_allclose(S, rS, rtol=2e-3, atol=2e-3)
        assert_tuple_contract(S, W, rS, rW, tim, vwn, tag=f"fast block {b}")
        pts = plan.points()
        sel = (S > THRE)
        assert pts.size == int(sel.sum())
        jj, ii = np.nonzero(sel)
        assert np.array_equal(pts["idm"], jj) and np.array_equal(pts["isamp"], ii)
        assert np.array_equal(pts["S"], S[jj, ii]) and np.array_equal(pts["wn"], W[jj, ii])
        return int(sel.sum())

    counts = stream(x, obs.nchans, ndump, nblocks, kw, check=check)
    assert sum(counts) > 0


def test_masked_channels_and_subband_products():
    ndump, nblocks = 512, 5
    obs, x, vwn, kw = small_case(seed=6)
    mask = np.ones(obs.nchans, np.uint8)
    mask[[3, 4, 30, 63]] = 0
    xo = x.astype(np.float32) - 128
    xo[:, mask == 0] = 0
    op = oracle_plan(obs, ndump, kw)

    def check(b, plan, ready):
        t = plan.tables()
        assert t["fcnt"].sum() == mask.sum()
        w0 = op.window_start(b)
        ref, sub = op.dedisperse(xo, w0, with_sub=True)
        assert np.array_equal(all_tim(plan), ref)
        for i in (0, 17, 39):
            assert np.array_equal(plan.subdata(i), sub[i]), (b, i)
        idm = np.array([5, 17, 39, 0], np.int32)
        isamp = np.array([100, 0, 639, 320], np.int32)
        wn = np.array([3, 7, 20, 1], np.int32)
        spec = plan.cand_spectra(idm, isamp, wn)
        for c in range(idm.size):
            lo, hi = max(0, isamp[c] - wn[c] + 1), min(sub.shape[2] - 1, isamp[c] + wn[c] - 1)
            expect = np.zeros(sub.shape[1], np.float32)
            for a in range(lo, hi + 1):
                expect += sub[idm[c], :, a] / t["fcnt"].astype(np.float32)
            assert np.array_equal(spec[c], expect), (b, c)
        return ready

    stream(x, obs.nchans, ndump, nblocks, kw, dict(chan_mask=mask), check=check)


def test_iqr_mode_bitexact():
    ndump, nblocks = 512, 5
    obs, x, vwn, kw = small_case(seed=8)

    def check(b, plan, ready):
        if not ready:
            return False
        tim = all_tim(plan)
        S, W = plan.rows(0, plan.info.ndm)
        rS, rW = oracle.boxcar(tim, vwn, iqr=True, variant=0)
        assert np.array_equal(S.view(np.uint32), rS.view(np.uint32))
        assert np.array_equal(W, rW)
        return True

    flags = stream(x, obs.nchans, ndump, nblocks, dict(kw, iqr=True, exact_order=True),
                   check=check)
    assert flags[-1]


@pytest.mark.parametrize("C,nsub,ndm,ddm,fbot,ftop,tsamp,ndump,group", [
    (256, 4, 120, 3.0, 1200, 1500, 5e-4, 2048, 0),     # 64 channels per sub-band
    (512, 128, 200, 5.0, 856, 1712, 1e-3, 2048, 0),    # wide groups, long delay spreads
    (384, 12, 90, 1.0, 1000, 1500, 2e-4, 1500, 7),     # odd sizes, explicit group
    (128, 128, 64, 2.0, 1200, 1500, 1e-3, 700, 0),     # one channel per sub-band
])
def test_series_bitexact_across_plan_shapes(C, nsub, ndm, ddm, fbot, ftop, tsamp, ndump, group):
    rng = np.random.default_rng(C + ndm)
    freqs = channel_freqs(C, fbot, ftop)
    op = oracle.Plan(freqs=freqs, tsamp=tsamp, ndump=ndump, dms=10.0, ddm=ddm, ndm=ndm,
                     overlap=0.1, nsub=nsub, group=group)
    info = op.info()
    nblocks = max(3, -(-(info.maxdelay + 2 * ndump) // ndump))
    x = rng.integers(0, 256, size=(nblocks * ndump, C), dtype=np.uint8)
    xo = x.astype(np.float32) - 128
    vwn = np.array([1, 2, 4, 8], np.int32)
    kw = dict(freqs=freqs, tsamp=tsamp, dms=10.0, ddm=ddm, ndm=ndm, vwn=vwn, overlap=0.1,
              nsub=nsub, group=group)
    say("case shapes", C, nsub, ndm, ddm, "blocks", nblocks)

    def check(b, plan, ready):
        if b in (1, nblocks // 2, nblocks):
            ref = op.dedisperse(xo, op.window_start(b))
            assert np.array_equal(all_tim(plan), ref), b
        return ready

    stream(x, C, ndump, nblocks, kw, check=check)


@pytest.mark.parametrize("ndump,maxw", [(7000, 5.0), (3000, 0.4), (2048, 1.5)])
def test_streaming_boxcar_long_boxes(ndump, maxw):
    """Default-mode boxcar (chunked running sum) against the oracle when the
    widest box reaches further than one chunk of 2048 samples."""
    C, tsamp = 64, 1e-3
    obs = Observation(nchans=C, fbot=1200, ftop=1500, tsamp=tsamp, nsamp=ndump * 4, sd=10,
                      seed=ndump, pulses=[Pulse(dm=20.0, t0=ndump * tsamp * 2.3, fwhm=0.3,
                                                snr=40)])
    x = generate(obs)
    vwn = width_ladder(tsamp, 1e-3, maxw, 0.1)
    kw = dict(freqs=obs.freqs, tsamp=tsamp, dms=0.0, ddm=1.0, ndm=24, vwn=vwn, overlap=0.1,
              nsub=8)
    say("case streaming boxcar ndump", ndump, "widest", int(vwn[-1]))

    def check(b, plan, ready):
        if not ready:
            return 0
        tim = all_tim(plan)
        S, W = plan.rows(0, plan.info.ndm)
        rS, rW = oracle.boxcar(tim, vwn)
        np.testing.assert_allclose(S, rS, rtol=2e-3, atol=2e-3)
        assert_tuple_contract(S, W, rS, rW, tim, vwn, tag=f"long boxes block {b}")
        return int((S > THRE).sum())

    counts = stream(x, C, ndump, 4, kw, check=check)
    assert sum(counts) > 0


def test_not_ready_until_offset_is_covered():
    ndump = 256
    obs, x, vwn, kw = small_case(ndump=ndump, nblocks=10)
    op = oracle_plan(obs, ndump, kw).info()
    flags = stream(x, obs.nchans, ndump, 10, kw)
    for b, f in enumerate(flags, start=1):
        assert f == (b * ndump >= op.offset - op.noverlap + ndump), b


def test_plain_1024ch_config_blocks():
    """1024 ch, 1200-1500 MHz, 98.304 us, DM 0-1000 step 1, nsub 32, ladder to 0.1 s."""
    C, tsamp = 1024, 98.304e-6
    ndump = int(2.0 / tsamp)
    nblocks = 3
    obs = Observation(nchans=C, fbot=1200, ftop=1500, tsamp=tsamp, nsamp=ndump * nblocks,
                      sd=16, seed=1,
                      pulses=[Pulse(dm=250.0, t0=2.5, fwhm=5e-3, snr=25),
                              Pulse(dm=420.0, t0=3.6, fwhm=2e-3, snr=15)])
    x = generate(obs)
    vwn = width_ladder(tsamp, 1e-4, 0.1, 0.1)
    assert len(vwn) == 32
    kw = dict(freqs=obs.freqs, tsamp=tsamp, dms=0.0, ddm=1.0, ndm=1000, vwn=vwn, nsub=32)
    op = oracle_plan(obs, ndump, kw)
    xo = x.astype(np.float32) - 128
    say("case plain 1024ch ndm 1000 ndump", ndump)

    def check(b, plan, ready):
        tim = all_tim(plan)
        ref = op.dedisperse(xo, op.window_start(b))
        assert np.array_equal(tim, ref), b
        if not ready:
            return None
        S, W = plan.rows(0, 1000)
        rS, rW = oracle.boxcar(ref, vwn)
        assert_tuple_contract(S, W, rS, rW, ref, vwn, tag=f"1024ch block {b}")
        np.testing.assert_allclose(S, rS, rtol=2e-3, atol=2e-3)
        i, t = np.unravel_index(np.argmax(S), S.shape)
        return i, t

    res = stream(x, C, ndump, nblocks, kw, check=check)
    peaks = [r for r in res if r is not None]
    assert peaks and abs(peaks[0][0] - 250) <= 3


@pytest.mark.parametrize("shapes", [{"s1_ncc": 4, "s1_q": 1}, {"s1_ncc": 16, "s1_q": 2},
                                    {"s2_shape": 1, "s2_lds": 40960}, {"s2_shape": 6},
                                    {"s2_shape": 3, "s2_lds": 16384}])
def test_launch_shapes_do_not_change_results(shapes):
    """Autotuner overrides pick other tiles; the series stay bit-exact."""
    ndump, nblocks = 512, 5
    obs, x, vwn, kw = small_case()
    op = oracle_plan(obs, ndump, kw)
    xo = x.astype(np.float32) - 128
    say("shapes", shapes)

    def check(b, plan, ready):
        assert np.array_equal(all_tim(plan), op.dedisperse(xo, op.window_start(b))), f"block {b}"
        return ready

    stream(x, obs.nchans, ndump, nblocks, dict(kw, shapes=shapes), check=check)


def test_series_bitexact_with_wide_stage2_spans():
    """High time resolution at high DM (16 us, 1000-1500 MHz, DM 3000-5000
    step 10): one DM step moves the bottom sub-band by ~1440 samples, so
    eight DMs span more than the default 80 KB of LDS and stage 2 takes the
    whole LDS of the workgroup.  Series bit-exact against the oracle for the
    first two ready blocks."""
    C, tsamp, ndump = 256, 16e-6, 8192
    kw0 = dict(tsamp=tsamp, dms=3000.0, ddm=10.0, ndm=200, nsub=16)
    freqs = Observation(nchans=C, fbot=1000, ftop=1500, tsamp=tsamp, nsamp=1).freqs
    op = oracle.Plan(freqs=freqs, tsamp=tsamp, ndump=ndump, dms=3000.0, ddm=10.0, ndm=200, nsub=16)
    ready_at = -(-(op.info().offset + ndump) // ndump)  # first block whose boxcar runs
    nblocks = ready_at + 1
    obs = Observation(nchans=C, fbot=1000, ftop=1500, tsamp=tsamp, nsamp=ndump * nblocks, sd=12, seed=9,
                      pulses=[Pulse(dm=4100.0, t0=0.05, fwhm=2e-3, snr=40)])
    x = generate(obs)
    vwn = width_ladder(tsamp, 16e-6, 5e-3, 0.1)
    kw = dict(kw0, freqs=freqs, vwn=vwn)
    xo = x.astype(np.float32) - 128
    say("case wide spans: offset", op.info().offset, "blocks", nblocks)

    def check(b, plan, ready):
        if b < ready_at - 1:
            return ready
        assert np.array_equal(all_tim(plan), op.dedisperse(xo, op.window_start(b))), b
        return ready

    res = stream(x, C, ndump, nblocks, kw, check=check)
    assert res[-1]


@pytest.mark.parametrize("dtype", ["u8", "f32"])
def test_boxcar_register_window_bit_identical(dtype):
    """The segmented boxcar's register window for the ladder's narrow widths
    (1-9, 11, 13, 16; s3_window 0, the default) gives the maps, widths and
    points of the all-LDS path (s3_window 1), bit for bit: rows long enough
    for whole segments inside every width's output range, pulses and ties
    at zero included."""
    ndump, nblocks, C = 8192, 4, 64
    obs = Observation(nchans=C, fbot=1200, ftop=1500, tsamp=1e-3, nsamp=ndump * nblocks, sd=10, seed=21,
                      pulses=[Pulse(dm=30.0, t0=9.1, fwhm=3e-3, snr=25), Pulse(dm=55.0, t0=20.3, fwhm=12e-3,
                                                                             snr=40)])
    x = generate(obs)
    if dtype == "f32":
        x = x.astype(np.float32) - 128.0
    vwn = width_ladder(1e-3, 1e-3, 0.05, 0.1)
    assert list(vwn[:12]) == [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 13, 16]
    kw = dict(freqs=obs.freqs, tsamp=obs.tsamp, dms=0.0, ddm=2.0, ndm=40, vwn=vwn, overlap=0.1, nsub=8)
    say("case window", dtype)

    def run(window_off):
        maps = []

        def check(b, plan, ready):
            if ready:
                S, W = plan.rows(0, plan.info.ndm)
                maps.append((S.copy(), W.copy(), plan.points()))
            return ready

        stream(x, C, ndump, nblocks, dict(kw, shapes={"s3_window": window_off}),
               beam_kw=dict(dtype=dtype), check=check)
        return maps

    a, b = run(0), run(1)
    assert len(a) == len(b) and len(a) >= 2
    hits = 0
    for (S1, W1, p1), (S2, W2, p2) in zip(a, b):
        assert np.array_equal(S1.view(np.uint32), S2.view(np.uint32))
        assert np.array_equal(W1, W2)
        assert np.array_equal(p1, p2)
        hits += len(p1)
    assert hits > 0


@pytest.mark.parametrize("mode", ["u8", "frontend", "frontend_in_load"])
def test_production_4096ch_rows_first_ready_block(mode):
    """The production beam at full size: 4096 ch, 856-1712 MHz, 76.56 us,
    2 s blocks with overlap 0.1, the three ddplan rows 0-500/0.5/64sub,
    500-1500/2/64sub, 1500-3000/5/128sub and the ladder to 0.1 s, with the
    production recipe's narrowband RFI channels and zero-DM impulses.  Each
    row's series are bit-exact at the block where its boxcar gate opens
    (up to 8 blocks in) and at the last block, and its S/N tuples match
    away from the 0.1 sigma margin; each row's first window holds a pulse.  mode "frontend": raw
    8-bit blocks through the device front end (equalized with the first
    block's statistics) into the f32 search, against the oracle's front
    end and f32 dedispersion; "frontend_in_load": the same with the front
    end applied in stage 1's load over an 8-bit ring."""
    import os

    import torch

    from transientx_amd.synth import gpu_generate

    C, tsamp = 4096, 76.56e-6
    ndump = int(2.0 / tsamp)
    rows = [(0.0, 0.5, 1000, 64), (500.0, 2.0, 500, 64), (1500.0, 5.0, 300, 128)]
    nblocks = 9
    # the production recipe: 0.5% of the channels with narrowband RFI at 5x,
    # zero-DM impulses, dispersed pulses in each row's first window
    obs = Observation(nchans=C, fbot=856.0, ftop=1712.0, tsamp=tsamp, nsamp=ndump * nblocks, mean=128.0,
                      sd=12.0, seed=8, rfi_frac=0.005, rfi_amp=5.0, impulses=3,
                      pulses=[Pulse(dm=230.0, t0=3.0, fwhm=3e-3, snr=30),
                              Pulse(dm=900.0, t0=2.5, fwhm=8e-3, snr=30),
                              Pulse(dm=2100.0, t0=1.2, fwhm=20e-3, snr=40)])
    vwn = width_ladder(tsamp, 1e-4, 0.1, 0.1)
    say("case production 4096ch", mode, "rows", rows, "ndump", ndump, "blocks", nblocks)
    data, _ = gpu_generate(obs, ndump=ndump)
    if mode.startswith("frontend"):
        off, gain = oracle.frontend_stats(data[:ndump].cpu().numpy(), 1, 1)
        beam = Beam(C, ndump, dtype="f32")
        beam.set_frontend(C, 1, 1, off, gain, in_load=mode == "frontend_in_load")
    else:
        beam = Beam(C, ndump)
    kws = [dict(freqs=obs.freqs, tsamp=tsamp, dms=d, ddm=dd, ndm=n, vwn=vwn, overlap=0.1, nsub=ns)
           for d, dd, n, ns in rows]
    plans = [Plan(beam, **kw) for kw in kws]
    got = []  # (row, block, series, maps, first ready block?)
    first = [None] * len(rows)
    try:
        for b in range(1, nblocks + 1):
            beam.push(data[(b - 1) * ndump:b * ndump])
            for r, p in enumerate(plans):
                p.dedisp_run()
                ready = p.boxcar_run(THRE)
                if ready and (first[r] is None or b == nblocks):
                    got.append((r, b, all_tim(p), p.rows(0, p.info.ndm), first[r] is None))
                    if first[r] is None:
                        first[r] = b
                        say("row", r, "ready at block", b)
    finally:
        beam.close()
    assert all(f is not None for f in first)
    x = data.cpu().numpy()
    del data
    torch.cuda.empty_cache()
    nth = min(16, os.cpu_count() or 1)
    for r, b, tim, maps, is_first in got:
        op = oracle_plan(obs, ndump, kws[r])
        o = op.info()
        w0 = op.window_start(b)
        a = max(0, w0)
        if mode.startswith("frontend"):
            xo = oracle.frontend_u8(x[a:w0 + o.nsamples + o.maxdelay], 1, 1, off, gain)
        else:
            xo = x[a:w0 + o.nsamples + o.maxdelay].astype(np.float32) - 128.0
        ref = op.dedisperse(xo, w0 - a, nthreads=nth)
        del xo
        assert np.array_equal(tim.view(np.uint32), ref.view(np.uint32)), r
        S, W = maps
        rS, rW = oracle.boxcar(ref, vwn, nthreads=nth)
        assert_tuple_contract(S, W, rS, rW, ref, vwn, tag=f"production {mode} row {r} block {b}")
        np.testing.assert_allclose(S, rS, rtol=2e-3, atol=2e-3)
        npts = int(np.count_nonzero(S > THRE))
        i, t = np.unravel_index(np.argmax(S), S.shape)
        say("row", r, "block", b, "points", npts, "peak DM", kws[r]["dms"] + i * kws[r]["ddm"], "S",
            float(S[i, t]))
        if is_first:
            assert npts > 0, r  # the row's pulse


@pytest.mark.parametrize("dtype,thre", [("u8", 1.5), ("f32", 1.5), ("u8", 7.0)])
def test_points_only_boxcar_equals_map_path(dtype, thre):
    """Without stored maps the segmented boxcar takes the maximum over the
    widths first and recovers the width only where it emits a point (the
    last width whose S/N equals the maximum).  Its points -- sample, S/N
    and width -- are bit-identical to those of the path that stores the
    maps, on two plans of one beam; a low threshold makes most outputs
    points, and long boxes put segment edges inside every width's reach.
    The points-only plan's rows, recomputed on demand, equal the stored
    maps."""
    ndump, nblocks, C = 8192, 4, 64
    obs = Observation(nchans=C, fbot=1200, ftop=1500, tsamp=1e-3, nsamp=ndump * nblocks, sd=10, seed=23,
                      pulses=[Pulse(dm=30.0, t0=9.1, fwhm=3e-3, snr=25), Pulse(dm=55.0, t0=20.3, fwhm=40e-3,
                                                                             snr=40)])
    x = generate(obs)
    if dtype == "f32":
        x = x.astype(np.float32) - 128.0
    vwn = width_ladder(1e-3, 1e-3, 0.3, 0.1)
    kw = dict(freqs=obs.freqs, tsamp=obs.tsamp, dms=0.0, ddm=2.0, ndm=40, vwn=vwn, overlap=0.1, nsub=8,
              max_points=1 << 22)
    say("case points-only", dtype, "thre", thre, "widths", len(vwn))
    beam = Beam(C, ndump, dtype=dtype)
    hits = 0
    try:
        pa = Plan(beam, store_maps=True, **kw)
        pb = Plan(beam, store_maps=False, **kw)
        for b in range(1, nblocks + 1):
            beam.push(x[(b - 1) * ndump:b * ndump])
            for p in (pa, pb):
                p.dedisp_run()
            ra, rb = pa.boxcar_run(thre), pb.boxcar_run(thre)
            assert ra == rb
            if not ra:
                continue
            a, c = pa.points(), pb.points()
            assert a.size == c.size, f"block {b}"
            assert np.array_equal(a, c), f"block {b}"
            S, W = pa.rows(0, pa.info.ndm)
            sel = S > thre
            assert a.size == int(sel.sum())
            # rows of the points-only plan, recomputed on demand: the stored maps
            for d0, d1 in ((0, pa.info.ndm), (5, 17)):
                S2, W2 = pb.rows(d0, d1)
                assert np.array_equal(S2.view(np.uint32), S[d0:d1].view(np.uint32)), (b, d0, d1)
                assert np.array_equal(W2, W[d0:d1]), (b, d0, d1)
            hits += a.size
    finally:
        beam.close()
    assert hits > 1000


def test_htr_4096ch_16us_four_rows_first_ready_block():
    """The high-time-resolution beam at full size: 4096 ch over 1000-1500
    MHz, tsamp 16 us, 0.5 s blocks (31,250 samples), DM 0-5000 in the four
    rows 0-500/0.2, 500-1500/1, 1500-3000/4, 3000-5000/10 (64, 64, 128, 128
    sub-bands) and the ladder 16 us - 5 ms (26 widths).  Points-only plans
    (the S/N rows recomputed on demand).  Each row's series are bit-exact
    against the oracle at the block where its boxcar gate opens (up to 24
    blocks, 12 s, in), its tuples -- points and widths -- meet the tuple
    contract, and a burst placed in each row's first window is found at its
    DM."""
    import os

    import torch

    from transientx_amd.synth import gpu_generate

    C, tsamp, ndump = 4096, 16e-6, 31250
    rows = [(0.0, 0.2, 2500, 64), (500.0, 1.0, 1000, 64), (1500.0, 4.0, 375, 128), (3000.0, 10.0, 200, 128)]
    burst_dm = [300.0, 1000.0, 2200.0, 4100.0]
    freqs = Observation(nchans=C, fbot=1000.0, ftop=1500.0, tsamp=tsamp, nsamp=1).freqs
    ops = [oracle.Plan(freqs=freqs, tsamp=tsamp, ndump=ndump, dms=d, ddm=dd, ndm=n, overlap=0.0, nsub=ns)
           for d, dd, n, ns in rows]
    ready = [-(-(o.info().offset + ndump) // ndump) for o in ops]
    nblocks = max(ready)
    # each burst in the middle of its row's first ready window (arrival at the top channel)
    pulses = [Pulse(dm=dm, t0=(o.window_start(b) + o.info().nsamples // 2) * tsamp, fwhm=3e-4, snr=30)
              for dm, o, b in zip(burst_dm, ops, ready)]
    obs = Observation(nchans=C, fbot=1000.0, ftop=1500.0, tsamp=tsamp, nsamp=ndump * nblocks, mean=128.0,
                      sd=12.0, seed=5, pulses=pulses)
    vwn = width_ladder(tsamp, 16e-6, 5e-3, 0.1)
    assert len(vwn) == 26
    say("case htr 4096ch 16us rows", rows, "ready at", ready)
    data, _ = gpu_generate(obs, ndump=ndump)
    beam = Beam(C, ndump)
    kws = [dict(freqs=obs.freqs, tsamp=tsamp, dms=d, ddm=dd, ndm=n, vwn=vwn, overlap=0.0, nsub=ns,
                store_maps=False) for d, dd, n, ns in rows]
    plans = [Plan(beam, **kw) for kw in kws]
    got = [None] * len(rows)
    try:
        for b in range(1, nblocks + 1):
            beam.push(data[(b - 1) * ndump:b * ndump])
            for r, p in enumerate(plans):
                p.dedisp_run()
                rd = p.boxcar_run(THRE)
                if rd and got[r] is None:
                    assert b == ready[r]
                    got[r] = (b, all_tim(p), p.rows(0, p.info.ndm), p.points())
                    say("row", r, "ready at block", b)
    finally:
        beam.close()
    x = data.cpu().numpy()
    del data
    torch.cuda.empty_cache()
    nth = min(16, os.cpu_count() or 1)
    for r, (b, tim, (S, W), pts) in enumerate(got):
        op = ops[r]
        o = op.info()
        w0 = op.window_start(b)
        a = max(0, w0)
        xo = x[a:w0 + o.nsamples + o.maxdelay].astype(np.float32) - 128.0
        ref = op.dedisperse(xo, w0 - a, nthreads=nth)
        del xo
        assert np.array_equal(tim.view(np.uint32), ref.view(np.uint32)), r
        rS, rW = oracle.boxcar(ref, vwn, nthreads=nth)
        assert_tuple_contract(S, W, rS, rW, ref, vwn, tag=f"htr row {r}")
        np.testing.assert_allclose(S, rS, rtol=2e-3, atol=2e-3)
        sel = S > THRE
        assert pts.size == int(sel.sum())
        assert np.array_equal(pts["S"], S[pts["idm"], pts["isamp"]])
        assert np.array_equal(pts["wn"], W[pts["idm"], pts["isamp"]])
        i, t = np.unravel_index(np.argmax(S), S.shape)
        dm_peak = rows[r][0] + i * rows[r][1]
        say("row", r, "points", pts.size, "peak DM", dm_peak, "S", float(S[i, t]))
        assert abs(dm_peak - burst_dm[r]) <= 3 * rows[r][1] + 0.02 * burst_dm[r], r
