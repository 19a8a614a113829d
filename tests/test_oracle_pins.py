"""Pins of the CPU oracle: closed forms, brute force, hand-derived cases.

Each test is chosen so that a plausible slip in the oracle (a dropped term,
a sign, an off-by-one index, a '>' for a '>=') makes it fail.
"""
import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components

from transientx_amd.search import width_ladder
from transientx_amd.synth import channel_freqs


# ------------------------------------------------------------- a1 dmdelay
def test_dmdelay_published_constant(ora):
    # Lorimer & Kramer (2005) eq. 4.7: dt = 4.15 ms * DM * (f/GHz)^-2 against
    # infinite frequency.  Pins units (s, MHz) and the inverse-square law.
    d = ora.dmdelay(1.0, 1e12, 1000.0)
    assert abs(d * 1e3 - 4.1487) < 2e-4
    assert abs(ora.dmdelay(100.0, 1e12, 2000.0) - 0.1037185) < 1e-6


def test_dmdelay_invariants(ora):
    a, b, c = 1500.0, 1300.0, 1100.0
    assert ora.dmdelay(10, a, c) == pytest.approx(ora.dmdelay(10, a, b) + ora.dmdelay(10, b, c))
    assert ora.dmdelay(10, a, c) == pytest.approx(-ora.dmdelay(10, c, a))
    assert ora.dmdelay(30, a, c) == pytest.approx(3 * ora.dmdelay(10, a, c))
    assert ora.dmdelay(10, a, a) == 0.0


# ------------------------------------------------------------- a3 ladder
LADDER_C0 = [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 13, 16, 19, 23, 28, 34, 41, 50, 61, 75, 92, 113, 139,
             171, 211, 260, 320, 395, 487, 601, 741, 914]


def test_ladder_sequences(ora):
    # worked out by hand from src/singlepulse/singlepulse.cpp:242-254:
    # next = int(float(prev) * float(1/0.81)), at least prev+1, stop past maxw
    c0 = list(ora.ladder(98.304e-6, 1e-4, 0.1, 0.1))
    assert c0 == LADDER_C0
    c1 = list(ora.ladder(76.56e-6, 1e-4, 0.1, 0.1))
    assert c1 == LADDER_C0 + [1128]
    c4 = list(ora.ladder(16e-6, 16e-6, 5e-3, 0.1))
    assert len(c4) == 26 and c4[-1] == 260 and c4[:9] == list(range(1, 10))


@pytest.mark.parametrize("tsamp,minw,maxw,loss", [
    (98.304e-6, 1e-4, 0.1, 0.1), (76.56e-6, 1e-4, 0.1, 0.1), (16e-6, 16e-6, 5e-3, 0.1),
    (64e-6, 1e-3, 0.02, 0.05), (1e-3, 1e-4, 1.0, 0.3), (256e-6, 2e-3, 0.5, 0.15)])
def test_ladder_product_matches_oracle(ora, tsamp, minw, maxw, loss):
    assert list(width_ladder(tsamp, minw, maxw, loss)) == list(ora.ladder(tsamp, minw, maxw, loss))


def test_ladder_minw_clamped_to_tsamp(ora):
    # minw below tsamp becomes tsamp -> first width 1
    assert ora.ladder(1e-3, 1e-5, 0.01, 0.1)[0] == 1
    assert ora.ladder(1e-3, 3.4e-3, 0.1, 0.1)[0] == 3


# ------------------------------------------------------- a4 plan / tables
def _plan(ora, C=64, nsub=8, group=0, ndm=40, ddm=2.0, dms=0.0, ndump=512, overlap=0.0,
          fbot=1200.0, ftop=1500.0, tsamp=1e-3):
    return ora.Plan(freqs=channel_freqs(C, fbot, ftop), tsamp=tsamp, ndump=ndump, dms=dms,
                    ddm=ddm, ndm=ndm, overlap=overlap, nsub=nsub, group=group)


def _ideal(freqs, tsamp, dm):
    fmax = freqs.max()
    k = 4.148741601e3
    return np.array([np.round(k * dm * (1 / f**2 - 1 / fmax**2) / tsamp) for f in freqs], int)


def test_plan_reduces_to_direct_when_one_channel_per_subband(ora):
    p = _plan(ora, C=32, nsub=32)
    d1, d2, dmc, fref = p.tables()
    assert np.all(d1 == 0)
    for i in range(p.ndm):
        assert np.array_equal(d2[i], _ideal(p.freqs, p.tsamp, p.dms + i * p.ddm))


def test_plan_reduces_to_direct_with_one_subband_and_group_one(ora):
    p = _plan(ora, C=32, nsub=1, group=1)
    d1, d2, dmc, fref = p.tables()
    assert np.all(d2 == 0)
    assert np.allclose(dmc, p.dms + p.ddm * np.arange(p.ndm))
    for i in range(p.ndm):
        assert np.array_equal(d1[i], _ideal(p.freqs, p.tsamp, p.dms + i * p.ddm))


def test_plan_group_matches_smearing_rule(ora):
    # 1024 channel centres over 1200-1500 MHz: fmin 1200.146, fmax 1499.854, the
    # lowest 32-channel sub-band tops out at 1209.229 MHz, so
    # (1/fmin^2-1/fmax^2)/(1/fmin^2-1/fref^2) = 2.4974e-7/1.0391e-8 = 24.03 -> g = 25
    p = _plan(ora, C=1024, nsub=32, ndm=1000, ddm=1.0, tsamp=98.304e-6, ndump=20345)
    o = p.info()
    assert (o.group, o.ndmc) == (25, 40)
    assert o.noverlap == 0 and o.nsamples == 20345
    p2 = _plan(ora, C=4096, nsub=64, ndm=1000, ddm=0.5, fbot=856, ftop=1712, tsamp=76.56e-6,
               ndump=26123, overlap=0.1)
    o2 = p2.info()
    assert o2.ndmc == 40 and o2.noverlap == 2612 and o2.nsamples == 28735
    assert o2.offset == o2.D1 + o2.D2 + o2.noverlap


# ----------------------------------------------------- a5 dedispersion
def _ints(rng, n, C):
    return rng.integers(-20, 21, size=(n, C)).astype(np.float32)


@pytest.mark.parametrize("nsub,group", [(32, 0), (1, 1)])
def test_dedisp_equals_direct_f64_in_reducing_cases(ora, nsub, group):
    rng = np.random.default_rng(5)
    p = _plan(ora, C=32, nsub=nsub, group=group, ndm=12, ddm=7.0, ndump=256, overlap=0.25)
    o = p.info()
    x = _ints(rng, 4 * 256 + o.maxdelay, 32)
    for blk in (1, 2, 4):
        w0 = p.window_start(blk)
        tim = p.dedisperse(x, w0)
        ref = ora.direct_f64(p.freqs, p.tsamp, p.dms, p.ddm, p.ndm, x, w0, o.nsamples)
        assert np.array_equal(tim.astype(np.float64), ref)


def test_dedisp_zero_dm_is_channel_sum(ora):
    rng = np.random.default_rng(6)
    p = _plan(ora, ndm=1, ddm=0.0)
    o = p.info()
    x = _ints(rng, 3 * 512 + o.maxdelay, 64)
    w0 = p.window_start(3)
    tim = p.dedisperse(x, w0)
    assert np.array_equal(tim[0], x[w0:w0 + o.nsamples].sum(axis=1))


def test_dedisp_linear(ora):
    rng = np.random.default_rng(7)
    p = _plan(ora)
    o = p.info()
    n = 3 * 512 + o.maxdelay
    x, y = _ints(rng, n, 64), _ints(rng, n, 64)
    w0 = p.window_start(3)
    assert np.array_equal(p.dedisperse(3 * x - 2 * y, w0),
                          3 * p.dedisperse(x, w0) - 2 * p.dedisperse(y, w0))


def test_dedisp_impulse_response(ora):
    # one sample in channel j at time t lands at tau = t - d2[i][s(j)] - d1[k(i)][j]
    p = _plan(ora, ndm=40, ddm=3.0)
    o = p.info()
    d1, d2, _, _ = p.tables()
    n = 4 * 512 + o.maxdelay
    w0 = p.window_start(4)
    for j in (0, 17, 63):
        x = np.zeros((n, 64), np.float32)
        t = w0 + o.maxdelay + 100
        x[t, j] = 1.0
        tim = p.dedisperse(x, w0)
        for i in range(p.ndm):
            tau = t - d2[i][j // o.cps] - d1[i // o.group][j] - w0
            expect = np.zeros(o.nsamples, np.float32)
            if 0 <= tau < o.nsamples:
                expect[tau] = 1
            assert np.array_equal(tim[i], expect), (j, i)


def test_dedisp_negative_times_are_zero(ora):
    p = _plan(ora)
    o = p.info()
    x = np.ones((512 + o.maxdelay + 1, 64), np.float32)
    tim = p.dedisperse(x, p.window_start(1))
    # every sample of the first window needs some input before t=0 at the top DM
    assert tim[-1].min() < 64 and tim[0].max() == 64


def test_dedisp_recovers_injected_pulse(ora):
    from transientx_amd.synth import Observation, Pulse, generate

    obs = Observation(nchans=128, fbot=1200, ftop=1500, tsamp=5e-4, nsamp=6000, sd=16, seed=3,
                      pulses=[Pulse(dm=210.0, t0=1.2, fwhm=4e-3, snr=40)])
    x = generate(obs).astype(np.float32) - 128
    p = ora.Plan(freqs=obs.freqs, tsamp=obs.tsamp, ndump=2000, dms=100, ddm=4.0, ndm=60,
                 nsub=16)
    w0 = p.window_start(2)
    assert w0 <= 2400 < w0 + p.info().nsamples  # the pulse (t0 = 1.2 s) is in this window
    tim = p.dedisperse(x, w0)
    mxS, _ = ora.boxcar(tim, [1, 2, 4, 8, 16, 32])
    i, t = np.unravel_index(np.argmax(mxS), mxS.shape)
    assert abs((100 + 4.0 * i) - 210.0) <= 4.0
    assert abs((w0 + t) * obs.tsamp - 1.2) < 0.01


# ------------------------------------------------------- a11-a13 boxcar
def test_boxcar_constant_row_is_zero(ora):
    S, w = ora.boxcar(np.full((1, 1000), 3.5, np.float32), LADDER_C0[:10])
    assert not S.any() and not w.any()


def test_boxcar_single_spike_closed_form(ora):
    N, p, h = 997, 500, 8.0
    x = np.zeros(N, np.float32)
    x[p] = h
    widths = [1, 3, 4, 9]
    vS, vw, mean, var = ora.boxcar_row_stats(x, widths)
    m = h / N
    v = h * h / N - m * m
    assert mean == pytest.approx(m, rel=1e-6) and var == pytest.approx(v, rel=1e-5)
    best = np.zeros(N)
    bw = np.zeros(N, int)
    for wn in widths:
        wl, wh = wn // 2 + 1, (wn - 1) // 2
        for i in range(wl, wl + (N - wh - wl) // 8 * 8):
            box = (h if i - wl < p <= i + wh else 0.0) - wn * m
            s = box / np.sqrt(wn * v)
            if s >= best[i]:
                best[i], bw[i] = s, wn
    assert np.allclose(vS, best, atol=2e-4)
    assert np.array_equal(vw[best > 1e-3], bw[best > 1e-3])
    # the spike itself peaks at wn=1, centred
    assert vw[p] == 1 and vS[p] == pytest.approx((h - m) / np.sqrt(v), rel=1e-4)


def test_boxcar_window_sums_match_f64(ora):
    rng = np.random.default_rng(11)
    x = rng.normal(0, 3, 4001).astype(np.float32)
    widths = [1, 2, 5, 13, 40]
    vS, vw, mean, var = ora.boxcar_row_stats(x, widths)
    xd = x.astype(np.float64)
    assert mean == pytest.approx(xd.mean(), abs=1e-5)
    assert var == pytest.approx(xd.var(), rel=1e-4)
    N = x.size
    S = np.full((len(widths), N), -np.inf)
    for k, wn in enumerate(widths):
        wl, wh = wn // 2 + 1, (wn - 1) // 2
        c = np.concatenate([[0.0], np.cumsum(xd - mean)])
        for i in range(wl, wl + (N - wh - wl) // 8 * 8):
            S[k, i] = (c[i + wh + 1] - c[i - wl + 1]) / np.sqrt(wn * var)
    best = np.maximum(S.max(axis=0), 0)
    top2 = np.sort(S, axis=0)[-2:]
    clear = (top2[1] - top2[0] > 1e-3) & (best > 0.01)
    assert np.allclose(vS, best, atol=2e-4)
    arg = np.array(widths)[np.argmax(S, axis=0)]
    assert np.array_equal(vw[clear], arg[clear])


def test_boxcar_ties_take_the_wider_width(ora):
    # +h and -h spikes: mean is exactly 0, so every box holding neither spike
    # sums to exactly 0 and S == 0 ties with the initial vS == 0 at every
    # width; S >= vS (boxcar.cpp:257,273) keeps the widest width covering i.
    N = 512
    x = np.zeros(N, np.float32)
    x[10], x[20] = 5, -5
    widths = [1, 2, 4, 8]
    vS, vw, mean, var = ora.boxcar_row_stats(x, widths)
    assert mean == 0.0
    i = 300
    assert vS[i] == 0.0 and vw[i] == 8
    # the scalar branch agrees on ties
    vS1, vw1, _, _ = ora.boxcar_row_stats(x, widths, variant=1)
    assert vw1[i] == 8


def test_boxcar_avx2_drops_tail(ora):
    # AVX2 branch (boxcar.cpp:264-279) covers 8*floor((N-wh-wl)/8) outputs per
    # width starting at wl; the scalar branch covers [wl, N-wh).
    rng = np.random.default_rng(2)
    N = 1003
    x = rng.normal(0, 1, N).astype(np.float32) + 5
    x[-2:] = 50.0  # strong samples inside the scalar range, past the AVX2 range
    vS0, vw0, _, _ = ora.boxcar_row_stats(x, [1], variant=0)
    vS1, vw1, _, _ = ora.boxcar_row_stats(x, [1], variant=1)
    wl, wh = 1, 0
    hi = wl + (N - wh - wl) // 8 * 8
    assert hi == 1001
    assert np.all(vS0[hi:] == 0) and np.all(vw0[hi:] == 0)
    assert np.array_equal(vS0[:hi], vS1[:hi])
    assert np.all(vS1[hi:N - wh] > 10) and np.all(vw1[hi:N - wh] == 1)
    assert vS0[0] == 0 and vw0[0] == 0  # i < wl is never written


def test_boxcar_iqr_statistics(ora):
    rng = np.random.default_rng(4)
    x = rng.normal(2, 5, 3001).astype(np.float32)
    _, _, mean, var = ora.boxcar_row_stats(x, [1, 2], iqr=True)
    s = np.sort(x)
    N = x.size
    q1, q2, q3 = s[N // 4], s[N // 2], s[N - 1 - N // 4]
    assert mean == q2
    assert var == np.float32(((float(q3) - float(q1)) / 1.349) ** 2)


def test_boxcar_fma_contraction_changes_stats(ora):
    # the reference build contracts var += x*x; the flag must matter somewhere
    rng = np.random.default_rng(9)
    diff = 0
    for _ in range(20):
        x = rng.normal(1000, 30, 5000).astype(np.float32)
        _, _, _, v1 = ora.boxcar_row_stats(x, [1], use_fma=True)
        _, _, _, v0 = ora.boxcar_row_stats(x, [1], use_fma=False)
        diff += v1 != v0
    assert diff > 0


# --------------------------------------------------------- a15 threshold
def test_boxcar_runnerup_against_brute_force(ora):
    """best = the map's S/N at the point; second = the largest S/N of the
    other widths, brute force over box sums of a zero-sum integer series
    (exact in FP32) with the oracle's own scale per width."""
    rng = np.random.default_rng(12)
    vwn = np.array([1, 2, 3, 5, 8, 13, 21], np.int32)
    rows = []
    for k in range(3):
        x = rng.integers(-40, 41, 700).astype(np.float64)
        x[100 + 50 * k:110 + 50 * k] += 60
        x[-1] -= x.sum()  # zero sum: mean 0, every running sum an exact integer
        rows.append(x.astype(np.float32))
    tim = np.stack(rows)
    mxS, mxwn = ora.boxcar(tim, vwn)
    idm, isamp = np.nonzero(mxS > 2.0)
    best, second = ora.boxcar_runnerup(tim, vwn, idm, isamp)
    assert idm.size > 20
    assert np.array_equal(best.view(np.uint32), mxS[idm, isamp].view(np.uint32))
    N = tim.shape[1]
    for p in range(idm.size):
        x = tim[idm[p]].astype(np.float64)
        _, _, mean, var = ora.boxcar_row_stats(tim[idm[p]], vwn)
        assert mean == 0.0
        cs = np.cumsum(x)
        i = isamp[p]
        vals = []
        for wn in vwn:
            wl, wh = wn // 2 + 1, (wn - 1) // 2
            hi = wl + (N - wh - wl) // 8 * 8
            if not (wl <= i < hi):
                vals.append(None)
                continue
            temp = np.float32(np.sqrt(1.0 / (np.float64(np.float32(wn) * np.float32(var)))))
            vals.append(np.float32((cs[i + wh] - cs[i - wl]) * np.float64(temp)))
        live = [(v, k) for k, v in enumerate(vals) if v is not None]
        top = max(v for v, _ in live)
        kwin = max(k for v, k in live if v == top)  # ties: the wider (later) width
        others = [v for v, k in live if k != kwin]
        assert best[p] == max(top, np.float32(0))
        assert second[p] == (max(others) if others else -np.finfo(np.float32).max)
        assert mxwn[idm[p], i] == vwn[kwin]


def test_dump_values_properties(ora):
    """--savetim values (declared rule): unit-variance channels summed over
    nlive come out with the requested mean and std; rounding is half away
    from zero; 8 and 16 bits clamp to the signed range; the 8- and 16-bit
    values are the rounded 32-bit ones wherever those are in range."""
    rng = np.random.default_rng(5)
    nlive, ndm, N, nov = 400, 3, 50000, 1000
    tim = (rng.normal(0, 1, (ndm, N)) * np.sqrt(nlive)).astype(np.float32)
    y32 = ora.dump(tim, nov, nlive, 5.0, 3.0, 32)
    assert y32.shape == (ndm, N - nov)
    assert abs(y32.mean() - 5.0) < 0.05 and abs(y32.std() - 3.0) < 0.05
    y8 = ora.dump(tim, nov, nlive, 5.0, 3.0, 8)
    y16 = ora.dump(tim, nov, nlive, 5.0, 3.0, 16)
    r = np.sign(y32.astype(np.float64)) * np.floor(np.abs(y32.astype(np.float64)) + 0.5)
    assert np.array_equal(y16.astype(np.float64), np.clip(r, -32768, 32767))
    assert np.array_equal(y8.astype(np.float64), np.clip(r, -128, 127))
    # exact halves: nlive 1, std 1 gives y = x + mean
    t = np.array([[0.0, 2.5, -2.5, 0.5, -0.5, 1e6, -1e6, 126.5, -128.5]], np.float32)
    assert ora.dump(t, 0, 1, 0.0, 1.0, 8).tolist() == [[0, 3, -3, 1, -1, 127, -128, 127, -128]]
    assert ora.dump(t, 0, 1, 0.0, 1.0, 16).tolist() == [[0, 3, -3, 1, -1, 32767, -32768, 127, -129]]
    # no live channel: every value is the mean
    assert (ora.dump(t, 0, 0, 7.0, 3.0, 8) == 7).all()


def test_fp64_restatements(ora):
    """txo_dedisp_window_f64 equals the FP32 window on integer data (every
    sum exact in both) and FP64 direct dedispersion with one channel per
    sub-band; txo_boxcar_f64 gives the single-spike closed form and stays
    within FP32 rounding of the FP32 boxcar's maximum."""
    from transientx_amd.synth import channel_freqs

    rng = np.random.default_rng(41)
    freqs = channel_freqs(32, 1200, 1500)
    op = ora.Plan(freqs=freqs, tsamp=1e-3, ndump=256, dms=0.0, ddm=3.0, ndm=20, overlap=0.25, nsub=4)
    o = op.info()
    x = rng.integers(-60, 61, size=(4 * 256 + o.maxdelay, 32)).astype(np.float32)
    w0 = 300
    assert np.array_equal(op.dedisperse_f64(x, w0), op.dedisperse(x, w0).astype(np.float64))
    one = ora.Plan(freqs=freqs, tsamp=1e-3, ndump=256, dms=0.0, ddm=3.0, ndm=20, overlap=0.25, nsub=32)
    xf = rng.normal(0, 1, size=(4 * 256 + one.info().maxdelay, 32)).astype(np.float32)
    d = ora.direct_f64(freqs, 1e-3, 0.0, 3.0, 20, xf, w0, one.info().nsamples)
    assert np.allclose(one.dedisperse_f64(xf, w0), d, rtol=0, atol=1e-12)
    # one spike of height h on zeros: S at the spike for wn = 1 is (h - mean)/sd
    N, h = 512, 50.0
    row = np.zeros(N)
    row[200] = h
    S = ora.boxcar_f64(row, np.array([1], np.int32))
    mean = h / N
    sd = np.sqrt(h * h / N - mean * mean)
    assert abs(S[0, 200] - (h - mean) / sd) < 1e-12
    vwn = np.array([1, 2, 3, 5, 8, 13], np.int32)
    t = rng.normal(0, 1, size=(4, 4000)).astype(np.float32)
    S64 = ora.boxcar_f64(t.astype(np.float64), vwn)
    S32, _ = ora.boxcar(t, vwn)
    assert np.abs(S64 - S32).max() < 1e-4


def test_threshold_row_major_strict(ora):
    rng = np.random.default_rng(1)
    S = rng.normal(0, 3, (20, 300)).astype(np.float32)
    S[3, 7] = 7.0  # equal to the threshold: not a point (strict >)
    W = rng.integers(1, 9, S.shape).astype(np.int32)
    idm, isamp, s, wn = ora.threshold(S, W, 7.0)
    jj, ii = np.nonzero(S > 7.0)
    assert np.array_equal(idm, jj) and np.array_equal(isamp, ii)
    assert np.array_equal(s, S[jj, ii]) and np.array_equal(wn, W[jj, ii])


# ----------------------------------------------------- a16 clustering
def _coords(idm, isamp, dms, ddm, fmax, fmin, tsamp):
    k = 4.148741601e3
    x = np.array([k * (dms + j * ddm) * (1 / fmin**2 - 1 / fmax**2) * 0.5 for j in idm])
    return x, isamp * tsamp


def test_cluster_is_connected_components(ora):
    rng = np.random.default_rng(8)
    n = 400
    idm = rng.integers(0, 60, n).astype(np.int32)
    isamp = rng.integers(0, 3000, n).astype(np.int32)
    keys = np.unique(idm.astype(np.int64) * 10000 + isamp)
    idm, isamp = (keys // 10000).astype(np.int32), (keys % 10000).astype(np.int32)
    n = idm.size
    S = rng.uniform(7, 30, n).astype(np.float32)
    wn = rng.integers(1, 30, n).astype(np.int32)
    pars = dict(dms=0.0, ddm=2.0, fmax=1500.0, fmin=1200.0, tsamp=1e-3, radius=0.02)
    cand, labels = ora.cluster(idm, isamp, S, wn, maxncand=10**6, **pars)
    x, y = _coords(idm, isamp, pars["dms"], pars["ddm"], pars["fmax"], pars["fmin"], pars["tsamp"])
    d2 = (x[:, None] - x[None]) ** 2 + (y[:, None] - y[None]) ** 2
    ncomp, comp = connected_components(csr_matrix(d2 <= pars["radius"] ** 2), directed=False)
    assert labels.min() >= 1 and labels.max() == ncomp
    # same partition
    pairs = {(a, b) for a, b in zip(labels, comp)}
    assert len(pairs) == ncomp
    # clusters are numbered by decreasing seed S/N; each candidate is its
    # cluster's maximum, earliest in row-major order on ties
    best = [S[labels == c].max() for c in range(1, ncomp + 1)]
    assert all(best[i] >= best[i + 1] for i in range(ncomp - 1))
    assert len(cand) == ncomp
    for c, pt in enumerate(cand, start=1):
        assert labels[pt] == c and S[pt] == best[c - 1]
        assert pt == np.flatnonzero((labels == c) & (S == best[c - 1]))[0]


def test_cluster_maxncand_and_radius_floor(ora):
    idm = np.array([0, 0, 0, 0], np.int32)
    isamp = np.array([0, 1, 100, 200], np.int32)
    S = np.array([8, 9, 10, 11], np.float32)
    wn = np.ones(4, np.int32)
    # radius below 1.5*tsamp is raised to 1.5*tsamp (cluster.cpp:117): samples 0,1 join
    cand, labels = ora.cluster(idm, isamp, S, wn, 0.0, 1.0, 1500.0, 1200.0, 1e-3, 1e-6,
                               maxncand=100)
    assert list(labels) == [3, 3, 2, 1] and list(cand) == [3, 2, 1]
    cand, labels = ora.cluster(idm, isamp, S, wn, 0.0, 1.0, 1500.0, 1200.0, 1e-3, 1e-6,
                               maxncand=2)
    assert list(labels) == [0, 0, 2, 1] and list(cand) == [3, 2]


def test_cluster_minpts_erase_shift(ora):
    # cluster.cpp:145-158 erases candcluster[k] inside the loop over k, so a
    # later k reads the size of the cluster after it.  Clusters (by seed S/N):
    # sizes [3, 10, 10], minpts 5:
    #   k=0: size 3 < 5 -> erase -> sizes [10, 10]
    #   k=1: sizes[1] == 10 -> keep cluster 1
    #   k=2: past the end -> size 0 -> dropped
    idm, isamp, S = [], [], []
    for c, (n, s0, t0) in enumerate([(3, 30, 0), (10, 20, 1000), (10, 10, 2000)]):
        for q in range(n):
            idm.append(0)
            isamp.append(t0 + q)
            S.append(s0 - 0.01 * q)
    idm = np.array(idm, np.int32)
    isamp = np.array(isamp, np.int32)
    S = np.array(S, np.float32)
    wn = np.ones(idm.size, np.int32)
    cand, labels = ora.cluster(idm, isamp, S, wn, 0.0, 1.0, 1500.0, 1200.0, 1e-3, 2e-3,
                               maxncand=100, minpts=5)
    assert [labels[p] for p in cand] == [2]
    cand, _ = ora.cluster(idm, isamp, S, wn, 0.0, 1.0, 1500.0, 1200.0, 1e-3, 2e-3,
                          maxncand=100, minpts=0)
    assert [labels[p] for p in cand] == [1, 2, 3]
    # --drop removes a cluster whose best width is the widest searched
    wn2 = wn.copy()
    wn2[0] = 7
    cand, _ = ora.cluster(idm, isamp, S, wn2, 0.0, 1.0, 1500.0, 1200.0, 1e-3, 2e-3,
                          maxncand=100, minpts=0, drop=True, maxwn=7)
    assert [labels[p] for p in cand] == [2, 3]


# --------------------------------------------------------- kadane / fl fh
def test_kadane_brute_force(ora):
    rng = np.random.default_rng(12)
    for _ in range(200):
        n = int(rng.integers(1, 40))
        a = rng.normal(0, 1, n).astype(np.float32)
        s, e = ora.kadane(a)
        best, arg = -np.inf, None
        for i in range(n):
            acc = np.float32(0)
            for j in range(i, n):
                acc = np.float32(acc + a[j])
                if acc > best:
                    best, arg = acc, (i, j)
        assert np.float32(a[s:e + 1].sum(dtype=np.float64)) == pytest.approx(best, abs=1e-5)


def test_cand_flfh_band_edges(ora):
    nsub, N = 8, 100
    fsub = np.linspace(1500, 1200, nsub)
    sub = np.zeros((nsub, N), np.float32)
    sub[2:5, 48:53] = 10.0       # emission in sub-bands 2..4 around sample 50
    sub[:, :] -= 0.5
    fl, fh = ora.cand_flfh(sub, np.full(nsub, 4.0), fsub, 50, 3)
    assert (fl, fh) == (pytest.approx(np.float32(fsub[2])), pytest.approx(np.float32(fsub[4])))
