"""Pins of the oracle's refinement grid (Candidate::dedisperse, optimize with
the dedispersion tree, matched_filter; src/singlepulse/candidate.cpp:108-140,899-1087,
include/hough_transform.h:20-125).

The tree is checked against the reference's own include/hough_transform.h,
compiled from source into oracle/_ref with the reference's flags; the rest
against closed forms, FP64 brute force and injected-pulse recovery."""
import numpy as np
import pytest

K_DM = 4.148741601e3


def desc_freqs(n, f0=1500.0, f1=1200.0):
    return np.linspace(f0, f1, n)


# ------------------------------------------------------------ Candidate::dedisperse
def test_cand_dedisperse_rotates_rows_by_rounded_delay(ora):
    rng = np.random.default_rng(1)
    f = desc_freqs(48)
    x = rng.normal(size=(48, 200)).astype(np.float32)
    tbin, dm = 1e-3, 123.4
    got = ora.cand_dedisperse(x, f, dm, tbin)
    for j in range(48):
        d = int(np.floor(K_DM * dm * (1 / f[j] ** 2 - 1 / f.max() ** 2) / tbin + 0.5))
        assert np.array_equal(got[j], np.roll(x[j], -(d % 200))), j


# --------------------------------------------------------------------- the tree
@pytest.mark.parametrize("nrow,ncol", [(2, 16), (8, 100), (64, 512), (256, 64), (1024, 32)])
@pytest.mark.parametrize("dedispersed", [False, True])
def test_tree_matches_reference_header(ora, nrow, ncol, dedispersed):
    ref = ora.ref_lib()
    if ref is None:
        pytest.skip("reference tree absent: oracle/_ref not built")
    rng = np.random.default_rng(nrow * 7 + ncol)
    x = rng.normal(size=(nrow, ncol)).astype(np.float32)
    fP = np.sort(rng.uniform(1100.0, 1600.0, nrow))[::-1].copy()
    dm0 = 317.3
    dms, ddm, tbin = rng.uniform(0, 300), rng.uniform(0.01, 2.0), 98.304e-6
    got = ora.hough(x, fP, dms, ddm, tbin, dm0 if dedispersed else 0.0)
    want = np.zeros_like(x)
    ref.ref_hough(x.ctypes.data, want.ctypes.data, nrow, ncol, fP.ctypes.data, dms, ddm, tbin,
                  dm0, int(dedispersed))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_tree_zero_dm_rows_are_channel_sums(ora):
    rng = np.random.default_rng(2)
    x = rng.integers(-20, 20, (128, 64)).astype(np.float32)   # sums exact in FP32
    out = ora.hough(x, desc_freqs(128), 0.0, 0.0, 1e-3)
    assert np.array_equal(out, np.tile(x.sum(0), (128, 1)))


def test_tree_is_linear(ora):
    rng = np.random.default_rng(3)
    a = rng.integers(-50, 50, (32, 40)).astype(np.float32)
    b = rng.integers(-50, 50, (32, 40)).astype(np.float32)
    args = (desc_freqs(32), 10.0, 3.0, 1e-3)
    assert np.array_equal(ora.hough(a + b, *args), ora.hough(a, *args) + ora.hough(b, *args))


def test_tree_recovers_a_linear_sweep(ora):
    """With 1/f^2 linear in channel and integer unit delays, trial p of the
    tree sums channel j shifted by p_l * 2^l over the levels l where j is
    the second child (p_l = p's top l+1 bits): an impulse swept at the
    matching rate lands in one output sample."""
    n, ncol, tbin = 16, 64, 1e-3
    a, bb = 1 / 1500.0 ** 2, 1e-8
    f = 1 / np.sqrt(a + bb * np.arange(n))
    ddm = tbin / (K_DM * bb) / 2  # delay step between adjacent channels = p/2 samples
    x = np.zeros((n, ncol), np.float32)
    p_true = 8                    # a power of two: exact at every level
    for j in range(n):
        x[j, (5 + (p_true * j) // 2) % ncol] = 1.0
    out = ora.hough(x, f, 0.0, ddm, tbin)
    assert out[p_true, 5] == n and out[p_true].sum() == n
    assert out.max() == n and np.argwhere(out == n).tolist() == [[p_true, 5]]


# ------------------------------------------------------------- matched_filter
def test_mf_ladder_is_the_search_ladder_from_one_sample(ora):
    # from 1 sample with snrloss 0.1 the ladder is the search ladder's
    # sequence 1..9, 11, 13, 16, ... (integer sequence worked out by hand)
    want = [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 13, 16, 19, 23, 28, 34, 41, 50, 61, 75, 92, 113,
            139, 171, 211]
    assert ora.mf_ladder(512, 0.1).tolist() == want
    assert ora.mf_ladder(40, 0.1).tolist() == [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 13, 16, 19]


def test_mf_constant_map_gives_zeros(ora):
    S, W, j, i, s = ora.refine_mf(np.full((8, 64), 3.0, np.float32), 0.1)
    assert not S.any() and not W.any() and (j, i, s) == (0, 0, 0.0)


@pytest.mark.parametrize("i0", [0, 1, 37, 63])
def test_mf_spike_closed_form_and_circular_wrap(ora, i0):
    n, h = 64, 10.0
    m = np.zeros((4, n), np.float32)
    m[2, i0] = h
    S, W, j, i, s = ora.refine_mf(m, 0.1)
    mean = h / n
    var = h * h / n - mean * mean
    # wn = 1 at the spike: (h - mean)/sqrt(var); wider boxes only add -mean
    assert (j, i) == (2, i0) and W[2, i0] == 1
    assert abs(s - (h - mean) / np.sqrt(var)) < 1e-5 * s


def test_mf_matches_fp64_circular_boxcar(ora):
    rng = np.random.default_rng(4)
    m = rng.normal(size=(16, 128)).astype(np.float32)
    m[5, 40:52] += 2.5
    S, W, j, i, s = ora.refine_mf(m, 0.1)
    vwn = ora.mf_ladder(128, 0.1)
    for r in range(16):
        x = m[r].astype(np.float64)
        mu, var = x.mean(), x.var()
        c = np.concatenate([[0.0], np.cumsum(np.tile(x - mu, 2))])  # c[k] = sum of first k
        best = np.zeros(128)
        for wn in vwn:
            wl, wh = wn // 2 + 1, (wn - 1) // 2
            for q in range(wl, 128 + wl):
                val = (c[q + wh + 1] - c[q - wl + 1]) / np.sqrt(wn * var)
                o = q % 128
                best[o] = max(best[o], val)
        assert np.max(np.abs(best - S[r])) < 2e-4 * max(1.0, np.abs(best).max())
    assert (j, 40 <= i < 52) == (5, True)


# --------------------------------------------------------------------- optimize
def test_optimize_is_band_order_invariant_and_pads(ora):
    rng = np.random.default_rng(5)
    f = desc_freqs(48)   # 48 channels -> 64 trials, 16 zero rows
    x = rng.normal(size=(48, 128)).astype(np.float32)
    d1, dms1, ddm1, n1 = ora.refine_optimize(x, f, 1e-3, 200.0)
    d2, dms2, ddm2, n2 = ora.refine_optimize(x[::-1], f[::-1], 1e-3, 200.0)
    assert (n1, n2) == (64, 64) and (dms1, ddm1) == (dms2, ddm2)
    assert np.array_equal(d1, d2)
    dmrange = 128 * 1e-3 / (K_DM * (1 / 1200.0 ** 2 - 1 / 1500.0 ** 2))
    assert dms1 == pytest.approx(200.0 - dmrange / 2) and dms1 + n1 * ddm1 == pytest.approx(200 + dmrange / 2)
    _, dms0, _, _ = ora.refine_optimize(x, f, 1e-3, 1.0)
    assert dms0 == 0.0  # clipped at zero


def test_grid_recovers_injected_pulse(ora):
    """A pulse at DM 300, cut out and dedispersed at DM 296 (the search's
    coarse DM), is refined back to within two grid steps of 300 and to its
    arrival sample."""
    nchan, nbin, tbin = 64, 256, 1e-3
    f = desc_freqs(nchan)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(nchan, nbin)).astype(np.float32)
    t0 = 100
    for j in range(nchan):
        d = K_DM * 300.0 * (1 / f[j] ** 2 - 1 / f.max() ** 2) / tbin
        k = int(np.floor(t0 + d + 0.5)) % nbin
        x[j, k] += 3.0
        x[j, (k + 1) % nbin] += 3.0
    xd = ora.cand_dedisperse(x, f, 296.0, tbin)
    dmt, dms, ddm, ndm = ora.refine_optimize(xd, f, tbin, 296.0, isdedispersed=True)
    S, W, j, i, s = ora.refine_mf(dmt, 0.1)
    assert abs(dms + j * ddm - 300.0) <= 2 * ddm + 1e-9
    assert abs(i - t0) <= 2 and s > 10
