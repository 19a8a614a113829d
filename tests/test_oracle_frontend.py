"""Pins of the oracle's front end (txo_frontend_u8, txo_frontend_stats): the
declared Downsample + Equalize of 8-bit blocks that SinglePulse::run applies
before dedispersion (src/singlepulse/singlepulse.cpp:289-291; the XLibs
classes themselves are not in the TransientX tree).  Downsample is pinned to
integer block sums and to Candidate::downsample's float loop
(src/singlepulse/candidate.cpp:381-405) on 8-bit data; Equalize to numpy
float64 statistics and the zero-mean unit-variance property that
docs/pipeline/preprocess.rst:15-22 states."""
import numpy as np
import pytest

import oracle


def _block(seed, nsi, nchi, lo=0, hi=256):
    return np.random.default_rng(seed).integers(lo, hi, size=(nsi, nchi)).astype(np.uint8)


def _neighbour_sums(x, td, fd):
    nso, ncho = x.shape[0] // td, x.shape[1] // fd
    return x[: nso * td, : ncho * fd].astype(np.int64).reshape(nso, td, ncho, fd).sum((1, 3))


def _candidate_downsample(x, td, fd):
    """Candidate::downsample's loops on the channel-major frame: float sums,
    channel group jj outer, time neighbour ii inner."""
    data = np.ascontiguousarray(x.T, np.float32)  # [nchan][nbin]
    nchan, nbin = data.shape
    nchan_new, nbin_new = nchan // fd, nbin // td
    out = np.zeros((nchan_new, nbin_new), np.float32)
    for j in range(nchan_new):
        for jj in range(fd):
            for ii in range(td):
                out[j] += data[j * fd + jj, ii: ii + td * nbin_new: td]
    return out.T


@pytest.mark.parametrize("nsi,nchi,td,fd", [(64, 32, 1, 1), (67, 33, 2, 1), (90, 40, 3, 4), (128, 64, 8, 2),
                                           (31, 7, 4, 3)])
def test_frontend_is_neighbour_sums_then_affine(nsi, nchi, td, fd):
    x = _block(nsi * 7 + nchi, nsi, nchi)
    ncho = nchi // fd
    rng = np.random.default_rng(td * 10 + fd)
    off = rng.uniform(0, 255 * td * fd, ncho).astype(np.float32)
    gain = rng.uniform(0.01, 2.0, ncho).astype(np.float32)
    mask = (rng.random(ncho) > 0.2).astype(np.uint8)
    got = oracle.frontend_u8(x, td, fd, off, gain, mask)
    y = _neighbour_sums(x, td, fd)
    exp = (y.astype(np.float32) - off) * gain  # float32 subtract, then float32 multiply
    exp[:, mask == 0] = 0.0
    assert got.shape == (nsi // td, ncho)
    assert np.array_equal(got.view(np.uint32), exp.astype(np.float32).view(np.uint32))
    # the neighbour sums are Candidate::downsample's float sums of the same samples
    ones = np.ones(ncho, np.float32)
    raw = oracle.frontend_u8(x, td, fd, np.zeros(ncho, np.float32), ones)
    assert np.array_equal(raw, _candidate_downsample(x, td, fd))


def test_frontend_constant_input_and_no_mask():
    x = np.full((40, 12), 200, np.uint8)
    off, gain = oracle.frontend_stats(x, 4, 3)
    assert np.all(off == 200 * 12) and np.all(gain == 1.0)  # variance 0: left unscaled
    assert np.count_nonzero(oracle.frontend_u8(x, 4, 3, off, gain)) == 0
    # the largest sums stay exact in float: 255 * td * fd
    x = np.full((16, 16), 255, np.uint8)
    got = oracle.frontend_u8(x, 16, 16, np.zeros(1, np.float32), np.ones(1, np.float32))
    assert got.shape == (1, 1) and got[0, 0] == 255 * 256


@pytest.mark.parametrize("td,fd", [(1, 1), (2, 2), (5, 1)])
def test_frontend_stats_match_float64_moments(td, fd):
    x = _block(td + 3 * fd, 503, 48, lo=90, hi=170)
    x[:, 5] = 17  # a constant channel
    off, gain = oracle.frontend_stats(x, td, fd)
    y = _neighbour_sums(x, td, fd).astype(np.float64)
    # integer data: numpy's sums are exact, so the mean is the correctly rounded S1/n
    assert np.array_equal(off, y.mean(0).astype(np.float32))
    sd = y.std(0)
    live = sd > 0
    ref = (1.0 / sd[live]).astype(np.float32)
    assert np.max(np.abs(gain[live] - ref) / ref) < 2.5e-7
    assert np.all(gain[~live] == 1.0)
    # the equalized block: zero mean and unit variance per live channel
    z = oracle.frontend_u8(x, td, fd, off, gain).astype(np.float64)
    assert np.max(np.abs(z.mean(0))) < 1e-5
    assert np.max(np.abs(z.std(0)[live] - 1.0)) < 1e-5
