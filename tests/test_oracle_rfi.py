"""Pins of the oracle's RFI cleaning (txo_zdot, txo_kadanef), restating
Candidate::zdot and Candidate::kadaneF (src/singlepulse/candidate.cpp:710-897):
zdot against an independent FP64 least-squares fit (numpy lstsq) and its
invariants; kadaneF against brute-force maximum runs of integer-valued data
(exact float sums) and planted runs, including the wnlimit and threshold
rules."""
import numpy as np
import pytest

import oracle


def _zdot_f64(x):
    """zdot in FP64 from the definition: per-channel standardisation, the
    channel mean s, the least-squares fit a*s + b of every channel removed."""
    x = np.asarray(x, np.float64)
    m = x.mean(1, keepdims=True)
    sd = x.std(1, keepdims=True)
    sd[sd == 0] = 1.0
    z = (x - m) / sd
    s = z.mean(0)
    A = np.stack([s, np.ones_like(s)], 1)
    coef, *_ = np.linalg.lstsq(A, z.T, rcond=None)
    return z - (A @ coef).T, s


@pytest.mark.parametrize("nchan,nbin,seed", [(16, 64, 1), (48, 200, 2), (5, 17, 3)])
def test_zdot_matches_fp64_least_squares(nchan, nbin, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(3.0, 2.0, size=(nchan, nbin))
    # broadband (zero-DM) interference with a different gain per channel
    rfi = rng.normal(size=nbin) * 5
    x += rng.uniform(0.2, 2.0, size=(nchan, 1)) * rfi[None, :]
    x = x.astype(np.float32)
    got = oracle.zdot(x).astype(np.float64)
    ref, s = _zdot_f64(x)
    assert np.max(np.abs(got - ref)) < 2e-5 * max(1.0, np.max(np.abs(ref)))
    # the residual is orthogonal to 1 and to the channel mean
    assert np.max(np.abs(got.sum(1))) < 1e-3
    assert np.max(np.abs(got @ s)) < 1e-3 * max(1.0, np.linalg.norm(s))
    # the interference, common to every channel up to gain and offset, is gone
    assert np.abs(np.corrcoef(got.mean(0), rfi)[0, 1]) < 0.2 or np.max(np.abs(got.mean(0))) < 1e-4


def test_zdot_identical_and_constant_channels():
    rng = np.random.default_rng(4)
    base = rng.normal(size=100).astype(np.float32)
    x = np.stack([base * 3 + 1, base * 0.5 - 2, base, np.full(100, 7.0, np.float32)])
    got = oracle.zdot(x)
    # proportional channels equal the channel mean up to a fit: nothing left
    assert np.max(np.abs(got[:3])) < 1e-5
    # a constant channel: standard deviation 0 -> divided by 1 -> all zero
    assert np.array_equal(got[3], np.zeros(100, np.float32))


def _runs_bruteforce(a):
    """Maximum-sum contiguous runs of integer data (exact int64 sums)."""
    n = a.size
    c = np.concatenate([[0], np.cumsum(a.astype(np.int64))])
    best, arg = None, []
    for s in range(n):
        for e in range(s, n):
            v = c[e + 1] - c[s]
            if best is None or v > best:
                best, arg = v, [(s, e)]
            elif v == best:
                arg.append((s, e))
    return best, arg


def _kadanef_expected(x, td, fd, thr, wnlimit):
    nchan, nbin = x.shape
    nsd, ncd = nbin // td, nchan // fd
    out = x.copy()
    ds = x[: ncd * fd, : nsd * td].reshape(ncd, fd, nsd, td).sum((1, 3))
    for j in range(ncd):
        for sg in (1, -1):
            best, arg = _runs_bruteforce(sg * ds[j])
            if len(arg) != 1:
                return None  # ambiguous maximum: the declared tie rule decides
            s, e = arg[0]
            wn = e - s + 1
            snr2 = best * best / (wn * td * fd) if wn > wnlimit else 0.0
            if abs(snr2 - thr * thr) < 1e-3 * thr * thr:
                return None  # too close to the threshold for float rounding
            if snr2 > thr * thr:
                out[j * fd:(j + 1) * fd, s * td:(e + 1) * td] = 0.0
    return out


@pytest.mark.parametrize("td,fd,wnlimit", [(1, 1, 0), (2, 1, 1), (1, 3, 2), (4, 2, 0), (3, 2, 5)])
def test_kadanef_matches_brute_force_runs(td, fd, wnlimit):
    rng = np.random.default_rng(10 * td + fd)
    checked = 0
    for trial in range(60):
        nchan, nbin = int(rng.integers(fd, 4 * fd + 3)), int(rng.integers(td, 12 * td + 5))
        # wide-ranging integers: float sums stay exact, equal-sum runs are rare
        x = rng.integers(-1000, 1001, size=(nchan, nbin)).astype(np.float32)
        # a strong run in one channel group
        j = int(rng.integers(0, nchan))
        a = int(rng.integers(0, nbin))
        x[j, a:a + int(rng.integers(1, max(2, nbin)))] += rng.choice([-3000.0, 3000.0])
        thr = float(rng.uniform(100.0, 3000.0))
        exp = _kadanef_expected(x, td, fd, thr, wnlimit)
        if exp is None:
            continue
        got = oracle.kadanef(x, td, fd, thr, wnlimit)
        assert np.array_equal(got, exp), (trial, td, fd, wnlimit)
        checked += 1
    assert checked >= 30


def test_kadanef_planted_runs_and_limits():
    # runs at the start of their rows: the running sum restarts only below
    # zero, so leading zeros would join a run
    nchan, nbin, td, fd = 8, 64, 2, 2
    x = np.zeros((nchan, nbin), np.float32)
    x[2:4, 0:20] = 3.0     # channels 2-3 (downsampled channel 1), samples 0-19
    x[6:8, 0:4] = -5.0     # channels 6-7 (downsampled channel 3), a short negative run
    got = oracle.kadanef(x, td, fd, 1.0, 0)
    assert np.count_nonzero(got) == 0
    # the short run (2 downsampled samples) is kept when wnlimit >= 2
    got = oracle.kadanef(x, td, fd, 1.0, 2)
    assert np.all(got[2:4, 0:20] == 0) and np.all(got[6:8, 0:4] == -5.0)
    # a threshold above both runs' S/N (sqrt(120^2/40) = 19, sqrt(40^2/8) = 14) keeps everything
    got = oracle.kadanef(x, td, fd, 20.0, 0)
    assert np.array_equal(got, x)
    got = oracle.kadanef(x, td, fd, 15.0, 0)
    assert np.all(got[2:4, 0:20] == 0) and np.all(got[6:8, 0:4] == -5.0)
    # leading zeros join a run: after 20 zero samples the run starts at 0
    y = np.zeros((2, 32), np.float32)
    y[0, 20:24] = 4.0
    got = oracle.kadanef(y, 1, 1, 1.0, 10)
    assert np.count_nonzero(got) == 0


def test_wnlimit_arithmetic():
    # int(nwidth*width/tbin/td) with nwidth*width in float
    assert oracle.kadane_wnlimit(20, 1e-3, 98.304e-6, 1) == 203
    assert oracle.kadane_wnlimit(0, 5e-3, 1e-4, 2) == 0
    assert oracle.kadane_wnlimit(3, 0.1, 0.01, 3) == int(np.float32(3 * np.float32(0.1)) / 0.01 / 3)
