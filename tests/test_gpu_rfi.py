"""GPU RFI cleaning (tx_rfi_zdot, tx_rfi_kadanef) against the oracle's
restatement of Candidate::zdot and Candidate::kadaneF
(src/singlepulse/candidate.cpp:710-897): bit-identical frames, in the
Candidate layout [nframes][nchan][nbin] and the DataBuffer layout
[nframes][nsamp][nchan], with zero-DM interference, narrowband bursts,
constant and empty channels, and sizes that td and fd do not divide."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def frames(seed, nf, nchan, nbin):
    rng = np.random.default_rng(seed)
    x = rng.normal(0.0, 1.0, size=(nf, nchan, nbin))
    for f in range(nf):
        # zero-DM interference with per-channel gains, a narrowband burst,
        # a constant channel and an empty one
        x[f] += rng.uniform(0.3, 2.0, size=(nchan, 1)) * (rng.normal(size=nbin) * 3)[None, :]
        j, a = int(rng.integers(2, nchan)), int(rng.integers(0, nbin // 2))
        x[f, j, a:a + nbin // 4] += rng.choice([-8.0, 8.0])
        x[f, 0] = 5.0
        x[f, 1] = 0.0
    return x.astype(np.float32)


@pytest.mark.parametrize("nchan,nbin,nf", [(64, 256, 3), (100, 333, 2), (7, 1000, 1)])
def test_zdot_bit_identical_both_layouts(nchan, nbin, nf):
    import torch

    from transientx_amd import rfi

    x = frames(nchan + nbin, nf, nchan, nbin)
    say("case zdot nchan", nchan, "nbin", nbin, "frames", nf)
    ref = np.stack([oracle.zdot(x[f]) for f in range(nf)])
    g = torch.from_numpy(x.copy()).cuda()
    rfi.zdot(g, layout="channel")
    got = g.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    t = torch.from_numpy(np.ascontiguousarray(x.transpose(0, 2, 1))).cuda()
    rfi.zdot(t, layout="time")
    got_t = t.cpu().numpy().transpose(0, 2, 1)
    assert np.array_equal(np.ascontiguousarray(got_t).view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("td,fd", [(1, 1), (2, 4), (3, 2), (8, 1)])
def test_kadanef_bit_identical_both_layouts(td, fd):
    import torch

    from transientx_amd import rfi

    nf, nchan, nbin = 3, 70, 301
    x = frames(7 * td + fd, nf, nchan, nbin)
    x = np.stack([oracle.zdot(x[f]) for f in range(nf)])  # normalised, as replot_fil has it
    wl = [0, 3, 40]
    thr = 3.0
    say("case kadaneF td", td, "fd", fd, "wnlimit", wl)
    ref = np.stack([oracle.kadanef(x[f], td, fd, thr, wl[f]) for f in range(nf)])
    assert np.count_nonzero(ref == 0) > np.count_nonzero(x == 0)  # something was zapped
    g = torch.from_numpy(x.copy()).cuda()
    rfi.kadane_f(g, td, fd, thr, wl, layout="channel")
    assert np.array_equal(g.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    t = torch.from_numpy(np.ascontiguousarray(x.transpose(0, 2, 1))).cuda()
    rfi.kadane_f(t, td, fd, thr, wl, layout="time")
    got_t = np.ascontiguousarray(t.cpu().numpy().transpose(0, 2, 1))
    assert np.array_equal(got_t.view(np.uint32), ref.view(np.uint32))


def test_cutout_cleaning_chain_matches_oracle():
    """zdot then kadaneF on a batch of cutouts, as replot_fil orders them,
    with wnlimit from Candidate's int(nwidth*width/tbin/td)."""
    import torch

    from transientx_amd import rfi

    nf, nchan, nbin, td, fd = 4, 128, 512, 4, 2
    x = frames(99, nf, nchan, nbin)
    widths = [1e-3, 2e-3, 5e-3, 1e-2]
    wl = [rfi.wnlimit(3, w, 98.304e-6, td) for w in widths]
    assert wl == [oracle.kadane_wnlimit(3, w, 98.304e-6, td) for w in widths]
    ref = np.stack([oracle.kadanef(oracle.zdot(x[f]), td, fd, 5.0, wl[f]) for f in range(nf)])
    g = torch.from_numpy(x.copy()).cuda()
    rfi.kadane_f(rfi.zdot(g), td, fd, 5.0, wl)
    assert np.array_equal(g.cpu().numpy().view(np.uint32), ref.view(np.uint32))
