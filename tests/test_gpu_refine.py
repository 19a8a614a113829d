"""GPU refinement grid (tx_refine_*) against the oracle's restatement of
Candidate::dedisperse / optimize / matched_filter: the DM-time map, the S/N
and width maps are bit-identical, and so are the argmax and the derived
fields.  The oracle's tree is itself bit-identical to the reference's
include/hough_transform.h (tests/test_oracle_refine.py)."""
import numpy as np
import pytest

import oracle
from transientx_amd.refine import Refiner

pytestmark = pytest.mark.gpu

K_DM = 4.148741601e3


def say(*a):
    print(*a, flush=True)


def cutouts(seed, ncand, nchan, nbin, tbin, f, dm_true, dm_cand, masked=0.03):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(ncand, nchan, nbin)).astype(np.float32)
    for c in range(ncand):
        t0 = nbin // 2 + rng.integers(-8, 9)
        w = rng.integers(1, 6)
        amp = rng.uniform(1.0, 3.0)
        for j in range(nchan):
            d = K_DM * dm_true[c] * (1 / f[j] ** 2 - 1 / f.max() ** 2) / tbin
            k = int(np.floor(t0 + d + 0.5))
            for u in range(w):
                x[c, j, (k + u) % nbin] += amp
        x[c, rng.random(nchan) < masked] = 0.0
    return x


@pytest.mark.parametrize("nchan,nbin,ascending,dedisp,isded", [
    (64, 256, False, True, True),
    (48, 128, True, True, True),        # padded rows, ascending band
    (1024, 512, False, True, True),     # the cutout shape of the refinement benchmark
    (512, 2048, False, True, True),     # 3+ tree passes
    (96, 200, False, False, False),     # no dedispersion, absolute trials
])
def test_refine_grid_bitexact(nchan, nbin, ascending, dedisp, isded):
    import torch

    tbin, snrloss = 98.304e-6, 0.1
    f = np.linspace(1500.0, 1200.0, nchan)
    if ascending:
        f = f[::-1].copy()
    ncand = 3
    rng = np.random.default_rng(nchan + nbin)
    dm_true = rng.uniform(50, 1500, ncand)
    dm_cand = dm_true + rng.uniform(-3, 3, ncand)
    x = cutouts(nchan * 3 + nbin, ncand, nchan, nbin, tbin, f, dm_true, dm_cand)
    say("case", nchan, nbin, "ascending", ascending, "dedisperse", dedisp)
    R = Refiner(f, nbin, tbin, snrloss, dedisperse=dedisp, isdedispersed=isded, max_batch=2)
    try:
        say("ndm", R.ndm, "passes", R.npass, "widths", R.nbox)
        res, (dmt, snr, wid) = R.run(torch.from_numpy(x).cuda(), dm_cand, maps=True)
        dmt, snr, wid = dmt.cpu().numpy(), snr.cpu().numpy(), wid.cpu().numpy()
    finally:
        R.close()
    for c in range(ncand):
        xc = oracle.cand_dedisperse(x[c], f, dm_cand[c], tbin) if dedisp else x[c]
        odmt, dms, ddm, ndm = oracle.refine_optimize(xc, f, tbin, dm_cand[c], isdedispersed=isded)
        assert np.array_equal(dmt[c].view(np.uint32), odmt.view(np.uint32)), f"dmtmap cand {c}"
        S, W, j, i, s = oracle.refine_mf(odmt, snrloss)
        assert np.array_equal(snr[c].view(np.uint32), S.view(np.uint32)), f"snr map cand {c}"
        assert np.array_equal(wid[c], W), f"width map cand {c}"
        r = res[c]
        assert (r["dms"], r["ddm"], r["ndm"]) == (dms, ddm, ndm)
        assert (r["maxdmid"], r["maxtid"]) == (j, i) and r["snr"] == np.float32(s)
        assert r["wn"] == W[j, i] and r["width"] == np.float32(W[j, i] * tbin)
        if dedisp and isded:
            assert abs(r["dm_maxsnr"] - dm_true[c]) < 4 * ddm + 0.5


def test_refine_without_maps_matches_with_maps():
    import torch

    nchan, nbin, tbin = 128, 256, 98.304e-6
    f = np.linspace(1500.0, 1200.0, nchan)
    x = cutouts(9, 5, nchan, nbin, tbin, f, np.full(5, 400.0), np.full(5, 401.0))
    R = Refiner(f, nbin, tbin, 0.1, max_batch=2)   # 3 chunks of <= 2 cutouts
    try:
        xt = torch.from_numpy(x).cuda()
        a = R.run(xt, np.full(5, 401.0))
        b, _ = R.run(xt, np.full(5, 401.0), maps=True)
    finally:
        R.close()
    assert a.tobytes() == b.tobytes()
