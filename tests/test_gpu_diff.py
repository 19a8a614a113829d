"""GPU differentiable S/N (tx_refine_fwd / tx_refine_adj / tx_refine_adj_runs): the forward
against an FP64 restatement of the same function, the adjoint against FP64
central differences of that restatement and against central differences
(h = 1e-3) of the GPU forward, torch autograd through RefineSNR, and
gradient ascent against the grid judge on the same S/N."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K_DM = 4.148741601e3


def say(*a):
    print(*a, flush=True)


def snr64(x, w, dm, t0, W, kf, s):
    """The S/N of tx_refine_fwd restated in FP64 (numpy)."""
    nchan, nbin = x.shape
    tau = dm * kf
    d = np.floor(tau)
    a = tau - d
    d = np.mod(d, nbin).astype(np.int64)
    t = np.arange(nbin)
    i0 = (t[None, :] + d[:, None]) % nbin
    i1 = (i0 + 1) % nbin
    xs = (1 - a)[:, None] * np.take_along_axis(x, i0, 1) + a[:, None] * np.take_along_axis(x, i1, 1)
    p = (w[:, None] * xs).sum(0)
    dl = np.mod(t - t0 + nbin / 2, nbin) - nbin / 2
    sig = lambda z: 1 / (1 + np.exp(-z))
    g = sig((dl + W / 2) / s) - sig((dl - W / 2) / s)
    gc = g - g.mean()
    return float((gc * p).sum() / (np.linalg.norm(gc) * np.linalg.norm(w)))


def make(seed, ncut, nchan=256, nbin=256, tbin=98.304e-6, dmhi=500.0):
    rng = np.random.default_rng(seed)
    f = np.linspace(1500.0, 1200.0 + 300.0 / nchan, nchan)
    x = rng.normal(size=(ncut, nchan, nbin))
    w = np.ones((ncut, nchan))
    w[rng.random((ncut, nchan)) < 0.03] = 0.0
    dm = rng.uniform(50, dmhi, ncut)
    t0 = nbin / 2 + rng.uniform(-8, 8, ncut)
    W = rng.uniform(3, 12, ncut)
    kf = K_DM * (1 / f ** 2 - 1 / f.max() ** 2) / tbin
    tt = np.arange(nbin)
    for c in range(ncut):
        for j in range(nchan):
            ctr = t0[c] + dm[c] * kf[j]
            dl = np.mod(tt - ctr + nbin / 2, nbin) - nbin / 2
            x[c, j] += rng.uniform(0.4, 0.8) * np.exp(-0.5 * (dl / (W[c] / 2.355)) ** 2)
    return f, x.astype(np.float32), w.astype(np.float32), dm, t0, W, kf, tbin


@pytest.mark.parametrize("nchan,nbin,dmhi", [(256, 256, 500.0), (1024, 512, 1500.0)])
def test_forward_and_adjoint_match_fp64(nchan, nbin, dmhi):
    """Also at the refinement workload's cutout shape: 1024 channels x 512
    samples, DM up to 1500, 3% of the channels masked."""
    import torch

    from transientx_amd.diff import DiffSNR

    f, x, w, dm, t0, W, kf, tbin = make(1, 4, nchan=nchan, nbin=nbin, dmhi=dmhi)
    ncut, nchan, nbin = x.shape
    rng = np.random.default_rng(2)
    pdm = dm + rng.uniform(-1, 1, ncut)
    pt0 = t0 + rng.uniform(-1.5, 1.5, ncut)
    pW = W * rng.uniform(0.7, 1.3, ncut)
    say("case ncut", ncut, "nchan", nchan, "nbin", nbin)
    m = DiffSNR(f, nbin, tbin, softness=1.0)
    try:
        xt, wt = torch.from_numpy(x).cuda(), torch.from_numpy(w).cuda()
        F = lambda v: torch.tensor(v, dtype=torch.float32, device="cuda")
        snr, gdm, gt0, gwd, gw, _ = m.adjoint(xt, wt, F(pdm), F(pt0), F(pW), torch.ones(ncut, device="cuda"))
        snr, gdm, gt0, gwd, gw = (v.cpu().numpy() for v in (snr, gdm, gt0, gwd, gw))
        snr_f = m.forward(xt, wt, F(pdm), F(pt0), F(pW)).cpu().numpy()
    finally:
        m.close()
    assert np.array_equal(snr, snr_f)  # the adjoint recomputes the same forward
    xd, wd = x.astype(np.float64), w.astype(np.float64)
    for c in range(ncut):
        args = (np.float64(np.float32(pdm[c])), np.float64(np.float32(pt0[c])), np.float64(np.float32(pW[c])))
        s64 = snr64(xd[c], wd[c], *args, kf, 1.0)
        assert abs(snr[c] - s64) <= 2e-5 * max(1.0, abs(s64)), (c, snr[c], s64)
        fd = []
        for k, h in enumerate((1e-6, 1e-5, 1e-5)):
            a1, a2 = list(args), list(args)
            a1[k] += h
            a2[k] -= h
            fd.append((snr64(xd[c], wd[c], *a1, kf, 1.0) - snr64(xd[c], wd[c], *a2, kf, 1.0)) / (2 * h))
        for got, ref in zip((gdm[c], gt0[c], gwd[c]), fd):
            assert abs(got - ref) <= 2e-3 * abs(ref) + 2e-3, (c, got, ref)
        for j in (0, 17, nchan // 2, nchan - 1):
            w1, w2 = wd[c].copy(), wd[c].copy()
            w1[j] += 1e-6
            w2[j] -= 1e-6
            ref = (snr64(xd[c], w1, *args, kf, 1.0) - snr64(xd[c], w2, *args, kf, 1.0)) / 2e-6
            assert abs(gw[c, j] - ref) <= 2e-3 * abs(ref) + 2e-4, (c, j, gw[c, j], ref)


def test_adjoint_matches_gpu_central_differences():
    import torch

    from transientx_amd.diff import DiffSNR

    f, x, w, dm, t0, W, kf, tbin = make(3, 3)
    ncut, nchan, nbin = x.shape
    m = DiffSNR(f, nbin, tbin, softness=1.0)
    try:
        xt, wt = torch.from_numpy(x).cuda(), torch.from_numpy(w).cuda()
        p = [torch.tensor(v + 0.3, dtype=torch.float32, device="cuda") for v in (dm, t0, W)]
        _, gdm, gt0, gwd, _, _ = m.adjoint(xt, wt, *p, torch.ones(ncut, device="cuda"), need_w=False)
        h = 1e-3
        for k, g in enumerate((gdm, gt0, gwd)):
            pp = [q.clone() for q in p]
            pm = [q.clone() for q in p]
            pp[k] += h
            pm[k] -= h
            fd = (m.forward(xt, wt, *pp) - m.forward(xt, wt, *pm)) / (2 * h)
            g, fd = g.cpu().numpy(), fd.cpu().numpy()
            say("param", k, "adjoint", g, "fd", fd)
            assert np.all(np.abs(g - fd) <= 0.05 * np.abs(g) + 0.05)
    finally:
        m.close()


@pytest.mark.parametrize("nchan,nbin,run,ncut", [(256, 256, 30, 3), (100, 500, 5, 4), (64, 1024, 9, 2),
                                                  (48, 512, 1, 3)])
def test_adjoint_runs_matches_adjoint(nchan, nbin, run, ncut):
    """tx_refine_adj_runs (a cutout's rows staged once for up to 8 of its
    evaluations) against tx_refine_adj with cand[e] = e // run: the same S
    and gradients up to FP32 summation order, and S against the FP64
    restatement; ragged runs (not a multiple of 8), a channel count that is
    not a multiple of the row chunk, nbin below a whole number of wavefronts,
    and upstream gradients other than 1."""
    import torch

    from transientx_amd.diff import DiffSNR

    f, x, w, dm, t0, W, kf, tbin = make(11 + run, ncut, nchan=nchan, nbin=nbin)
    rng = np.random.default_rng(run)
    n = ncut * run
    c = np.arange(n) // run
    pdm = (dm[c] + rng.uniform(-1, 1, n)).astype(np.float32)
    pt0 = (t0[c] + rng.uniform(-2, 2, n)).astype(np.float32)
    pW = (W[c] * rng.uniform(0.6, 1.5, n)).astype(np.float32)
    gs = rng.uniform(0.5, 2.0, n).astype(np.float32)
    say("case nchan", nchan, "nbin", nbin, "run", run, "ncut", ncut)
    m = DiffSNR(f, nbin, tbin, softness=1.0)
    try:
        xt, wt = torch.from_numpy(x).cuda(), torch.from_numpy(w).cuda()
        F = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()
        a = m.adjoint_runs(xt, wt, F(pdm), F(pt0), F(pW), F(gs), run=run)
        b = m.adjoint(xt, wt, F(pdm), F(pt0), F(pW), F(gs), cand=F(c.astype(np.int32)), need_w=False)[:4]
        a = [v.cpu().numpy() for v in a]
        b = [v.cpu().numpy() for v in b]
    finally:
        m.close()
    for k, (u, v) in enumerate(zip(a, b)):
        tol = 2e-5 * np.maximum(1.0, np.abs(v)) if k == 0 else 1e-4 * np.abs(v) + 1e-4
        assert np.all(np.abs(u - v) <= tol), (k, np.max(np.abs(u - v)))
    xd, wd = x.astype(np.float64), w.astype(np.float64)
    for e in range(0, n, max(1, n // 5)):
        s64 = snr64(xd[c[e]], wd[c[e]], float(pdm[e]), float(pt0[e]), float(pW[e]), kf, 1.0)
        assert abs(a[0][e] - s64) <= 2e-5 * max(1.0, abs(s64)), (e, a[0][e], s64)


@pytest.mark.parametrize("nbin", [256, 1024])
def test_runs_second_derivatives_match_central_differences(nbin):
    """tx_refine_adj_runs' d2S/dt0^2 and d2S/dDM dt0 against central
    differences (h = 0.01 samples in t0) of its own dS/dt0 and dS/dDM, with
    upstream gradients other than 1."""
    import torch

    from transientx_amd.diff import DiffSNR

    run, ncut = 6, 3
    f, x, w, dm, t0, W, kf, tbin = make(31 + nbin, ncut, nchan=64, nbin=nbin)
    rng = np.random.default_rng(nbin)
    n = ncut * run
    c = np.arange(n) // run
    pdm = (dm[c] + rng.uniform(-0.5, 0.5, n)).astype(np.float32)
    pt0 = (t0[c] + rng.uniform(-3, 3, n)).astype(np.float32)
    pW = (W[c] * rng.uniform(0.7, 1.4, n)).astype(np.float32)
    gs = rng.uniform(0.5, 2.0, n).astype(np.float32)
    say("case second derivatives nbin", nbin)
    m = DiffSNR(f, nbin, tbin, softness=1.0)
    try:
        xt, wt = torch.from_numpy(x).cuda(), torch.from_numpy(w).cuda()
        F = lambda v: torch.from_numpy(np.ascontiguousarray(v, np.float32)).cuda()
        _, gd, gt, _, gtt, gdt = m.adjoint_runs(xt, wt, F(pdm), F(pt0), F(pW), F(gs), run=run, second=True)
        h = 0.01
        _, gdp, gtp, _ = m.adjoint_runs(xt, wt, F(pdm), F(pt0 + h), F(pW), F(gs), run=run)
        _, gdn, gtn, _ = m.adjoint_runs(xt, wt, F(pdm), F(pt0 - h), F(pW), F(gs), run=run)
        gtt, gdt = gtt.cpu().numpy(), gdt.cpu().numpy()
        fd_tt = ((gtp - gtn) / (2 * h)).cpu().numpy()
        fd_dt = ((gdp - gdn) / (2 * h)).cpu().numpy()
    finally:
        m.close()
    say("d2S/dt0^2", gtt[:4], fd_tt[:4])
    say("d2S/dDM dt0", gdt[:4], fd_dt[:4])
    assert np.all(np.abs(gtt - fd_tt) <= 0.02 * np.abs(fd_tt) + 2e-3)
    assert np.all(np.abs(gdt - fd_dt) <= 0.02 * np.abs(fd_dt) + 2e-3)
    assert np.max(np.abs(gtt)) > 1e-2  # the check is not vacuous


def test_autograd_and_channel_data_gradient():
    import torch

    from transientx_amd.diff import DiffSNR, refine_snr

    f, x, w, dm, t0, W, kf, tbin = make(4, 2, nchan=64, nbin=128)
    ncut, nchan, nbin = x.shape
    m = DiffSNR(f, nbin, tbin, softness=1.5)
    try:
        xt = torch.from_numpy(x).cuda().requires_grad_(True)
        wt = torch.from_numpy(w).cuda().requires_grad_(True)
        dmt = torch.tensor(dm, dtype=torch.float64, device="cuda", requires_grad=True)
        t0t = torch.tensor(t0, dtype=torch.float32, device="cuda", requires_grad=True)
        Wt = torch.tensor(W, dtype=torch.float32, device="cuda", requires_grad=True)
        S = refine_snr(m, xt, wt, dmt, t0t, Wt)
        (S * torch.tensor([1.0, -2.0], device="cuda")).sum().backward()
        assert dmt.grad.dtype == torch.float64
        _, gdm, gt0, gwd, gw, gx = m.adjoint(xt.detach(), wt.detach(), dmt.detach().float(), t0t.detach(),
                                             Wt.detach(), torch.tensor([1.0, -2.0], device="cuda"),
                                             need_x=True)
        assert torch.equal(dmt.grad.float(), gdm) and torch.equal(t0t.grad, gt0) and torch.equal(Wt.grad, gwd)
        assert torch.equal(wt.grad, gw) and torch.equal(xt.grad, gx)
        # dS/dx against FP64 central differences of the restatement on a few elements
        xd, wd = x.astype(np.float64), w.astype(np.float64)
        gxn = gx.cpu().numpy()
        for (c, j, k) in [(0, 0, 0), (0, 5, 64), (1, 63, 127), (1, 30, 70)]:
            x1, x2 = xd[c].copy(), xd[c].copy()
            x1[j, k] += 1e-6
            x2[j, k] -= 1e-6
            a = (np.float32(dm[c]), np.float32(t0[c]), np.float32(W[c]))
            ref = (snr64(x1, wd[c], *a, kf, 1.5) - snr64(x2, wd[c], *a, kf, 1.5)) / 2e-6
            ref *= 1.0 if c == 0 else -2.0
            assert abs(gxn[c, j, k] - ref) <= 2e-3 * abs(ref) + 1e-5, (c, j, k, gxn[c, j, k], ref)
    finally:
        m.close()


def test_gradient_ascent_matches_grid_with_fewer_evaluations():
    import torch

    from transientx_amd.diff import DiffSNR, ascend, grid_search

    f, x, w, dm, t0, W, kf, tbin = make(5, 6, nchan=256, nbin=256)
    ncut, nchan, nbin = x.shape
    rng = np.random.default_rng(7)
    dm_start = dm + rng.uniform(-1.5, 1.5, ncut)   # the search's coarse DM
    t0_start = np.full(ncut, nbin / 2.0)
    m = DiffSNR(f, nbin, tbin, softness=1.0)
    try:
        xt, wt = torch.from_numpy(x).cuda(), torch.from_numpy(w).cuda()
        F = lambda v: torch.tensor(v, dtype=torch.float32, device="cuda")
        adm, at0, aw, asnr, nev = ascend(m, xt, wt, F(dm_start), F(t0_start), F(np.full(ncut, 4.0)))
        gdm, gt0, gw_, gsnr, ngrid = grid_search(m, xt, wt, dm_start, t0_start, ddm=0.05, dm_half=2.0,
                                                 dt=0.25, t_half=8.0)
    finally:
        m.close()
    adm, asnr = adm.cpu().numpy(), asnr.cpu().numpy()
    say("ascent dm", adm, "grid dm", gdm, "true", dm)
    say("ascent snr", asnr, "grid snr", gsnr, "evaluations", nev, ngrid)
    assert ngrid >= 10 * nev
    assert np.all(np.abs(adm - gdm) <= 0.1)
    assert np.all(asnr >= gsnr - 0.05 * np.abs(gsnr))


def test_grid_matches_forward_brute_force():
    """tx_refine_grid: the best S per (cutout, DM trial) and its (t0, W)
    index equal a brute-force maximum over tx_refine_fwd evaluations of the
    same grid (S within FP32 rounding; the index wherever the top two S
    differ by more than that), for whole-sample and 1/4-sample t0 steps."""
    import torch

    from transientx_amd.diff import DiffSNR

    f, x, w, dm, t0, W, kf, tbin = make(11, 3, nchan=128, nbin=256)
    ncut, nchan, nbin = x.shape
    m = DiffSNR(f, nbin, tbin, softness=1.0)
    try:
        xt, wt = torch.from_numpy(x).cuda(), torch.from_numpy(w).cuda()
        widths = np.float32([1.0, 2.5, 6.0, 14.0, 33.0])
        for ndm, ddm, nt0, dt in [(7, 0.5, 9, 1.0), (5, 0.25, 13, 0.25), (3, 1.0, 5, 2.0)]:
            say("grid", ndm, ddm, nt0, dt)
            d0 = np.float32(dm - 1.0)
            tz = np.float32(t0 - 3.0)
            best, idx = m.grid(xt, wt, torch.from_numpy(d0).cuda(), torch.from_numpy(tz).cuda(), ndm,
                               ddm, nt0, dt, torch.from_numpy(widths).cuda())
            best, idx = best.cpu().numpy(), idx.cpu().numpy()
            mm, ii, kk = np.meshgrid(np.arange(ndm), np.arange(nt0), np.arange(len(widths)), indexing="ij")
            for c in range(ncut):
                dmv = (d0[c] + mm.astype(np.float32) * np.float32(ddm)).reshape(-1)
                t0v = (tz[c] + ii.astype(np.float32) * np.float32(dt)).reshape(-1)
                wv = widths[kk.reshape(-1)]
                cand = torch.full((dmv.size,), c, dtype=torch.int32, device="cuda")
                F = lambda v: torch.from_numpy(np.ascontiguousarray(v, np.float32)).cuda()
                s = m.forward(xt, wt, F(dmv), F(t0v), F(wv), cand).cpu().numpy().reshape(ndm, -1)
                for r in range(ndm):
                    top = np.sort(s[r])[::-1]
                    assert abs(best[c, r] - top[0]) <= 1e-4 * abs(top[0]) + 1e-5, (c, r, best[c, r], top[0])
                    if top[0] - top[1] > 1e-4 * abs(top[0]) + 1e-5:
                        assert idx[c, r] == int(np.argmax(s[r])), (c, r, idx[c, r], np.argmax(s[r]))
    finally:
        m.close()


def test_coarse_grid_then_ascent_matches_fine_grid_on_wide_pulses():
    """Wide pulses (S/N nearly flat in DM): the coarse-grid start plus the
    ridge bisection lands within 0.1 pc/cc of the fine grid judge (DM step
    0.05, t0 step 0.25, 16 widths) with at least 10x fewer S/N evaluations,
    on at least 90% of the cutouts; any other cutout ends on a maximum at
    least as high as the judge's (to 1e-4): a tie, whose DM the S/N does not
    fix to 0.1."""
    import torch

    from transientx_amd.diff import DiffSNR, grid_search, refine_gradient

    rng = np.random.default_rng(21)
    ncut, nchan, nbin, tbin = 12, 256, 512, 98.304e-6
    f = np.linspace(1500.0, 1200.0 + 300.0 / nchan, nchan)
    kf = K_DM * (1 / f ** 2 - 1 / f.max() ** 2) / tbin
    dm = rng.uniform(50, 1500, ncut)
    t0 = nbin / 2 + rng.integers(-8, 9, ncut)
    fw = rng.uniform(1e-3, 20e-3, ncut) / tbin
    x = rng.normal(size=(ncut, nchan, nbin))
    tt = np.arange(nbin)
    for c in range(ncut):
        ctr = t0[c] + dm[c] * kf
        dl = np.mod(tt[None, :] - ctr[:, None] + nbin / 2, nbin) - nbin / 2
        prof = np.exp(-0.5 * (dl / (fw[c] / 2.3548)) ** 2)
        x[c] += rng.uniform(8, 40) / np.sqrt(prof.sum(-1, keepdims=True) * nchan) * prof
    w = (rng.random((ncut, nchan)) >= 0.03).astype(np.float32)
    x = (x * w[:, :, None]).astype(np.float32)
    dm_start = dm + rng.uniform(-1.5, 1.5, ncut)
    m = DiffSNR(f, nbin, tbin, softness=1.0)
    try:
        xt, wt = torch.from_numpy(x).cuda(), torch.from_numpy(w).cuda()
        F = lambda v: torch.tensor(v, dtype=torch.float32, device="cuda")
        adm, at0, aw, asnr, nev = refine_gradient(m, xt, wt, F(dm_start), F(np.full(ncut, nbin / 2.0)))
        gdm, gt0, gw_, gsnr, ngrid = grid_search(m, xt, wt, dm_start, np.full(ncut, nbin / 2.0), ddm=0.05,
                                                 dm_half=2.0, dt=0.25, t_half=8.0,
                                                 widths=np.geomspace(1.0, 256.0, 16))
    finally:
        m.close()
    adm, asnr = adm.cpu().numpy(), asnr.cpu().numpy()
    say("ascent dm", np.round(adm, 3), "grid dm", np.round(gdm, 3), "true", np.round(dm, 3))
    say("ascent snr", np.round(asnr, 3), "grid snr", np.round(gsnr, 3), "evaluations", nev, ngrid)
    assert ngrid >= 10 * nev
    close = np.abs(adm - gdm) <= 0.1
    assert np.mean(close) >= 0.9
    assert np.all(close | (asnr >= gsnr * (1 - 1e-4)))


def refinement_cutouts(ncand, seed=4):
    """The refinement workload: float32 cutouts of 1024 channels x 512
    samples (1200-1500 MHz, 98.304 us), unit-variance noise, 3% of channels
    zeroed, one dispersed pulse each (DM U[50,1500], arrival +-8 samples
    from the centre, FWHM 1-20 ms, S/N 8-40), a start DM off by up to 2."""
    import torch

    nchan, nbin, tbin = 1024, 512, 98.304e-6
    f = np.linspace(1500.0, 1200.0 + 300.0 / nchan, nchan)
    g = torch.Generator(device="cuda").manual_seed(seed)
    rng = np.random.default_rng(seed)
    dm_true = rng.uniform(50, 1500, ncand)
    dm_cand = dm_true + rng.uniform(-2, 2, ncand)
    t0 = nbin // 2 + rng.integers(-8, 9, ncand)
    fwhm = rng.uniform(1e-3, 20e-3, ncand) / tbin
    snr = rng.uniform(8, 40, ncand)
    x = torch.randn((ncand, nchan, nbin), generator=g, device="cuda", dtype=torch.float32)
    dl = torch.tensor(K_DM * (1 / f ** 2 - 1 / f.max() ** 2) / tbin, device="cuda")
    tt = torch.arange(nbin, device="cuda", dtype=torch.float64)
    for c0 in range(0, ncand, 256):
        c1 = min(ncand, c0 + 256)
        ctr = torch.tensor(t0[c0:c1], device="cuda")[:, None] + \
            torch.tensor(dm_true[c0:c1], device="cuda")[:, None] * dl[None, :]
        sig = torch.tensor(fwhm[c0:c1] / 2.3548, device="cuda")[:, None, None]
        d = (tt[None, None, :] - ctr[:, :, None] + nbin / 2) % nbin - nbin / 2
        prof = torch.exp(-0.5 * (d / sig) ** 2)
        amp = torch.tensor(snr[c0:c1], device="cuda")[:, None, None] / torch.sqrt(prof.sum(-1, keepdim=True) * nchan)
        x[c0:c1] += (amp * prof).float()
    mask = torch.rand((ncand, nchan), generator=g, device="cuda") < 0.03
    x[mask] = 0.0
    return f, nbin, tbin, x, (~mask).float().contiguous(), dm_cand, dm_true


def test_gradient_refinement_at_the_refinement_workload():
    """4096 cutouts of 1024 x 512 (the refinement workload): gradient
    refinement lands within 0.1 pc/cc of the grid judge (DM step 0.05, t0
    step 0.25 samples, 16 widths, 84,240 S/N evaluations) on at least 99% of
    the cutouts, the others on a maximum at least as high as the judge's
    (to 1e-4: a tie, whose DM the S/N does not fix to 0.1), with at least
    10x fewer S/N evaluations."""
    import torch

    from transientx_amd.diff import DiffSNR, grid_search, refine_gradient

    f, nbin, tbin, x, w, dm_cand, dm_true = refinement_cutouts(4096)
    ncut = x.shape[0]
    say("case refinement workload", ncut, "cutouts of", tuple(x.shape[1:]))
    m = DiffSNR(f, nbin, tbin, softness=1.0)
    try:
        F = lambda v: torch.tensor(v, dtype=torch.float32, device="cuda")
        adm, at0, aw, asnr, nev = refine_gradient(m, x, w, F(dm_cand), F(np.full(ncut, nbin / 2.0)))
        jdm, jt0, jw, jsnr, njev = grid_search(m, x, w, dm_cand, np.full(ncut, nbin / 2.0), ddm=0.05, dm_half=2.0,
                                               dt=0.25, t_half=8.0, widths=np.geomspace(1.0, 256.0, 16))
    finally:
        m.close()
    adm, asnr = adm.cpu().numpy(), asnr.cpu().numpy()
    close = np.abs(adm - jdm) <= 0.1
    tie = asnr >= jsnr * (1 - 1e-4)
    say("within 0.1", close.mean(), "within 0.1 or tie", (close | tie).mean(), "evaluations", nev, njev,
        "median |dDM| vs truth", float(np.median(np.abs(adm - dm_true))))
    assert njev >= 10 * nev
    assert np.mean(close | tie) >= 0.99
    assert np.mean(close) >= 0.95
