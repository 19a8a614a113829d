"""Differentiable candidate S/N and gradient refinement (include/txgpu.h
tx_diff_* / tx_refine_fwd / tx_refine_adj).

The S/N of a soft-edged boxcar (arrival t0, width W, samples) on the
fractional-delay (linear interpolation) dedispersed profile of a cutout at a
DM, with channel weights w.  :class:`RefineSNR` exposes it to torch autograd
through the hand-written adjoint kernel; :func:`ascend` refines (DM, t0, W)
by gradient ascent, the replot's grid search (the refinement grid of
transientx_amd.refine, src/singlepulse/candidate.cpp:899-1087) being the
fallback and judge.  :func:`refine_gradient` starts the ascent from a coarse
grid of the same S/N; :func:`grid_search` is the fine grid judge.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class DiffSNR:
    """S/N model for cutouts of one shape [nchan][nbin]."""

    def __init__(self, freqs, nbin: int, tbin: float, softness: float = 1.0, device: int = 0):
        self._f = np.ascontiguousarray(freqs, np.float64)
        self.nchan, self.nbin, self.tbin = self._f.size, int(nbin), float(tbin)
        d = _lib.DiffDesc(self.nchan, self.nbin, self._f.ctypes.data_as(C.POINTER(C.c_double)),
                          self.tbin, softness, device)
        h = C.c_void_p()
        _lib.check(_lib.lib().tx_diff_create(C.byref(d), C.byref(h)))
        self.h = h
        fmax = self._f.max()
        # samples of delay per unit DM, channel by channel
        self.kf = 4.148741601e3 * (1 / self._f ** 2 - 1 / fmax ** 2) / self.tbin

    @staticmethod
    def _p(t):
        return t.data_ptr() if t is not None else None

    def forward(self, x, w, dm, t0, width, cand=None):
        """S/N of neval evaluations (device tensors; evaluation e reads cutout
        cand[e], or e)."""
        import torch

        n = dm.numel()
        snr = torch.empty(n, dtype=torch.float32, device=dm.device)
        torch.cuda.synchronize(dm.device)
        _lib.check(_lib.lib().tx_refine_fwd(self.h, n, self._p(cand), x.data_ptr(), w.data_ptr(),
                                            dm.data_ptr(), t0.data_ptr(), width.data_ptr(),
                                            snr.data_ptr()))
        return snr

    def adjoint(self, x, w, dm, t0, width, gsnr, cand=None, need_w=True, need_x=False):
        import torch

        n = dm.numel()
        dev = dm.device
        f32 = dict(dtype=torch.float32, device=dev)
        snr, gdm, gt0, gwd = (torch.empty(n, **f32) for _ in range(4))
        gw = torch.empty((n, self.nchan), **f32) if need_w else None
        gx = torch.empty((n, self.nchan, self.nbin), **f32) if need_x else None
        torch.cuda.synchronize(dev)
        _lib.check(_lib.lib().tx_refine_adj(
            self.h, n, self._p(cand), x.data_ptr(), w.data_ptr(), dm.data_ptr(), t0.data_ptr(),
            width.data_ptr(), gsnr.data_ptr(), snr.data_ptr(), gdm.data_ptr(), gt0.data_ptr(),
            gwd.data_ptr(), self._p(gw), self._p(gx)))
        return snr, gdm, gt0, gwd, gw, gx

    def adjoint_runs(self, x, w, dm, t0, width, gsnr, run: int, second: bool = False):
        """The adjoint without dS/dw and dS/dx for evaluations in runs of
        `run` per cutout (evaluation e reads cutout e // run): each
        workgroup stages a cutout's rows once for up to 8 of its evaluations
        (tx_refine_adj_runs).  Returns (snr, gdm, gt0, gwidth), and with
        second=True also d2S/dt0^2 and d2S/dDM dt0 (upstream-scaled)."""
        import torch

        n = dm.numel()
        if n % run:
            raise ValueError("the evaluation count must be a whole number of runs")
        dev = dm.device
        f32 = dict(dtype=torch.float32, device=dev)
        snr, gdm, gt0, gwd = (torch.empty(n, **f32) for _ in range(4))
        gtt = gdt = None
        if second:
            gtt, gdt = torch.empty(n, **f32), torch.empty(n, **f32)
        torch.cuda.synchronize(dev)
        _lib.check(_lib.lib().tx_refine_adj_runs(
            self.h, n // run, run, x.data_ptr(), w.data_ptr(), dm.data_ptr(), t0.data_ptr(),
            width.data_ptr(), gsnr.data_ptr(), snr.data_ptr(), gdm.data_ptr(), gt0.data_ptr(),
            gwd.data_ptr(), self._p(gtt), self._p(gdt)))
        if second:
            return snr, gdm, gt0, gwd, gtt, gdt
        return snr, gdm, gt0, gwd

    def grid(self, x, w, dm0, t00, ndm: int, ddm: float, nt0: int, dt: float, widths):
        """Highest S per (cutout, DM trial) over the (t0, W) grid: DM trials
        dm0[c] + m*ddm, t0 trials t00[c] + i*dt, widths [nwidth] (device
        tensors, float32).  Returns (best [ncut][ndm], idx [ncut][ndm] =
        i*nwidth + k of the first maximum)."""
        import torch

        ncut = x.shape[0]
        dev = x.device
        best = torch.empty((ncut, ndm), dtype=torch.float32, device=dev)
        idx = torch.empty((ncut, ndm), dtype=torch.int32, device=dev)
        torch.cuda.synchronize(dev)
        _lib.check(_lib.lib().tx_refine_grid(self.h, ncut, x.data_ptr(), w.data_ptr(), dm0.data_ptr(),
                                             t00.data_ptr(), ndm, ddm, nt0, dt, widths.numel(),
                                             widths.data_ptr(), best.data_ptr(), idx.data_ptr()))
        return best, idx

    def close(self):
        if self.h:
            _lib.lib().tx_diff_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _function():
    import torch

    class RefineSNR(torch.autograd.Function):
        """S = RefineSNR.apply(model, x, w, dm, t0, width): one S/N per
        cutout; backward runs the adjoint kernel (gradients for x, w, dm, t0
        and width, x only when it requires grad)."""

        @staticmethod
        def forward(ctx, model, x, w, dm, t0, width):
            ctx.dtypes = (x.dtype, w.dtype, dm.dtype, t0.dtype, width.dtype)
            x, w = x.contiguous().float(), w.contiguous().float()
            dm, t0, width = (v.contiguous().float() for v in (dm, t0, width))
            ctx.model = model
            ctx.save_for_backward(x, w, dm, t0, width)
            return model.forward(x, w, dm, t0, width)

        @staticmethod
        def backward(ctx, gS):
            x, w, dm, t0, width = ctx.saved_tensors
            m = ctx.model
            _, gdm, gt0, gwd, gw, gx = m.adjoint(
                x, w, dm, t0, width, gS.contiguous().float(), need_w=ctx.needs_input_grad[2],
                need_x=ctx.needs_input_grad[1])
            grads = [gx, gw, gdm, gt0, gwd]
            return (None,) + tuple(g.to(dt) if g is not None else None
                                   for g, dt in zip(grads, ctx.dtypes))

    return RefineSNR


_RefineSNR = None


def refine_snr(model: DiffSNR, x, w, dm, t0, width):
    """Differentiable S/N (torch autograd)."""
    global _RefineSNR
    if _RefineSNR is None:
        _RefineSNR = _function()
    return _RefineSNR.apply(model, x, w, dm, t0, width)


def _adam(model, x, w, th, steps, lr, dmu, kbar, cand=None, box=None):
    """Adam-style ascent on S over th = (DM/dmu, tc, log W), tc = t0 +
    DM*kbar; one adjoint evaluation per step, step sizes decaying on a cosine
    to 5% (a zero learning rate holds a coordinate).  box = (dm_lo, dm_hi,
    t0_lo, t0_hi) per row keeps DM and t0 inside a window.  Returns the best
    point seen and its S."""
    import torch

    m = torch.zeros_like(th)
    v = torch.zeros_like(th)
    lrv = torch.tensor(lr, dtype=torch.float64, device=th.device)
    best = None
    best_th = th.clone()

    def params(t):
        d = t[:, 0] * dmu
        return d.float(), (t[:, 1] - d * kbar).float(), torch.exp(t[:, 2]).float()

    def project(t):
        if box is None:
            return t
        d = (t[:, 0] * dmu).clamp(box[0], box[1])
        t0 = (t[:, 1] - d * kbar).clamp(box[2], box[3])
        return torch.stack([d / dmu, t0 + d * kbar, t[:, 2]], 1)

    th = project(th)
    for k in range(1, steps + 1):
        dmk, t0k, wk = params(th)
        ones = torch.ones_like(dmk)
        snr, gdm, gt0, gwd, _, _ = model.adjoint(x, w, dmk, t0k, wk, ones, cand=cand, need_w=False)
        gdm, gt0 = gdm.double(), gt0.double()
        g = torch.stack([(gdm - gt0 * kbar) * dmu, gt0, gwd.double() * wk.double()], 1)
        if best is None:
            best = snr.clone()
            best_th = th.clone()
        else:
            better = snr > best
            best = torch.where(better, snr, best)
            best_th = torch.where(better[:, None], th, best_th)
        m = 0.8 * m + 0.2 * g
        v = 0.95 * v + 0.05 * g * g
        mh = m / (1 - 0.8 ** k)
        vh = v / (1 - 0.95 ** k)
        decay = 0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * (k - 1) / max(1, steps - 1)))
        th = th + decay * lrv * mh / (torch.sqrt(vh) + 1e-9)
        th[:, 2] = th[:, 2].clamp(0.0, float(np.log(model.nbin / 2)))
        th = project(th)
    dmb, t0b, wb = params(best_th)
    return dmb, t0b, wb, best


def _bisect(model, x, w, lo, hi, tc, logw, iters, kbar, cand=None, t0box=None, h=0.25, tmove=2.0,
            run=None):
    """Bisection on the sign of dS/dDM along the ridge of S in DM, at fixed
    width.  The fractional delay is a linear interpolation, so S is piecewise
    linear in DM (a kink wherever a channel's delay crosses a whole sample):
    its maximum sits at a kink, where Newton and decaying gradient steps
    dither.  S is smooth in the arrival time, so each step also takes one
    Newton step in tc = t0 + DM*kbar (the ridge coordinate: dS/dDM at fixed
    tc is the slope along the ridge).  One batched adjoint call per step
    evaluates (mid, tc) and (mid, tc + h); 2 evaluations per step.  Every
    evaluated point counts towards the best S returned.  run: rows come in
    runs of `run` per cutout (row r reads cutout r // run), which lets the
    adjoint share each cutout's rows among its evaluations and return the
    second derivatives in t0: then one evaluation per step, at (mid, tc),
    gives the Newton step and the slope's change along it."""
    import torch

    n = lo.shape[0]
    dev = lo.device
    c2 = None if cand is None else cand[:, None].expand(n, 2).reshape(-1).contiguous()
    best = torch.full((n,), -float("inf"), dtype=torch.float32, device=dev)
    bd, bt = lo.clone(), tc - lo * kbar
    wv = torch.exp(logw).float()
    w2 = wv[:, None].expand(n, 2).reshape(-1).contiguous()

    def clamp_tc(d, t):
        if t0box is None:
            return t
        return (t - d * kbar).clamp(t0box[0], t0box[1]) + d * kbar

    ones = torch.ones(n, dtype=torch.float32, device=dev)
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        tc = clamp_tc(mid, tc)
        if run is not None:
            t0v = tc - mid * kbar
            snr, gdm, gt0, _, gtt, gdt = model.adjoint_runs(
                x, w, mid.float().contiguous(), t0v.float().contiguous(), wv.contiguous(), ones, run=run,
                second=True)
            better = snr > best
            best = torch.where(better, snr, best)
            bd = torch.where(better, mid, bd)
            bt = torch.where(better, t0v, bt)
            gd, gt, curv, cdt = gdm.double(), gt0.double(), gtt.double(), gdt.double()
            newton = -gt / torch.where(curv < 0, curv, -torch.ones_like(curv))
            move = torch.where(curv < 0, newton, torch.sign(gt) * 0.5).clamp(-tmove, tmove)
            # the ridge slope (dS/dDM at fixed tc) carried to the stepped tc
            slope = (gd - gt * kbar) + (cdt - curv * kbar) * move
            up = slope > 0
            lo = torch.where(up, mid, lo)
            hi = torch.where(up, hi, mid)
            tc = tc + move
            continue
        tcs = torch.stack([tc, tc + h], 1)
        dm2 = mid[:, None].expand(n, 2)
        t02 = tcs - dm2 * kbar
        snr, gdm, gt0, _, _, _ = model.adjoint(x, w, dm2.reshape(-1).float().contiguous(),
                                               t02.reshape(-1).float().contiguous(), w2,
                                               torch.ones(2 * n, dtype=torch.float32, device=dev),
                                               cand=c2, need_w=False)
        sv, gd, gt = snr.view(n, 2), gdm.double().view(n, 2), gt0.double().view(n, 2)
        for j in range(2):
            better = sv[:, j] > best
            best = torch.where(better, sv[:, j], best)
            bd = torch.where(better, mid, bd)
            bt = torch.where(better, t02[:, j], bt)
        # Newton step in tc from the two arrival-time slopes, and the DM slope
        # at fixed tc (gdm - gt0*kbar) carried to the stepped tc along its own
        # tc derivative: the ridge slope, free of the arrival-time error
        curv = (gt[:, 1] - gt[:, 0]) / h
        newton = -gt[:, 0] / torch.where(curv < 0, curv, -torch.ones_like(curv))
        move = torch.where(curv < 0, newton, torch.sign(gt[:, 0]) * 0.5).clamp(-tmove, tmove)
        s0 = gd[:, 0] - gt[:, 0] * kbar
        s1 = gd[:, 1] - gt[:, 1] * kbar
        slope = s0 + (s1 - s0) / h * move
        up = slope > 0
        lo = torch.where(up, mid, lo)
        hi = torch.where(up, hi, mid)
        tc = tc + move
    return bd.float(), bt.float(), wv, best


def _units(model):
    # DM change that moves the band edge by one sample; the band's mean delay per unit DM
    return float(1.0 / np.max(model.kf)), float(np.mean(model.kf))


def ascend(model: DiffSNR, x, w, dm, t0, width, steps: int = 20, lr=(2.0, 1.0, 0.25),
           dm_span: float = 1.5, tc_span: float = 6.0):
    """Gradient ascent on S/N over (DM, t0, log W) for all cutouts at once,
    from the best of a 5 x 3 stencil around the start.

    DM and t0 are strongly coupled (t0 is the arrival at the top of the
    band, so a DM change slides the dedispersed pulse), so the ascent moves
    in (DM, tc, log W) with tc = t0 + DM*kbar the arrival at the band's mean
    delay: DM in units of one sample of delay across the band (1/max kf),
    tc in samples, the width in log samples.  The stencil is DM +-dm_span,
    tc +-tc_span (one batched forward call).  Returns (dm, t0, width, snr,
    S/N evaluations per cutout)."""
    import torch

    dmu, kbar = _units(model)
    ncut = dm.numel()
    dev = dm.device
    tc = t0.double() + dm.double() * kbar
    od = torch.tensor([-1.0, -0.5, 0.0, 0.5, 1.0], dtype=torch.float64, device=dev) * dm_span
    ot = torch.tensor([-1.0, 0.0, 1.0], dtype=torch.float64, device=dev) * tc_span
    sd = (dm.double()[:, None, None] + od[None, :, None]).expand(ncut, 5, 3)
    st = (tc[:, None, None] + ot[None, None, :]).expand(ncut, 5, 3)
    sdm = sd.reshape(-1)
    st0 = (st.reshape(-1) - sdm * kbar)
    sw = width.double()[:, None].expand(ncut, 15).reshape(-1)
    cand = torch.arange(ncut, device=dev, dtype=torch.int32)[:, None].expand(ncut, 15).reshape(-1).contiguous()
    s0 = model.forward(x, w, sdm.float().contiguous(), st0.float().contiguous(), sw.float().contiguous(),
                       cand).view(ncut, 15)
    pick = torch.argmax(s0, 1)
    sdm, st = sd.reshape(ncut, 15), st.reshape(ncut, 15)
    dm0 = sdm.gather(1, pick[:, None])[:, 0]
    tc0 = st.gather(1, pick[:, None])[:, 0]
    th = torch.stack([dm0 / dmu, tc0, torch.log(width.double())], 1)
    dmb, t0b, wb, best = _adam(model, x, w, th, steps, lr, dmu, kbar)
    return dmb, t0b, wb, best, steps + 15


def _basins(best, k: int, sep: int):
    """Up to k DM trials per cutout, strongest first, each at least sep
    trials from the ones before (best: [ncut][ndm] S per DM trial)."""
    import torch

    b = best.clone()
    ndm = b.shape[1]
    picks = []
    ar = torch.arange(ndm, device=b.device)[None, :]
    for _ in range(k):
        m = torch.argmax(b, 1)
        picks.append(m)
        b = torch.where((ar - m[:, None]).abs() < sep, torch.full_like(b, -float("inf")), b)
    return torch.stack(picks, 1)  # [ncut][k]


def refine_gradient(model: DiffSNR, x, w, dm, t0, steps: int = 20, dm_half: float = 2.0,
                    ddm: float = 0.5, t_half: float = 8.0, dt: float = 2.0, widths=None,
                    wstride: int = 1, starts: int = 2, nsplit: int = 8, wsteps: int = 2):
    """Candidate refinement by gradient ascent, inside the grid search's
    window (DM +-dm_half, t0 +-t_half) and on its width ladder.

    1. A coarse grid of the same S/N (DM step ddm, t0 step dt, every
       wstride-th ladder width; each DM trial's profile formed once,
       tx_refine_grid) locates up to `starts` basins in DM (the strongest
       DM trials at least two steps apart) and their widths.
    2. In each basin, at the basin's width and its two ladder neighbours,
       `steps` adjoint evaluations of :func:`_bisect` (DM bisection on the
       sign of the ridge slope dS/dDM at fixed arrival time, Newton steps in
       the arrival time from d2S/dt0^2) over `nsplit` sub-brackets of one
       coarse step either side; all rows in one batched call per step.
    3. Ascent in width: from the best point, the ladder neighbours' S at the
       same DM and arrival time (their dS/dW sign is the direction); while
       one is higher, move there and bisect again in a bracket of a quarter
       coarse step (`wsteps` times at most).

    S/N is piecewise linear in DM (a kink wherever a channel's delay crosses
    a whole sample) and nearly flat for wide pulses, which is why DM is
    bisected on the slope's sign rather than followed by gradient steps.
    Returns (dm, t0, width, snr, S/N evaluations per cutout)."""
    import torch

    dev = x.device
    if widths is None:
        widths = np.geomspace(1.0, model.nbin / 2, 16)
    wl = np.asarray(widths, np.float32)
    nw = wl.size
    wt = torch.tensor(wl, device=dev)
    sub = np.arange(0, nw, wstride)
    if sub[-1] != nw - 1:
        sub = np.append(sub, nw - 1)
    wsub = torch.tensor(wl[sub], device=dev)
    subk = torch.tensor(sub, device=dev, dtype=torch.long)
    nws = sub.size
    ndm = int(round(2 * dm_half / ddm)) + 1
    nt0 = int(round(2 * t_half / dt)) + 1
    ncut = dm.numel()
    d0 = (dm.float() - dm_half).contiguous()
    tz = (t0.float() - t_half).contiguous()
    best, idx = model.grid(x, w, d0, tz, ndm, ddm, nt0, dt, wsub)
    ks = min(starts, ndm)
    mi = _basins(best, ks, 2)                                   # [ncut][ks] DM trials
    ii = idx.gather(1, mi).long()                                # t0*nws + k of each
    dmg = d0.double()[:, None] + mi.double() * ddm              # [ncut][ks]
    t0g = tz.double()[:, None] + (ii // nws).double() * dt
    kw = subk[ii % nws]                                          # full-ladder width index
    dmu, kbar = _units(model)
    # rows per cutout: ks basins x 3 widths x nsplit sub-brackets
    nr = ks * 3 * nsplit
    kk = torch.stack([(kw - 1).clamp(0, nw - 1), kw, (kw + 1).clamp(0, nw - 1)], 2)   # [ncut][ks][3]
    kk = kk[:, :, :, None].expand(ncut, ks, 3, nsplit).reshape(-1)
    lo0 = torch.maximum(dmg - ddm, d0.double()[:, None])
    hi0 = torch.minimum(dmg + ddm, d0.double()[:, None] + (ndm - 1) * ddm)
    frac = torch.arange(nsplit, device=dev, dtype=torch.float64) / nsplit
    span = (hi0 - lo0)[:, :, None]
    lo = (lo0[:, :, None] + span * frac)[:, :, None, :].expand(ncut, ks, 3, nsplit).reshape(-1)
    hi = (lo0[:, :, None] + span * (frac + 1.0 / nsplit))[:, :, None, :].expand(ncut, ks, 3, nsplit).reshape(-1)
    tc0 = (t0g + dmg * kbar)[:, :, None, None].expand(ncut, ks, 3, nsplit).reshape(-1)
    rep = lambda v, n: v[:, None].expand(ncut, n).reshape(-1)
    cand = torch.arange(ncut, device=dev, dtype=torch.int32)[:, None].expand(ncut, nr).reshape(-1).contiguous()
    t0box = (rep(tz.double(), nr), rep(tz.double() + (nt0 - 1) * dt, nr))
    iters = max(1, steps // 2)
    # rows come in runs of nr per cutout: the adjoint shares each cutout's
    # rows among its evaluations where the cutout shape allows
    runs = model.nbin <= 1024 and model.nbin % 4 == 0 and x.data_ptr() % 16 == 0
    dmb, t0b, wb, snr = _bisect(model, x, w, lo, hi, tc0, torch.log(wt.double()[kk]), iters, kbar, cand=cand,
                                t0box=t0box, run=nr if runs else None)
    per_step = 1 if runs else 2
    nev = ndm * nt0 * nws + nr * per_step * iters
    pick = torch.argmax(snr.view(ncut, nr), 1)
    sel = lambda v: v.view(ncut, nr).gather(1, pick[:, None])[:, 0]
    bdm, bt0, bw, bsnr = sel(dmb).double(), sel(t0b).double(), sel(wb).double(), sel(snr)
    bk = sel(kk.view(-1)).long()
    # width ascent on the ladder
    cand2 = torch.arange(ncut, device=dev, dtype=torch.int32)[:, None].expand(ncut, 2).reshape(-1).contiguous()
    for _ in range(wsteps):
        nb = torch.stack([(bk - 1).clamp(0, nw - 1), (bk + 1).clamp(0, nw - 1)], 1)        # [ncut][2]
        s2 = model.forward(x, w, bdm[:, None].expand(ncut, 2).reshape(-1).float().contiguous(),
                           bt0[:, None].expand(ncut, 2).reshape(-1).float().contiguous(),
                           wt[nb.reshape(-1)].contiguous(), cand2).view(ncut, 2)
        nev += 2
        j = torch.argmax(s2, 1)
        up = s2.gather(1, j[:, None])[:, 0] > bsnr
        nk = torch.where(up, nb.gather(1, j[:, None])[:, 0], bk)
        q = 0.25 * ddm
        lo2 = torch.maximum(bdm - q, d0.double())
        hi2 = torch.minimum(bdm + q, d0.double() + (ndm - 1) * ddm)
        it2 = max(1, iters // 2)
        dm2, t02, w2, sn2 = _bisect(model, x, w, lo2, hi2, bt0 + bdm * kbar, torch.log(wt.double()[nk]), it2, kbar,
                                    cand=torch.arange(ncut, device=dev, dtype=torch.int32).contiguous(),
                                    t0box=(tz.double(), tz.double() + (nt0 - 1) * dt))
        nev += 2 * it2
        better = up & (sn2 > bsnr)
        bdm = torch.where(better, dm2.double(), bdm)
        bt0 = torch.where(better, t02.double(), bt0)
        bw = torch.where(better, w2.double(), bw)
        bsnr = torch.where(better, sn2, bsnr)
        bk = torch.where(better, nk, bk)
    return bdm.float(), bt0.float(), bw.float(), bsnr, nev


def grid_search(model: DiffSNR, x, w, dm0, t0c, ddm=0.05, dm_half=2.0, dt=0.25, t_half=8.0,
                widths=None):
    """The grid judge on the same S/N: DM dm0 +- dm_half step ddm, t0 t0c +-
    t_half step dt, 16 geometric widths; per cutout, first maximum in (DM,
    t0, W) order (tx_refine_grid).  Returns (dm, t0, width, snr, evaluations
    per cutout) as numpy arrays."""
    import torch

    dev = x.device
    if widths is None:
        widths = np.geomspace(1.0, model.nbin / 8, 16)
    wt = torch.tensor(np.asarray(widths, np.float32), device=dev)
    nd = int(round(2 * dm_half / ddm)) + 1
    nt = int(round(2 * t_half / dt)) + 1
    d0 = torch.tensor(np.asarray(dm0, np.float32) - np.float32(dm_half), device=dev)
    tz = torch.tensor(np.asarray(t0c, np.float32) - np.float32(t_half), device=dev)
    best, idx = model.grid(x, w, d0, tz, nd, ddm, nt, dt, wt)
    mi = torch.argmax(best, 1)
    ii = idx.gather(1, mi[:, None])[:, 0].long()
    dmv = (d0 + mi.float() * np.float32(ddm)).cpu().numpy().astype(np.float64)
    t0v = (tz + (ii // wt.numel()).float() * np.float32(dt)).cpu().numpy().astype(np.float64)
    wv = wt[ii % wt.numel()].cpu().numpy().astype(np.float64)
    sv = best.gather(1, mi[:, None])[:, 0].cpu().numpy().astype(np.float64)
    return dmv, t0v, wv, sv, nd * nt * wt.numel()
