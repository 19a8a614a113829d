"""CPU checks of the refinement's host side: cutout extraction, the burst
generator's scattering and S/N calibration, and the gradient refinement's
optimizer (coarse grid, ridge bisection, Newton steps in arrival time) on a
synthetic S/N with a known maximum, through a stand-in for DiffSNR."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from transientx_amd.cutout import dedispersed_cutouts, group_freqs  # noqa: E402
from transientx_amd.synth import KDM, Observation, Pulse, burst_amplitude, gpu_bursts  # noqa: E402


def test_cutout_aligns_a_dispersed_impulse():
    C, ts, T = 16, 1e-3, 4000
    f = np.linspace(1500.0, 1200.0, C)
    dm, t_top = 120.0, 1000
    x = np.full((T, C), 128, np.uint8)
    for j in range(C):
        d = KDM * dm * (1 / f[j] ** 2 - 1 / f.max() ** 2) / ts
        x[t_top + int(np.sign(d) * np.floor(abs(d) + 0.5)), j] = 200
    out = dedispersed_cutouts(torch.from_numpy(x), f, ts, [dm], [t_top], 64, fdown=4, zero_point=128.0, sd=2.0)
    o = out[0].numpy()
    assert o.shape == (4, 64)
    # every group: its 4 impulses in column 32, (200-128)*4/(2*sqrt(4)) = 72
    assert np.allclose(o[:, 32], 72.0) and np.count_nonzero(o) == 4
    assert np.allclose(group_freqs(f, 4), f.reshape(4, 4).mean(1))


def test_cutout_reads_zero_point_outside_the_data():
    f = np.linspace(1500.0, 1200.0, 8)
    x = np.full((100, 8), 130, np.uint8)
    o = dedispersed_cutouts(torch.from_numpy(x), f, 1e-3, [0.0], [5], 32, zero_point=128.0, sd=1.0)[0].numpy()
    # samples -11..-1 read the zero point, 0..20 read 130
    assert np.all(o[:, :11] == 0.0) and np.all(o[:, 11:] == 2.0)


def test_scattered_burst_keeps_area_and_calibrated_snr():
    C, ts = 64, 16e-6
    obs = Observation(nchans=C, fbot=1000, ftop=1500, tsamp=ts, nsamp=1, mean=100.0, sd=1e-6, seed=3)
    p = Pulse(dm=0.0, t0=0.004, fwhm=100e-6, snr=25.0)
    amp0 = 100.0
    clean = gpu_bursts(obs, [p], 0, 2000, device="cpu", amps=[amp0]).numpy().astype(float) - 100
    scat = gpu_bursts(obs, [p], 0, 2000, device="cpu", scatter=[200e-6], amps=[amp0]).numpy().astype(float) - 100
    # noiseless 8-bit samples round the faint end of the tail away (noise
    # would dither it): the area is kept to a few percent, and the centroid
    # moves by 85-105% of tau, 200 us at the top and (1500/1000)^4 * 200 us
    # at the bottom of the band
    a0, a1 = clean.sum(0), scat.sum(0)
    assert np.all(a1 >= 0.95 * a0) and np.all(a1 <= 1.01 * a0)
    t = np.arange(scat.shape[0])
    cen = lambda a: (a * t).sum() / a.sum()
    f = obs.freqs
    for j in (0, C - 1):
        tau = 200e-6 * (f[j] / f[0]) ** -4 / ts
        lag = cen(scat[:, j]) - cen(clean[:, j])
        assert 0.85 * tau <= lag <= 1.05 * tau, (j, lag, tau)
    # calibration: the noiseless burst's best boxcar S/N is the requested one
    sd = 12.0
    amp = burst_amplitude(p, 200e-6, obs.freqs, ts, sd)
    obs2 = Observation(nchans=C, fbot=1000, ftop=1500, tsamp=ts, nsamp=1, mean=100.0, sd=1e-6, seed=3)
    b = gpu_bursts(obs2, [p], 0, 2000, device="cpu", scatter=[200e-6], amps=[amp]).numpy().astype(float) - 100
    prof = b.sum(1)
    cs = np.concatenate([[0.0], np.cumsum(prof)])
    best = max((cs[w:] - cs[:-w]).max() / (sd * np.sqrt(C * w)) for w in range(1, 400))
    assert abs(best - 25.0) < 0.6


class _Model:
    """A stand-in for DiffSNR: S(dm, t0, W) = 10 - |dm - 1.234|*3 (piecewise
    linear in DM, as the interpolated S/N is) - (t0 - tc(dm))^2/4 - (log W -
    log 9)^2, tc(dm) = 256 + 0.3 - dm*kbar."""

    nbin = 512

    def __init__(self):
        self.kf = np.linspace(0.0, 10.0, 16)
        self.kbar = float(np.mean(self.kf))

    def _s(self, dm, t0, w):
        dm, t0, w = (np.asarray(v, np.float64) for v in (dm, t0, w))
        tc = 256.3 - (dm - 1.234) * self.kbar
        return 10 - 3 * np.abs(dm - 1.234) - (t0 - tc) ** 2 / 4 - (np.log(w) - np.log(9.0)) ** 2

    def forward(self, x, w, dm, t0, width, cand=None):
        return torch.tensor(self._s(dm.numpy(), t0.numpy(), width.numpy()), dtype=torch.float32)

    def adjoint(self, x, w, dm, t0, width, gsnr, cand=None, need_w=True, need_x=False):
        d, t, wd = (v.double().numpy() for v in (dm, t0, width))
        h = 1e-5
        g = lambda a, b: torch.tensor((self._s(*a) - self._s(*b)) / (2 * h), dtype=torch.float32)
        s = torch.tensor(self._s(d, t, wd), dtype=torch.float32)
        return (s, g((d + h, t, wd), (d - h, t, wd)), g((d, t + h, wd), (d, t - h, wd)),
                g((d, t, wd + h), (d, t, wd - h)), None, None)

    def adjoint_runs(self, x, w, dm, t0, width, gsnr, run, second=False):
        assert dm.numel() % run == 0
        out = self.adjoint(x, w, dm, t0, width, gsnr, need_w=False)[:4]
        if not second:
            return out
        d, t, wd = (v.double().numpy() for v in (dm, t0, width))
        h = 1e-3
        f = lambda a, b, c: self._s(a, b, c)
        gtt = (f(d, t + h, wd) - 2 * f(d, t, wd) + f(d, t - h, wd)) / h ** 2
        gdt = (f(d + h, t + h, wd) - f(d + h, t - h, wd) - f(d - h, t + h, wd) + f(d - h, t - h, wd)) / (4 * h * h)
        return out + (torch.tensor(gtt, dtype=torch.float32), torch.tensor(gdt, dtype=torch.float32))

    def grid(self, x, w, dm0, t00, ndm, ddm, nt0, dt, widths):
        n = dm0.numel()
        best = torch.empty((n, ndm))
        idx = torch.empty((n, ndm), dtype=torch.int32)
        wl = widths.numpy()
        for c in range(n):
            for m in range(ndm):
                dm = float(dm0[c]) + m * ddm
                s = np.array([[self._s(dm, float(t00[c]) + i * dt, wv) for wv in wl] for i in range(nt0)])
                k = int(np.argmax(s))
                best[c, m] = float(s.reshape(-1)[k])
                idx[c, m] = k
        return best, idx


def test_refine_gradient_finds_the_kink_maximum():
    from transientx_amd.diff import refine_gradient

    m = _Model()
    x = torch.zeros((2, 16, 512))
    w = torch.ones((2, 16))
    dm = torch.tensor([0.2, 2.5], dtype=torch.float32)   # starts within the +-2 window
    t0 = torch.tensor([256.0, 256.0], dtype=torch.float32)
    ladder = np.geomspace(1.0, 256.0, 16)
    adm, at0, aw, asnr, nev = refine_gradient(m, x, w, dm, t0, widths=ladder)
    # the coarse grid (9 DM x 9 t0 x 16 widths), then 10 bisection steps of
    # one evaluation (the runs adjoint's second derivatives give the Newton
    # step) per (basin, width, sub-bracket) = 2 x 3 x 8 rows, then two width
    # moves of 2 neighbours + 5 two-evaluation bisection steps each
    assert nev == 9 * 9 * 16 + 2 * 3 * 8 * 10 + 2 * (2 + 2 * 5)
    assert np.all(np.abs(adm.numpy() - 1.234) < 0.02)
    # the arrival time on the ridge at that DM, within the Newton steps' reach
    tc = 256.3 - (adm.numpy() - 1.234) * m.kbar
    assert np.all(np.abs(at0.numpy() - tc) < 0.2)
    # the ladder width nearest log 9
    assert np.all(np.abs(np.log(aw.numpy()) - np.log(9.0)) <= 0.5 * np.log(256.0) / 15 + 1e-6)
