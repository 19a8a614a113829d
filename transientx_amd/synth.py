"""Seeded synthetic 8-bit filterbanks with dispersed pulses.

Channel frequencies are the centres of `nchans` equal channels spanning
[fbot, ftop], highest first (SIGPROC order with negative foff).  Pulse
arrival times follow the cold-plasma delay of include/candidate.h:116-119,
4.148741601e3*DM*(1/f^2 - 1/fmax^2) s, relative to the top channel.

This numpy generator is used for inputs small enough to build on the host;
bench.py builds its full-size observation on the device (gpu_synth).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

KDM = 4.148741601e3


def channel_freqs(nchans: int, fbot: float, ftop: float) -> np.ndarray:
    bw = (ftop - fbot) / nchans
    return ftop - (np.arange(nchans) + 0.5) * bw


def dmdelay(dm, fh, fl):
    return KDM * dm * (1.0 / (fl * fl) - 1.0 / (fh * fh))


@dataclass
class Pulse:
    dm: float
    t0: float        # arrival time at the top channel, s
    fwhm: float      # s
    snr: float       # approximate band-integrated boxcar S/N


@dataclass
class Observation:
    nchans: int
    fbot: float
    ftop: float
    tsamp: float
    nsamp: int
    mean: float = 128.0
    sd: float = 16.0
    seed: int = 1
    pulses: list = field(default_factory=list)
    rfi_frac: float = 0.0       # fraction of channels with persistent narrowband RFI
    rfi_amp: float = 5.0        # RFI amplitude in noise sd
    impulses: int = 0           # zero-DM broadband impulses

    @property
    def freqs(self) -> np.ndarray:
        return channel_freqs(self.nchans, self.fbot, self.ftop)


def pulse_amplitude(p: Pulse, nchans: int, tsamp: float, sd: float) -> float:
    """Per-channel peak (counts) giving roughly p.snr after dedispersion + boxcar."""
    sig = p.fwhm / 2.3548200450309493 / tsamp
    energy = np.sqrt(np.sqrt(np.pi) * max(sig, 0.5))
    return p.snr * sd / (np.sqrt(nchans) * energy)


def generate(obs: Observation, t_begin: int = 0, t_end: int | None = None) -> np.ndarray:
    """uint8 time-major block [t_end-t_begin][nchans] of the observation."""
    if t_end is None:
        t_end = obs.nsamp
    rng = np.random.default_rng([obs.seed, t_begin])
    n = t_end - t_begin
    x = rng.normal(obs.mean, obs.sd, size=(n, obs.nchans))
    f = obs.freqs
    fmax = f.max()
    t = (t_begin + np.arange(n)) * obs.tsamp
    if obs.rfi_frac > 0:
        rr = np.random.default_rng([obs.seed, 7])
        nrfi = max(1, int(round(obs.rfi_frac * obs.nchans)))
        chans = rr.choice(obs.nchans, nrfi, replace=False)
        for c in chans:
            period = rr.uniform(0.01, 1.0)
            x[:, c] += obs.rfi_amp * obs.sd * np.sin(2 * np.pi * t / period)
    if obs.impulses > 0:
        rr = np.random.default_rng([obs.seed, 11])
        for _ in range(obs.impulses):
            ti = rr.uniform(0, obs.nsamp * obs.tsamp)
            i = int(ti / obs.tsamp) - t_begin
            if 0 <= i < n:
                x[i, :] += 6 * obs.sd
    for p in obs.pulses:
        amp = pulse_amplitude(p, obs.nchans, obs.tsamp, obs.sd)
        sig = p.fwhm / 2.3548200450309493
        arr = p.t0 + dmdelay(p.dm, fmax, f)
        lo = np.searchsorted(t, arr - 5 * sig)
        hi = np.searchsorted(t, arr + 5 * sig)
        for j in range(obs.nchans):
            if hi[j] > lo[j]:
                tt = t[lo[j]:hi[j]]
                x[lo[j]:hi[j], j] += amp * np.exp(-0.5 * ((tt - arr[j]) / sig) ** 2)
    return np.clip(np.rint(x), 0, 255).astype(np.uint8)


def gpu_generate(obs: Observation, device: str = "cuda", flush_blocks: int = 0, ndump: int = 0):
    """The whole observation as a uint8 [nsamp][nchans] tensor built on the device.

    Recipe (seeded by obs.seed): Gaussian noise N(mean, sd^2); obs.rfi_frac
    of the channels carry persistent narrowband RFI, modelled as noise of
    rfi_amp times the normal amplitude; obs.impulses zero-DM broadband
    impulses of 5 sd in every channel; dispersed Gaussian pulses per
    obs.pulses.  Values are rounded and clipped to [0, 255].  Also returns
    flush_blocks blocks of plain noise (the end-of-file flush of
    src/singlepulse/dedisperse_fil.cpp:340-354, seeded here).
    """
    import torch

    C, T = obs.nchans, obs.nsamp
    f = torch.tensor(obs.freqs, dtype=torch.float64, device=device)
    fmax = float(obs.freqs.max())
    delay1 = KDM * (1.0 / (f * f) - 1.0 / (fmax * fmax))  # s per unit DM
    g = torch.Generator(device=device)
    rr = np.random.default_rng([obs.seed, 7])
    nrfi = int(round(obs.rfi_frac * C))
    rfi = torch.zeros(C, dtype=torch.float32, device=device)
    if nrfi:
        rfi[torch.tensor(rr.choice(C, nrfi, replace=False), device=device)] = 1.0
    scale = obs.sd * (1.0 + (obs.rfi_amp - 1.0) * rfi)
    ri = np.random.default_rng([obs.seed, 11])
    imp = np.sort(ri.integers(0, T, obs.impulses)) if obs.impulses else np.zeros(0, int)
    out = torch.empty((T, C), dtype=torch.uint8, device=device)
    step = 1 << 15
    for t0 in range(0, T, step):
        t1 = min(T, t0 + step)
        g.manual_seed(obs.seed * 1000003 + t0)
        x = torch.randn((t1 - t0, C), generator=g, device=device, dtype=torch.float32)
        x = x * scale + obs.mean
        for i in imp[(imp >= t0) & (imp < t1)]:
            x[i - t0] += 5 * obs.sd
        tt = (torch.arange(t0, t1, device=device, dtype=torch.float64) * obs.tsamp)
        for p in obs.pulses:
            sig = p.fwhm / 2.3548200450309493
            arr = p.t0 + p.dm * delay1
            lo, hi = float(arr.min()) - 5 * sig, float(arr.max()) + 5 * sig
            if hi < t0 * obs.tsamp or lo > t1 * obs.tsamp:
                continue
            a = int(max(t0, np.floor(lo / obs.tsamp))) - t0
            b = int(min(t1, np.ceil(hi / obs.tsamp) + 1)) - t0
            amp = pulse_amplitude(p, C, obs.tsamp, obs.sd)
            z = (tt[a:b, None] - arr[None, :]) / sig
            x[a:b] += (amp * torch.exp(-0.5 * z * z)).to(torch.float32)
        out[t0:t1] = torch.clamp(torch.round(x), 0, 255).to(torch.uint8)
    flush = None
    if flush_blocks:
        g.manual_seed(obs.seed * 1000003 + 999331)
        x = torch.randn((flush_blocks * ndump, C), generator=g, device=device) * obs.sd + obs.mean
        flush = torch.clamp(torch.round(x), 0, 255).to(torch.uint8)
    return out, flush


def burst_amplitude(p: Pulse, scatter_top: float, freqs: np.ndarray, tsamp: float, sd: float) -> float:
    """Per-channel peak (counts) of the unscattered Gaussian such that the
    burst, scattered (f^-4, scatter_top s at the top channel, area kept) and
    dedispersed at its DM, has S/N p.snr under its best boxcar of the width
    ladder 1..1024 samples: the S/N of the injected signal itself."""
    from scipy.special import erfc, erfcx

    f = np.asarray(freqs, np.float64)
    sig = p.fwhm / 2.3548200450309493 / tsamp                        # samples
    taus = scatter_top * (f / f.max()) ** -4 / tsamp                  # samples per channel
    L = int(np.ceil(10 * sig + 10 * taus.max())) + 8
    t = np.arange(-int(np.ceil(5 * sig)) - 2, L, dtype=np.float64)[None, :]
    tau = np.maximum(taus, 1e-9)[:, None]
    area = sig * np.sqrt(2 * np.pi)
    z = (sig / tau - t / sig) / np.sqrt(2.0)
    gauss = np.exp(-0.5 * (t / sig) ** 2)
    with np.errstate(over="ignore", invalid="ignore"):
        pos = area / (2 * tau) * gauss * erfcx(np.maximum(z, 0.0))
        neg = area / (2 * tau) * np.exp(np.minimum(sig * sig / (2 * tau * tau) - t / tau, 700.0)) * \
            erfc(np.minimum(z, 0.0))
    emg = np.where(z > 0, pos, neg)
    emg = np.where(tau < 0.05 * sig, gauss, emg)
    prof = emg.sum(0)
    cs = np.concatenate([[0.0], np.cumsum(prof)])
    best = 0.0
    wn = 1
    while wn <= min(1024, prof.size):
        best = max(best, (cs[wn:] - cs[:-wn]).max() / np.sqrt(f.size * wn))
        wn = max(wn + 1, int(wn * 1.2))
    return p.snr * sd / best


def gpu_bursts(obs: Observation, bursts: list, t_begin: int, t_end: int, device: str = "cuda",
               scatter: list | None = None, amps: list | None = None):
    """uint8 block [t_end - t_begin][nchans] on the device: the seeded noise
    of obs (seed obs.seed, keyed by t_begin) plus dispersed bursts with an
    optional f^-4 scattering tail.

    A burst is a Gaussian of FWHM p.fwhm arriving at p.t0 + DM delay;
    scatter[i] (s, at the top channel, scaling as (f/fmax)^-4) convolves it
    with a one-sided exponential, keeping its area (the exponentially
    modified Gaussian, in the erfcx form that stays finite).  amps[i]: the
    per-channel peak of burst i (burst_amplitude); by default
    pulse_amplitude's unscattered calibration.  Each burst is
    added only over its own window in every channel, so a block costs one
    noise draw plus a few thousand samples per burst and channel."""
    import torch

    C = obs.nchans
    n = t_end - t_begin
    g = torch.Generator(device=device)
    g.manual_seed(obs.seed * 1000003 + t_begin)
    x = torch.randn((n, C), generator=g, device=device, dtype=torch.float32)
    x.mul_(obs.sd).add_(obs.mean)
    f = torch.tensor(obs.freqs, dtype=torch.float64, device=device)
    fmax = float(obs.freqs.max())
    delay1 = KDM * (1.0 / (f * f) - 1.0 / (fmax * fmax))
    ts = obs.tsamp
    for i, p in enumerate(bursts):
        tau_top = (scatter[i] if scatter else 0.0)
        sig = p.fwhm / 2.3548200450309493
        arr = p.t0 + p.dm * delay1                                  # [C] s
        tau = tau_top * (f / fmax) ** -4                             # [C] s
        lo_t = float(arr.min()) - 5 * sig
        hi_t = float(arr.max()) + 5 * sig + 8 * float(tau.max())
        if hi_t < t_begin * ts or lo_t > t_end * ts:
            continue
        K = int(np.ceil((10 * sig + 8 * float(tau.max())) / ts)) + 3
        start = torch.floor((arr - 5 * sig) / ts).long()             # [C]
        idx = start[:, None] + torch.arange(K, device=device)[None, :]  # [C][K] absolute samples
        d = idx.double() * ts - arr[:, None]
        amp = amps[i] if amps else pulse_amplitude(p, C, ts, obs.sd)
        if tau_top <= 0.0:
            val = amp * torch.exp(-0.5 * (d / sig) ** 2)
        else:
            tj = tau[:, None].clamp_min(1e-12)
            area = amp * sig * np.sqrt(2 * np.pi)
            z = (sig / tj - d / sig) / np.sqrt(2.0)
            pos = area / (2 * tj) * torch.exp(-0.5 * (d / sig) ** 2) * torch.special.erfcx(z.clamp_min(0))
            neg = area / (2 * tj) * torch.exp(sig * sig / (2 * tj * tj) - d / tj) * torch.erfc(z.clamp_max(0))
            val = torch.where(z > 0, pos, neg)
            # narrow tails: the Gaussian itself
            val = torch.where(tj < 0.05 * sig, amp * torch.exp(-0.5 * (d / sig) ** 2), val)
        inside = (idx >= t_begin) & (idx < t_end)
        if not bool(inside.any()):
            continue
        rows = (idx - t_begin)[inside]
        cols = torch.arange(C, device=device)[:, None].expand(C, K)[inside]
        x.index_put_((rows, cols), val[inside].float(), accumulate=True)
    return torch.clamp(torch.round(x), 0, 255).to(torch.uint8)


def random_pulses(seed: int, n: int, dm_lo: float, dm_hi: float, t_lo: float, t_hi: float,
                  fwhm_lo: float, fwhm_hi: float, snr_lo: float, snr_hi: float,
                  log_dm: bool = False) -> list:
    rng = np.random.default_rng([seed, 3])
    out = []
    for _ in range(n):
        dm = (np.exp(rng.uniform(np.log(dm_lo), np.log(dm_hi))) if log_dm
              else rng.uniform(dm_lo, dm_hi))
        out.append(Pulse(dm=float(dm), t0=float(rng.uniform(t_lo, t_hi)),
                         fwhm=float(rng.uniform(fwhm_lo, fwhm_hi)),
                         snr=float(rng.uniform(snr_lo, snr_hi))))
    return out
