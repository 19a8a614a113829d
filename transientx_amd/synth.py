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
