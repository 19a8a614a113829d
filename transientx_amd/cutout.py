"""Candidate cutouts for refinement: the filterbank around a candidate,
dedispersed at its search DM and downsampled in frequency, as replot_fil
prepares it before Candidate::optimize (src/singlepulse/replot_fil.cpp:376-394:
the captured samples, Candidate::dedisperse of src/singlepulse/candidate.cpp:108-140,
then downsample).

Channel j of a cutout is the 8-bit input at samples
t_top + round(delay_j(dm)/tsamp) - nbin/2 + k, k < nbin, with delay_j the
cold-plasma delay of include/candidate.h:116-119 relative to the top channel
and round half away from zero; the zero point is subtracted and groups of
`fdown` adjacent channels are summed and scaled by 1/(sd*sqrt(fdown)), so
noise is unit-variance.  Samples outside the data read as the zero point.
"""
from __future__ import annotations

import numpy as np

KDM = 4.148741601e3


def group_freqs(freqs: np.ndarray, fdown: int) -> np.ndarray:
    """Mean frequency of each group of fdown adjacent channels."""
    f = np.asarray(freqs, np.float64)
    return f[: f.size // fdown * fdown].reshape(-1, fdown).mean(1)


def dedispersed_cutouts(data, freqs, tsamp: float, dm, t_top, nbin: int, fdown: int = 1,
                        zero_point: float = 128.0, sd: float = 1.0, chunk: int = 8):
    """Cutouts [ncand][nchan/fdown][nbin] (float32, on data's device).

    data: time-major uint8 [T][C] tensor; dm, t_top: per candidate (t_top in
    samples, the arrival at the top channel, which lands at index nbin//2)."""
    import torch

    dev = data.device
    T, Cn = data.shape
    f = np.asarray(freqs, np.float64)
    fmax = f.max()
    dm = np.asarray(dm, np.float64)
    t_top = np.asarray(t_top, np.int64)
    ncand = dm.size
    G = Cn // fdown
    out = torch.empty((ncand, G, nbin), dtype=torch.float32, device=dev)
    kk = torch.arange(nbin, device=dev, dtype=torch.int64)
    cols = torch.arange(G * fdown, device=dev, dtype=torch.int64)
    scale = 1.0 / (sd * np.sqrt(fdown))
    for c0 in range(0, ncand, chunk):
        c1 = min(ncand, c0 + chunk)
        x = KDM * dm[c0:c1, None] * (1.0 / f[None, : G * fdown] ** 2 - 1.0 / fmax ** 2) / tsamp
        d = np.sign(x) * np.floor(np.abs(x) + 0.5)  # half away from zero
        start = t_top[c0:c1, None] + d.astype(np.int64) - nbin // 2  # [n][C]
        idx = torch.from_numpy(start).to(dev)[:, :, None] + kk[None, None, :]  # [n][C][nbin]
        ok = (idx >= 0) & (idx < T)
        v = data[idx.clamp(0, T - 1), cols[None, :, None]].float()
        v = torch.where(ok, v - zero_point, torch.zeros_like(v))
        out[c0:c1] = v.view(c1 - c0, G, fdown, nbin).sum(2) * scale
    return out
