#!/usr/bin/env python
"""Throughput of the TransientX search hot path on one MI355X.

Workload (the 4096-channel production single beam): 8-bit filterbank,
4096 channels over 856-1712 MHz, tsamp 76.56 us, 300 s, read in 2.0 s
blocks with 0.1 overlap; ddplan rows DM 0-500 step 0.5 (64 sub-bands),
500-1500 step 2 (64), 1500-3000 step 5 (128); boxcar ladder to 0.1 s with
snrloss 0.1 (33 widths); threshold 7.  Noise N(128, 12^2), 0.5% of the
channels with persistent narrowband RFI at 5x amplitude, 20 zero-DM
impulses, 50 dispersed pulses (DM log-uniform 10-3000, FWHM 0.1-50 ms,
S/N 8-50).  The observation is generated on the device and is resident in
HBM before timing starts.

One step = one full search of the observation: stream reset, 150 blocks
through all three rows (dedispersion, boxcar, thresholded-point
compaction), the end-of-file flush (row r runs offset/ndump + 1 flush
blocks, 2, 4 and 7 here, as transientx_fil flushes each row), and every
thresholded point in pinned host memory.  value = DM-trial*samples/s counting ndump samples per
block per DM over the 150 data blocks; rtf = 300 s / step time.  By
default the boxcar emits the thresholded points without materializing the
S/N and width maps (--maps on stores them).

Besides the headline (blocks resident in HBM), the line carries legs of
the same search: with the maps stored, host-fed (every block pushed from
pinned host memory inside the step, copies overlapping the search), and
raw 8-bit blocks through the device front end (Equalize) into the f32
search; the error of the FP32 results against FP64 per row; and the CPU
baseline on every host CPU this process may use.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "DM-trial·samples/s + real-time factor, 4096-ch, 1 GPU; batched vs plain"
HBM_PEAK = 8.0e12
# B/s: measured gfx950 LDS read rates of the form each kernel uses (probes of
# 2048 workgroups reading a 16 KB array in a loop): stage 1 ds_read2_b32 of
# two adjacent words with lanes 4 bytes apart, stage 2 ds_read2st64_b32, the
# boxcar with maps one ds_read_b32 per running sum
LDS_PEAK = {"stage1": 82.1e12, "stage2": 130.5e12, "boxcar": 76.6e12}

ROWS = [  # (dms, ddm, ndm, nsub)
    (0.0, 0.5, 1000, 64),
    (500.0, 2.0, 500, 64),
    (1500.0, 5.0, 300, 128),
]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(seconds: float):
    from transientx_amd.synth import Observation, random_pulses

    tsamp = 76.56e-6
    nsamp = int(seconds / tsamp)
    # 50 pulses and 20 impulses per 300 s; shorter runs keep the same rates
    npulse = max(1, round(50 * seconds / 300.0))
    nimp = max(1, round(20 * seconds / 300.0))
    t_lo, t_hi = min(5.0, 0.1 * seconds), seconds - min(20.0, 0.1 * seconds)
    pulses = random_pulses(2, npulse, 10.0, 3000.0, t_lo, t_hi, 1e-4, 50e-3, 8.0, 50.0,
                           log_dm=True)
    obs = Observation(nchans=4096, fbot=856.0, ftop=1712.0, tsamp=tsamp, nsamp=nsamp, mean=128.0,
                      sd=12.0, seed=2, pulses=pulses, rfi_frac=0.005, rfi_amp=5.0, impulses=nimp)
    return obs


def algorithmic_bytes(info, nchans, nblocks_data, in_bytes=1, maps=True):
    """Per-kernel algorithmic bytes of one row over the data blocks.

    Per DM-trial*sample (rho = N/ndump): input s_in*C/ndm (in_bytes per
    sample: 1 for u8, 4 for f32), sub-band
    series written 4*nsub*ndmc/ndm and read 4*nsub*ndmc/ndm*rho, series
    written and read 4*rho each, mxS + mxwn written 8*rho when the maps are
    stored (points only: the boxcar reads the series and writes points).
    """
    ndm, nsub, ndmc, ndump, N = info.ndm, info.nsub, info.ndmc, info.ndump, info.nsamples
    per_block = {
        "stage1": nchans * ndump * in_bytes + 4 * nsub * ndmc * ndump,
        "stage2": 4 * nsub * ndmc * N + 4 * ndm * N,
        "boxcar": 4 * ndm * N + (8 * ndm * N if maps else 0),
    }
    return {k: v * nblocks_data for k, v in per_block.items()}


def width_ladder_of(tsamp):
    from transientx_amd import width_ladder

    return width_ladder(tsamp, 1e-4, 0.1, 0.1)


def host_threads():
    """The CPUs this process may run on (the box's CPU share)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def autotune_rows(obs, ndump, vwn, dtype, args):
    """Launch shapes per row timed on the device (None: the cost model's)."""
    if args.no_autotune:
        return [None] * len(ROWS)
    from transientx_amd.autotune import tune_row

    t0 = time.time()
    # timed afresh on every run unless TX_AUTOTUNE_CACHE names a cache file
    cached = bool(os.environ.get("TX_AUTOTUNE_CACHE"))
    shapes = [tune_row(obs.nchans, obs.freqs, obs.tsamp, ndump, r, vwn, 0.1, use_cache=cached, dtype=dtype)
              for r in ROWS]
    log(f"autotuned {dtype} launch shapes {shapes} in {time.time() - t0:.1f} s")
    return shapes


def search_setup(obs, ndump, vwn, dtype, maps, shapes, exact=False):
    from transientx_amd import Beam, Plan

    beam = Beam(obs.nchans, ndump, dtype=dtype, zero_point=128)
    plans = [Plan(beam, freqs=obs.freqs, tsamp=obs.tsamp, dms=dms, ddm=ddm, ndm=ndm, vwn=vwn, overlap=0.1,
                  nsub=nsub, store_maps=maps, shapes=sh, exact_order=exact)
             for (dms, ddm, ndm, nsub), sh in zip(ROWS, shapes)]
    return beam, plans


def search_leg(obs, ndump, vwn, dtype, maps, shapes, blocks, thre, args, pipeline="graph", events="off",
               host=False, frontend=None, exact=False, nleft=None, maps_to_host=False):
    """One configuration of the production search: warm-up steps, then
    args.steps timed steps (a step = stream reset, every block through every
    row, every thresholded point stored by the kernels into pinned host
    lists, and the wait for the last one).  host: the
    blocks are pinned host tensors pushed asynchronously.  frontend:
    (offset, gain) of the device front end feeding an f32 beam with raw
    8-bit blocks.  nleft: the end-of-file flush blocks each row runs (the
    last max(nleft) blocks are the flush; row r runs the first nleft[r] of
    them, offset/ndump + 1 as src/singlepulse/dedisperse_fil.cpp:347 has
    it); with a graph the blocks every row still runs are graph launches
    and the rest direct launches.  maps_to_host: after every ready boxcar
    the row's S/N and width maps are copied to pinned host arrays
    (tx_boxcar_rows), as the drop-in used unchanged hands Cluster::run the
    raw mxS / mxwn arrays (direct launches only).  Returns the step time, per-kernel times (summed over the
    timed steps) and the points per row.  exact: the boxcar in exact_order
    mode (the reference's sequential order, bit-identical maps)."""
    import torch

    beam, plans = search_setup(obs, ndump, vwn, dtype, maps, shapes, exact=exact)
    if frontend is not None:
        beam.set_frontend(obs.nchans, 1, 1, frontend[0], frontend[1])
    cap = 1 << 25
    hostbuf = [torch.empty(cap * 24, dtype=torch.uint8, pin_memory=True) for _ in plans]
    nb = len(blocks)
    mapbuf = []
    if maps_to_host:
        import ctypes as C

        from transientx_amd import _lib as L

        assert maps and pipeline == "direct"
        mapbuf = [(torch.empty(p.info.ndm * p.info.nsamples, dtype=torch.float32, pin_memory=True),
                   torch.empty(p.info.ndm * p.info.nsamples, dtype=torch.int32, pin_memory=True)) for p in plans]
    built = [False]

    def step(use_graph):
        beam.reset()
        # the points go straight into the pinned host lists as the kernels
        # find them (tx_session_begin_host): no D2H at the end of the step
        for p, h in zip(plans, hostbuf):
            p.session_begin(cap, host=h)
        ngraph = 0
        if use_graph:
            # one hipGraph launch per block: transpose, the rows as parallel
            # branches, clock tick (tx_beam_graph)
            if not built[0]:
                beam.build_graph(thre)
                built[0] = True
            ngraph = len(blocks) - (max(nleft) - min(nleft) if nleft else 0)
            for blk in blocks[:ngraph]:
                beam.graph_push(blk, host_async=host)
        for b, blk in enumerate(blocks[ngraph:], ngraph):
            if host:
                beam.push_async(blk)
            else:
                beam.push(blk)
            for r, p in enumerate(plans):
                if nleft and b >= nb - max(nleft) + nleft[r]:
                    continue  # this row has run its flush blocks
                p.dedisp_run()
                if p.boxcar_run(thre) and maps_to_host:
                    S, W = mapbuf[r]
                    L.check(L.lib().tx_boxcar_rows(p.h, 0, p.info.ndm, C.c_void_p(S.data_ptr()),
                                                   C.c_void_p(W.data_ptr())))
        return [p.session_end_into(h.data_ptr(), cap) for p, h in zip(plans, hostbuf)]

    graph = pipeline == "graph"
    for _ in range(args.warmup):
        step(graph)
    beam.sync()
    torch.cuda.synchronize()
    beam.timing()  # clear
    # graph launches carry no per-kernel events: kernels are timed in a
    # separate direct-launch search
    ev = "separate" if graph and events == "timed" else events
    beam.set_timing(ev == "timed")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pts = step(graph)
    beam.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    beam.set_timing(False)
    kt = beam.timing()
    if ev == "separate":
        beam.set_timing(True)
        step(False)
        beam.set_timing(False)
        kt = beam.timing()
        kt = {k: (v[0] * args.steps, v[1] * args.steps) for k, v in kt.items()}
    beam.close()
    return {"ms_per_step": round((t1 - t0) * 1e3 / args.steps, 3), "kernels": kt,
            "points": [int(n) for n in pts]}


def fp64_errors(obs, data, ndump, vwn, thre, shapes):
    """Error of the device's FP32 results against FP64 on this workload:
    per row, at the block where its boxcar gate opens, the dedispersed
    series against the same two-stage sums in FP64
    (series_max_rel = max |x - x64| / max |x64|) and the S/N maximum per
    output (the rows recomputed by tx_boxcar_rows) against Boxcar::match in
    FP64 (snr_max_abs).  8-bit input: both through the 8-bit search and
    through the device front end (f32 sums)."""
    import oracle
    import torch

    from transientx_amd import frontend

    x = None
    out = {}
    off, gain = frontend.stats(data[:ndump], 1, 1)
    offn, gainn = off.cpu().numpy(), gain.cpu().numpy()
    for mode in ("u8", "frontend"):
        beam, plans = search_setup(obs, ndump, vwn, "u8" if mode == "u8" else "f32", False, [None] * len(ROWS))
        if mode == "frontend":
            beam.set_frontend(obs.nchans, 1, 1, offn, gainn)
        got = [None] * len(plans)
        b = 0
        while any(g is None for g in got):
            b += 1
            beam.push(data[(b - 1) * ndump:b * ndump])
            for r, p in enumerate(plans):
                p.dedisp_run()
                if p.boxcar_run(thre) and got[r] is None:
                    o = p.info
                    tim = np.stack([p.timdata(i) for i in range(o.ndm)])
                    S, _ = p.rows(0, o.ndm)
                    got[r] = (b, tim, S)
        beam.close()
        if x is None:
            x = data[: b * ndump].cpu().numpy()
        rows = []
        for r, (dms, ddm, ndm, nsub) in enumerate(ROWS):
            blk, tim, S = got[r]
            op = oracle.Plan(freqs=obs.freqs, tsamp=obs.tsamp, ndump=ndump, dms=dms, ddm=ddm, ndm=ndm,
                             overlap=0.1, nsub=nsub)
            o = op.info()
            w0 = op.window_start(blk)
            a = max(0, w0)
            seg = x[a:w0 + o.nsamples + o.maxdelay]
            xo = oracle.frontend_u8(seg, 1, 1, offn, gainn) if mode == "frontend" else \
                seg.astype(np.float32) - 128.0
            t64 = op.dedisperse_f64(xo, w0 - a, nthreads=host_threads())
            del xo
            s64 = oracle.boxcar_f64(t64, vwn, nthreads=host_threads())
            rows.append({"block": blk,
                         "series_max_rel": float(np.abs(tim - t64).max() / np.abs(t64).max()),
                         "snr_max_abs": float(np.abs(S.astype(np.float64) - s64).max())})
            del t64, s64
        out[mode] = {"series_max_rel_vs_fp64": max(r["series_max_rel"] for r in rows),
                     "snr_max_abs_vs_fp64": max(r["snr_max_abs"] for r in rows), "rows": rows}
    torch.cuda.empty_cache()
    return out


def cpu_baseline(obs, data_gpu, vwn, thre, nthreads, max_blocks):
    """The oracle's streaming CPU search (the reference path restated in C,
    OpenMP over DMs) on a bounded sample: each row is warmed up to boxcar
    readiness untimed, then max_blocks ready blocks are timed."""
    import oracle

    ndump = int(2.0 / obs.tsamp)
    total_units, total_t = 0, 0.0
    for dms, ddm, ndm, nsub in ROWS:
        p = oracle.Plan(freqs=obs.freqs, tsamp=obs.tsamp, ndump=ndump, dms=dms, ddm=ddm, ndm=ndm,
                        overlap=0.1, nsub=nsub)
        info = p.info()
        ready_at = -(-(info.offset - info.noverlap + ndump) // ndump)
        st = oracle.Stream(p, vwn, 128, nthreads)
        for b in range(1, ready_at + max_blocks):
            blk = data_gpu[(b - 1) * ndump:b * ndump].cpu().numpy()
            t0 = time.perf_counter()
            st.run(blk, thre)
            dt = time.perf_counter() - t0
            if b >= ready_at:
                total_t += dt
                total_units += ndm * ndump
        st.close()
        log(f"cpu baseline row dms={dms}: ready at block {ready_at}, cumulative {total_t:.2f} s")
    return total_units / total_t, total_t


COPY_BYTES = 1 << 30


def copy_calibration():
    """A plain elementwise copy of COPY_BYTES read + COPY_BYTES written per
    dispatch: the known-size kernel that scales FETCH_SIZE/WRITE_SIZE."""
    import torch

    x = torch.randn(COPY_BYTES // 4, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(5):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": COPY_BYTES, "bytes_written": COPY_BYTES,
                                           "dispatches": 5}}), flush=True)


def load_traffic(kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy calibration), recorded in
    profiles/traffic.txt as `kernel bytes_per_launch` lines; None if absent."""
    f = ROOT / "profiles" / "traffic.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        parts = line.split()
        if len(parts) == 2 and parts[0] == kernel and not line.startswith("#"):
            return int(float(parts[1]))
    return None


def refine_bench(args):
    """The refinement grid (Candidate::dedisperse + optimize + matched_filter)
    on a batch of cutouts: 4096 float32 cutouts of 1024 channels x 512
    samples (1200-1500 MHz, 98.304 us), unit-variance noise, 3% of channels
    zeroed, one dispersed pulse each (DM U[50,1500], arrival +-8 samples
    from the centre, FWHM 1-20 ms, S/N 8-40), each cutout dedispersed at a
    DM off by up to 2 pc/cc.  Generated on the device; one step refines the
    whole batch."""
    import torch

    from transientx_amd.refine import Refiner

    ncand, nchan, nbin, tbin = args.cutouts, 1024, 512, 98.304e-6
    f = np.linspace(1500.0, 1200.0 + 300.0 / nchan, nchan)
    g = torch.Generator(device="cuda").manual_seed(4)
    rng = np.random.default_rng(4)
    dm_true = rng.uniform(50, 1500, ncand)
    dm_cand = dm_true + rng.uniform(-2, 2, ncand)
    t0 = nbin // 2 + rng.integers(-8, 9, ncand)
    fwhm = rng.uniform(1e-3, 20e-3, ncand) / tbin
    snr = rng.uniform(8, 40, ncand)
    x = torch.randn((ncand, nchan, nbin), generator=g, device="cuda", dtype=torch.float32)
    K = 4.148741601e3
    dl = torch.tensor(K * (1 / f ** 2 - 1 / f.max() ** 2) / tbin, device="cuda")
    tt = torch.arange(nbin, device="cuda", dtype=torch.float64)
    for c0 in range(0, ncand, 256):
        c1 = min(ncand, c0 + 256)
        ctr = torch.tensor(t0[c0:c1], device="cuda")[:, None] + \
            torch.tensor(dm_true[c0:c1], device="cuda")[:, None] * dl[None, :]
        sig = torch.tensor(fwhm[c0:c1] / 2.3548, device="cuda")[:, None, None]
        d = (tt[None, None, :] - ctr[:, :, None] + nbin / 2) % nbin - nbin / 2
        prof = torch.exp(-0.5 * (d / sig) ** 2)
        amp = torch.tensor(snr[c0:c1], device="cuda")[:, None, None] / torch.sqrt(
            prof.sum(-1, keepdim=True) * nchan) * 1.0
        x[c0:c1] += (amp * prof).float()
    mask = torch.rand((ncand, nchan), generator=g, device="cuda") < 0.03
    x[mask] = 0.0
    torch.cuda.synchronize()
    R = Refiner(f, nbin, tbin, 0.1, dedisperse=True, isdedispersed=True, max_batch=args.refine_batch)
    for _ in range(args.warmup):
        res = R.run(x, dm_cand)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        res = R.run(x, dm_cand)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    ms = (t2 - t1) * 1e3 / args.steps
    err = np.abs(res["dm_maxsnr"] - dm_true)
    # gradient refinement of the same cutouts: coarse grid start + ridge bisection
    from transientx_amd.diff import DiffSNR, refine_gradient

    D = DiffSNR(f, nbin, tbin, softness=1.0)
    w = (~mask).float().contiguous()
    F = lambda v: torch.tensor(v, dtype=torch.float32, device="cuda")
    # warm-up on 8 cutouts (first launches load the kernels), not timed
    refine_gradient(D, x[:8].contiguous(), w[:8].contiguous(), F(dm_cand[:8]), F(np.full(8, nbin / 2.0)))
    torch.cuda.synchronize()
    ta = time.perf_counter()
    adm, at0, aw, asnr, nev = refine_gradient(D, x, w, F(dm_cand), F(np.full(ncand, nbin / 2.0)))
    torch.cuda.synchronize()
    ta = time.perf_counter() - ta
    adm = adm.cpu().numpy()
    # the grid judge on the same S/N (DM step 0.05, t0 step 0.25, 16 widths) on a sample
    from transientx_amd.diff import grid_search

    nj = min(256, ncand)
    tj = time.perf_counter()
    jdm, _, _, jsnr, njev = grid_search(D, x[:nj].contiguous(), w[:nj].contiguous(), dm_cand[:nj],
                                        np.full(nj, nbin / 2.0), ddm=0.05, dm_half=2.0, dt=0.25,
                                        t_half=8.0, widths=np.geomspace(1.0, 256.0, 16))
    tj = time.perf_counter() - tj
    D.close()
    grad = {"ms": round(ta * 1e3, 3), "cutouts_per_s": round(ncand / ta, 1),
            "snr_evaluations_per_cutout": nev,
            "grid_evaluations_per_cutout": int(R.ndm * nbin * R.nbox),
            "judge_sample": nj, "judge_evaluations_per_cutout": int(njev),
            "dm_within_0p1_of_judge": round(float(np.mean(np.abs(adm[:nj] - jdm) <= 0.1)), 4),
            "dm_within_0p1_or_snr_tie": round(float(np.mean((np.abs(adm[:nj] - jdm) <= 0.1) |
                                                            (asnr.cpu().numpy()[:nj] >= jsnr * (1 - 1e-4)))), 4),
            "dm_vs_judge_max_abs": round(float(np.max(np.abs(adm[:nj] - jdm))), 4),
            "dm_vs_judge_median_abs": round(float(np.median(np.abs(adm[:nj] - jdm))), 4),
            "judge_ms": round(tj * 1e3, 3),
            "snr_vs_judge_min_ratio": round(float(np.min(asnr.cpu().numpy()[:nj] / jsnr)), 4),
            "dm_vs_tree_grid_median_abs": round(float(np.median(np.abs(adm - res["dm_maxsnr"]))), 4),
            "dm_error_median": round(float(np.median(np.abs(adm - dm_true))), 3)}
    cpu = None
    if args.cpu_blocks > 0:
        import oracle
        from concurrent.futures import ThreadPoolExecutor

        nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)
        nsample = min(ncand, 4 * nthreads)
        xs = x[:nsample].cpu().numpy()

        def one(c):
            xd = oracle.cand_dedisperse(xs[c], f, dm_cand[c], tbin)
            dmt = oracle.refine_optimize(xd, f, tbin, dm_cand[c], True)[0]
            return oracle.refine_mf(dmt, 0.1)[2:4]

        tc = time.perf_counter()
        with ThreadPoolExecutor(nthreads) as ex:
            cres = list(ex.map(one, range(nsample)))
        tc = time.perf_counter() - tc
        assert all((a, b) == (r["maxdmid"], r["maxtid"]) for (a, b), r in zip(cres, res[:nsample]))
        cpu = {"value": round(nsample / tc, 2), "unit": "cutouts/s", "cores": nthreads,
               "kind": "port",
               "sample": f"{nsample} of the cutouts, oracle Candidate::dedisperse + optimize + "
                         f"matched_filter, one cutout per thread ({tc:.1f} s wall)"}
    line = {
        "metric": "refinement grid: cutouts/s (1024 ch x 512 samples)",
        "value": round(ncand / (ms / 1e3), 1), "unit": "cutouts/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (generated on the device, resident in HBM)",
        "config": {"workload": f"{ncand} cutouts of 1024 ch x 512 samples, 1200-1500 MHz, "
                               f"98.304 us, DM U[50,1500], grid of {R.ndm} trials x {R.nbox} widths"},
        "dm_error_median": round(float(np.median(err)), 3),
        "dm_grid_step": round(float(np.median(res["ddm"])), 4),
        "gradient_refinement": grad,
        "cpu_baseline": cpu,
    }
    R.close()
    print(json.dumps(line), flush=True)


def batch_specs(nbeams: int, scale: float = 1.0):
    """The batching workload: nbeams beams, seed 3+beam; channel counts from
    {1024, 2048, 4096} with the 1200-1500 MHz / 98.304 us band (1024) or the
    856-1712 MHz / 76.56 us band; lengths U[60, 600] s times scale; blocks of
    2 s drawn from per-band pools of 24 s (noise, 0.5% narrowband RFI
    channels at 5x, 2 zero-DM impulses -- the production recipe's rate of
    about one per 15 s -- and 10 dispersed pulses) generated on the device,
    each beam starting at its own pool offset; the 1024-channel beams search
    DM 0-1000 step 1 (32 sub-bands), the others the three production rows."""
    from transientx_amd import width_ladder
    from transientx_amd.batch import BeamSpec
    from transientx_amd.synth import Observation, gpu_generate, random_pulses

    pools = {}
    specs = []
    for b in range(nbeams):
        rng = np.random.default_rng(3 + b)
        C = int(rng.choice([1024, 2048, 4096]))
        secs = float(rng.uniform(60.0, 600.0)) * scale
        if C == 1024:
            fb, ft, tsamp, rows, ov = 1200.0, 1500.0, 98.304e-6, [(0.0, 1.0, 1000, 32)], 0.0
        else:
            fb, ft, tsamp, ov = 856.0, 1712.0, 76.56e-6, 0.1
            rows = [(0.0, 0.5, 1000, 64), (500.0, 2.0, 500, 64), (1500.0, 5.0, 300, 128)]
            if C == 2048:
                rows = [(d, dd, n, ns // 2) for d, dd, n, ns in rows]
        ndump = int(2.0 / tsamp)
        if C not in pools:
            npool = 12
            pulses = random_pulses(20 + C, 10, 10.0, 1000.0, 1.0, 2.0 * npool - 2.0, 1e-3, 20e-3,
                                   10.0, 40.0, log_dm=True)
            obs = Observation(nchans=C, fbot=fb, ftop=ft, tsamp=tsamp, nsamp=ndump * npool, mean=128.0,
                              sd=12.0, seed=C, pulses=pulses, rfi_frac=0.005, rfi_amp=5.0, impulses=2)
            data, _ = gpu_generate(obs, flush_blocks=0, ndump=ndump)
            pools[C] = ([data[k * ndump:(k + 1) * ndump] for k in range(npool)], obs.freqs)
        blocks, freqs = pools[C]
        nb = max(1, int(secs / 2.0))
        specs.append(BeamSpec(freqs=freqs, tsamp=tsamp, ndump=ndump, nblocks=nb, rows=rows,
                              vwn=width_ladder(tsamp, 1e-4, 0.1, 0.1), overlap=ov,
                              blocks=[blocks[(5 * b + k) % len(blocks)] for k in range(nb)]))
    return specs


def batch_compare(specs, graphs: bool = True, host: bool = True, drain_every: int = 1, workers: int = 8):
    """Plain (beam after beam, points to the host after every block, as
    transientx_fil clusters each block) against batched (rounds of one block
    per live beam, one stream per beam, each beam's points drained every
    round): resident blocks, then blocks pushed from pinned host memory
    inside the timed region (plain: copy then wait, as the drop-in pushes;
    batched: queued copies on double-buffered staging).  The batched paths
    drain each beam's points every drain_every rounds and deal the beams to
    `workers` host threads (also timed: one thread).  Candidate lists
    are compared per (beam, row) by count and order-independent digest
    (batch.points_digest), reduced after each timed path, so that a full
    length run (about 10^9 points) is never held twice."""
    import torch

    from transientx_amd.batch import BeamSearch, collect, run_batched, run_plain

    searches = [BeamSearch(s, cap=1 << 23) for s in specs]
    units = sum(s.units() for s in searches)
    # warm-up (plan first-use allocations, the pinned drain buffer): one
    # block of every beam, untimed
    for s in searches:
        s.begin(7.0)
        s.step(7.0)
        s.drain()
        s.end(digest=True)
    torch.cuda.synchronize()
    out = {"beams": len(specs), "blocks": sum(s.nblocks for s in specs), "units": int(units)}

    def timed(name, fn, *a, **k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn(searches, 7.0, *a, **k)
        torch.cuda.synchronize()
        out[name + "_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
        log(f"batch {name}: {out[name + '_ms']:.1f} ms")
        return collect(searches, digest=True)

    def same(a, b):
        return a == b

    plain = timed("plain", run_plain, drain=True)
    batched = timed("batched", run_batched, drain=True, drain_every=drain_every, workers=workers)
    ok = same(plain, batched)
    ok = ok and same(plain, timed("batched_one_thread", run_batched, drain=True, drain_every=drain_every))
    out["drain_every_rounds"] = drain_every
    out["host_threads"] = workers
    if graphs:
        for s in searches:
            s.graph = True
        ok = ok and same(plain, timed("batched_graph", run_batched, drain=True, drain_every=drain_every,
                                      workers=workers))
        for s in searches:
            s.graph = False
    if host:
        hpool = {}
        for s in searches:
            sp = s.spec
            for blk in sp.blocks:
                if blk.data_ptr() not in hpool:
                    hpool[blk.data_ptr()] = blk.cpu().pin_memory()
            sp.host_blocks = [hpool[blk.data_ptr()] for blk in sp.blocks]
        ok = ok and same(plain, timed("plain_host", run_plain, "host", drain=True))
        ok = ok and same(plain, timed("batched_host", run_batched, "host_async", drain=True,
                                      drain_every=drain_every, workers=workers))
        out["host_bytes"] = int(sum(s.spec.nblocks * s.spec.ndump * len(s.spec.freqs) for s in searches))
    for s in searches:
        s.close()
    out["identical_candidates"] = bool(ok)
    out["points"] = int(sum(n for pa in plain for n, _ in pa))
    out["batched_over_plain"] = round(out["plain_ms"] / out["batched_ms"], 3)
    if host:
        out["batched_over_plain_host_fed"] = round(out["plain_host_ms"] / out["batched_host_ms"], 3)
    return out


def batch_bench(args):
    """64 beams of uneven length (60-600 s) and channel count (batch_specs)
    searched beam by beam (plain) and all together (batched); candidate
    lists must be identical."""
    import torch

    specs = batch_specs(args.beams, args.batch_scale)
    torch.cuda.synchronize()
    log(f"{len(specs)} beams, {sum(s.nblocks for s in specs)} blocks, channels "
        f"{sorted(set(len(s.freqs) for s in specs))}")
    r = batch_compare(specs)
    line = {
        "metric": "batched search: DM-trial\u00b7samples/s over 64 beams of uneven length and channels",
        "value": round(r["units"] / (r["batched_ms"] / 1e3), 1), "unit": "DM-trial\u00b7samples/s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": r["batched_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic (generated on the device, resident in HBM; host-fed legs from pinned memory)",
        "config": {"workload": f"{args.beams} beams, channels 1024/2048/4096, "
                               f"{60 * args.batch_scale:g}-{600 * args.batch_scale:g} s, 2 s blocks, "
                               f"points drained every block (plain) or {r['drain_every_rounds']} round(s) "
                               f"(batched, beams on {r['host_threads']} host threads)",
                   "path": "batched"},
        "plain_value": round(r["units"] / (r["plain_ms"] / 1e3), 1),
        "batch": r,
    }
    print(json.dumps(line), flush=True)
    if not r["identical_candidates"]:
        sys.exit(1)


def rfi_bench(args):
    """RFI cleaning as replot_fil runs it on candidate cutouts before the
    refinement grid (src/singlepulse/replot_fil.cpp:409-447): Candidate::zdot,
    then Candidate::kadaneF with td 4, fd 2, threshold 5 and wnlimit 0 (the
    three-argument call), on 4096 float32 cutouts of 1024 channels x 512
    samples: unit-variance noise, zero-DM interference with per-channel gains,
    a narrowband burst per cutout.  A step cleans the whole batch in place
    (the cutouts are restored from a pristine copy before each step, outside
    the timed region).  Also timed: zdot of one production search block
    (26,123 samples x 4096 channels, time-major), the front end's "zdot"."""
    import torch

    from transientx_amd import rfi

    ncut, nchan, nbin, td, fd, thr = args.cutouts, 1024, 512, 4, 2, 5.0
    g = torch.Generator(device="cuda").manual_seed(6)
    x0 = torch.randn((ncut, nchan, nbin), generator=g, device="cuda")
    x0 += torch.rand((ncut, nchan, 1), generator=g, device="cuda") * 2 * (
        torch.randn((ncut, 1, nbin), generator=g, device="cuda") * 3)
    j = torch.randint(0, nchan, (ncut,), generator=g, device="cuda")
    x0[torch.arange(ncut, device="cuda"), j, nbin // 4:nbin // 2] += 6.0
    x = torch.empty_like(x0)
    wl = np.zeros(ncut, np.int32)
    times = []
    for step in range(args.warmup + args.steps):
        x.copy_(x0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rfi.kadane_f(rfi.zdot(x), td, fd, thr, wl)
        torch.cuda.synchronize()
        if step >= args.warmup:
            times.append(time.perf_counter() - t0)
    ms = 1e3 * sum(times) / len(times)
    zapped = float((x == 0).float().mean()) - float((x0 == 0).float().mean())
    # per-kernel event timing of one pass (the dominant kernel for the roofline)
    x.copy_(x0)
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    e[0].record()
    rfi.zdot(x)
    e[1].record()
    rfi.kadane_f(x, td, fd, thr, wl)
    e[2].record()
    torch.cuda.synchronize()
    zms, kms = e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2])
    nel = ncut * nchan * nbin
    # algorithmic bytes of zdot: every sample read once and written once
    zach = 8.0 * nel / (zms / 1e3) / 1e9
    # one production search block, time-major, the front end's zdot
    blk = torch.randn((26123, 4096), generator=g, device="cuda")
    rfi.zdot(blk, layout="time")
    torch.cuda.synchronize()
    tb = time.perf_counter()
    for _ in range(3):
        rfi.zdot(blk, layout="time")
    torch.cuda.synchronize()
    bms = (time.perf_counter() - tb) * 1e3 / 3
    cpu = None
    if args.cpu_blocks > 0:
        import oracle
        from concurrent.futures import ThreadPoolExecutor

        nthreads = args.cpu_threads or min(16, os.cpu_count() or 1)
        nsample = min(ncut, 4 * nthreads)
        xs = x0[:nsample].cpu().numpy()
        tc = time.perf_counter()
        with ThreadPoolExecutor(nthreads) as ex:
            cres = list(ex.map(lambda c: oracle.kadanef(oracle.zdot(xs[c]), td, fd, thr, 0), range(nsample)))
        tc = time.perf_counter() - tc
        ref = x[:nsample].cpu().numpy()
        assert all(np.array_equal(ref[c].view(np.uint32), cres[c].view(np.uint32)) for c in range(nsample))
        cpu = {"value": round(nsample / tc, 2), "unit": "cutouts/s", "cores": nthreads, "kind": "port",
               "sample": f"{nsample} of the cutouts, oracle Candidate::zdot + kadaneF, one cutout per "
                         f"thread ({tc:.2f} s wall); results bit-identical to the GPU's"}
    line = {
        "metric": "RFI cleaning of candidate cutouts: cutouts/s (zdot + kadaneF, 1024 ch x 512 samples)",
        "value": round(ncut / (ms / 1e3), 1), "unit": "cutouts/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32/f64", "data": "synthetic (generated on the device, resident in HBM)",
        "config": {"workload": f"{ncut} cutouts of {nchan} ch x {nbin} samples; zdot, then kadaneF td {td} "
                               f"fd {fd} threshold {thr} wnlimit 0"},
        "zapped_fraction": round(zapped, 5),
        "kernels_ms": {"zdot": round(zms, 3), "kadaneF": round(kms, 3)},
        "roofline": {"kernel": "zdot", "bound": "hbm", "achieved": round(zach, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(zach / 8000.0, 4), "traffic": None,
                     "note": "algorithmic bytes 8 per sample (read once, written once); the kernels make "
                             "four passes in the reference's sequential FP64 order"},
        "search_block_zdot": {"ms": round(bms, 3), "block": "26123 x 4096 time-major (2 s at 76.56 us)",
                              "real_time_factor": round(2000.0 / bms, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


# DM 0-5000 in four rows with steps 0.2 / 1 / 4 / 10, on the production ddplan's boundaries
HTR_ROWS = [(0.0, 0.2, 2500, 64), (500.0, 1.0, 1000, 64), (1500.0, 4.0, 375, 128), (3000.0, 10.0, 200, 128)]


def htr_bench(args):
    """High-time-resolution combined run: beams of 8-bit data, 4096 channels
    over 1000-1500 MHz, tsamp 16 us, searched in 0.5 s blocks over DM 0-5000
    in four rows (steps 0.2/1/4/10), width ladder 16 us - 5 ms (26 widths).
    Seed 5+beam: 20 bursts per beam, DM U[20, 4900], FWHM 20-500 us, S/N
    8-30, an f^-4 scattering tail of U[0, 1] ms at 1000 MHz, arrival times
    that keep each burst's sweep inside the data.  Beams run batched (one
    stream per beam, groups of --htr-group beams resident at a time, each
    group's data generated before its timed search), launch shapes
    autotuned per row.  Every cluster above 7 sigma is then cut out
    (dedispersed at its search DM, 1024 channels) and refined by gradient
    ascent; the refinement is compared with the injected truth and with the
    grid judge on the same S/N."""
    import torch

    from transientx_amd import host, width_ladder
    from transientx_amd.autotune import tune_row
    from transientx_amd.batch import BeamSearch, BeamSpec, collect, run_batched
    from transientx_amd.cutout import dedispersed_cutouts, group_freqs
    from transientx_amd.diff import DiffSNR, grid_search, refine_gradient
    from transientx_amd.synth import KDM, Observation, Pulse, burst_amplitude, dmdelay, gpu_bursts

    C, fb, ft, ts = 4096, 1000.0, 1500.0, 16e-6
    ndump = int(round(0.5 / ts))
    nbeams, group, secs = args.htr_beams, args.htr_group, args.htr_seconds
    nblk = int(round(secs / 0.5))
    vwn = width_ladder(ts, 16e-6, 5e-3, 0.1)
    freqs = Observation(nchans=C, fbot=fb, ftop=ft, tsamp=ts, nsamp=1).freqs
    fmax, fmin = float(freqs.max()), float(freqs.min())
    cached = bool(os.environ.get("TX_AUTOTUNE_CACHE"))
    t0 = time.time()
    shapes = [tune_row(C, freqs, ts, ndump, r, vwn, 0.0, use_cache=cached) for r in HTR_ROWS]
    log(f"htr: autotuned launch shapes {shapes} in {time.time() - t0:.1f} s")
    # flush blocks (the end-of-file flush, seeded): a pool of noise blocks
    full_delay = dmdelay(5000.0, fmax, fmin)
    nflush = int(np.ceil((full_delay / ts + 16 * 1024) / ndump)) + 1
    fobs = Observation(nchans=C, fbot=fb, ftop=ft, tsamp=ts, nsamp=1, mean=128.0, sd=12.0, seed=5 + 999)
    fpool = [gpu_bursts(fobs, [], k * ndump, (k + 1) * ndump) for k in range(4)]
    t_search = 0.0
    units = 0
    obs_s = 0.0
    cands = []   # (beam, row, dm, tau, wn, S)
    truth = []   # (beam, dm, t0 s, fwhm, snr, scatter)
    gf = group_freqs(freqs, 4)
    fref = float(gf.max())
    kt_all = None
    for g0 in range(0, nbeams, group):
        beams = list(range(g0, min(nbeams, g0 + group)))
        datas, specs = [], []
        tg = time.time()
        for b in beams:
            rng = np.random.default_rng(5 + b)
            bursts, scat = [], []
            for _ in range(20):
                dm = float(rng.uniform(20.0, 4900.0))
                sweep = dmdelay(dm, fmax, fmin)
                tb = float(rng.uniform(0.2, max(0.3, secs - 0.3 - sweep - 0.02)))
                p = Pulse(dm=dm, t0=tb, fwhm=float(rng.uniform(20e-6, 500e-6)), snr=float(rng.uniform(8, 30)))
                sc = float(rng.uniform(0.0, 1e-3)) * (fmin / fmax) ** 4  # at the top channel
                bursts.append(p)
                scat.append(sc)
                truth.append((b, p.dm, p.t0, p.fwhm, p.snr, sc))
            obs = Observation(nchans=C, fbot=fb, ftop=ft, tsamp=ts, nsamp=nblk * ndump, mean=128.0, sd=12.0,
                              seed=5 + b)
            # S/N of the injected (scattered, dedispersed) signal under its best boxcar
            amps = [burst_amplitude(p, sc, freqs, ts, 12.0) for p, sc in zip(bursts, scat)]
            data = torch.empty((nblk * ndump, C), dtype=torch.uint8, device="cuda")
            for k in range(nblk):
                data[k * ndump:(k + 1) * ndump] = gpu_bursts(obs, bursts, k * ndump, (k + 1) * ndump,
                                                               scatter=scat, amps=amps)
            datas.append(data)
            blocks = [data[k * ndump:(k + 1) * ndump] for k in range(nblk)] + \
                     [fpool[k % len(fpool)] for k in range(nflush)]
            specs.append(BeamSpec(freqs=freqs, tsamp=ts, ndump=ndump, nblocks=nblk + nflush, rows=HTR_ROWS,
                                  vwn=vwn, overlap=0.0, blocks=blocks, shapes=shapes))
        torch.cuda.synchronize()
        log(f"htr: beams {beams} generated in {time.time() - tg:.1f} s")
        searches = [BeamSearch(sp, cap=1 << 21) for sp in specs]
        run_batched(searches[:1], 7.0)  # first-use allocations, untimed
        collect(searches[:1])
        torch.cuda.synchronize()
        ta = time.perf_counter()
        run_batched(searches, 7.0)
        torch.cuda.synchronize()
        t_search += time.perf_counter() - ta
        units += sum(sum(p.info.ndm for p in s.plans) * ndump * nblk for s in searches)
        obs_s += len(beams) * nblk * ndump * ts
        pts = collect(searches)
        if kt_all is None:
            # per-kernel launch times of one beam searched alone (direct launches)
            s0 = searches[0]
            s0.beam.timing()
            s0.beam.set_timing(True)
            run_batched([s0], 7.0)
            s0.beam.set_timing(False)
            kt_all = (s0.beam.timing(), [p.info for p in s0.plans], s0.spec.nblocks)
            collect([s0])
        # clusters of every (beam, row, block), then their cutouts while the
        # group's data is resident
        per = {}
        for bi, (b, s) in enumerate(zip(beams, searches)):
            for r, (plan, pr) in enumerate(zip(s.plans, pts[bi])):
                info = plan.info
                dms, ddm = HTR_ROWS[r][0], HTR_ROWS[r][1]
                for blk in np.unique(pr["block"]):
                    sel = pr[pr["block"] == blk]
                    p4 = np.zeros(sel.size, _lib_point_dtype())
                    for k in ("idm", "isamp", "S", "wn"):
                        p4[k] = sel[k]
                    cl = host.cluster(p4, dms=dms, ddm=ddm, fmax=fmax, fmin=fmin, tsamp=ts, radius=1e-3,
                                      maxncand=100, minpts=5)
                    w0 = int(blk) * ndump - info.maxdelay - info.nsamples
                    for c in cl:
                        per.setdefault((bi, r), []).append(
                            (b, r, dms + int(c["idm"]) * ddm, w0 + int(c["isamp"]), int(c["wn"]), float(c["S"])))
        for (bi, r), cl in per.items():
            x = dedispersed_cutouts(datas[bi], freqs, ts, [c[2] for c in cl], [c[3] for c in cl], HTR_NBIN[r],
                                    fdown=4, zero_point=128.0, sd=12.0)
            HTR_CUTS.append((r, cl, x))
            cands.extend(cl)
        for s in searches:
            s.close()
        del datas, searches, specs
        torch.cuda.empty_cache()
    value = units / t_search
    # refinement of every candidate, per row (cutout length and DM window by row)
    tr = time.perf_counter()
    res = []
    nev = njev = 0
    for r in range(len(HTR_ROWS)):
        items = [(cl, x) for rr, cl, x in HTR_CUTS if rr == r]
        if not items:
            continue
        cl = [c for it in items for c in it[0]]
        x = torch.cat([it[1] for it in items]).contiguous()
        nbin = HTR_NBIN[r]
        dm_half = HTR_DMHALF[r]
        model = DiffSNR(gf, nbin, ts, softness=1.0)
        w = torch.ones((x.shape[0], gf.size), dtype=torch.float32, device="cuda")
        F = lambda v: torch.tensor(v, dtype=torch.float32, device="cuda")
        z = F(np.zeros(len(cl)))
        tc = F(np.full(len(cl), nbin / 2.0))
        lad = np.geomspace(1.0, 256.0, 16)
        # a scattered burst's boxcar centre trails its Gaussian arrival by
        # tens of samples: t0 window +-32 samples for both
        adm, at0, aw, asnr, ne = refine_gradient(model, x, w, z, tc, dm_half=dm_half, t_half=32.0, widths=lad)
        torch.cuda.synchronize()
        jd, jt, jw, js, nj = grid_search(model, x, w, np.zeros(len(cl)), np.full(len(cl), nbin / 2.0), ddm=0.05,
                                         dm_half=dm_half, dt=0.25, t_half=32.0, widths=lad)
        model.close()
        nev, njev = max(nev, ne), max(njev, nj)
        adm, at0, asnr = adm.cpu().numpy(), at0.cpu().numpy(), asnr.cpu().numpy()
        for k, c in enumerate(cl):
            res.append((c, c[2] + float(adm[k]), float(at0[k]), float(asnr[k]), nbin, c[2] + float(jd[k]),
                        float(js[k])))
    t_ref = time.perf_counter() - tr
    # match candidates to the injected bursts: the highest-S candidate of
    # the same beam within 2 ms and 2 row DM steps of the truth
    errs, errs_search, terrs, jclose, jtie = [], [], [], [], []
    found = 0
    for (b, dm, tt, fw, snr, sc) in truth:
        best = None
        for (c, rdm, rt0, rs, nbin, jdm, js) in res:
            if c[0] != b:
                continue
            ddm_row = HTR_ROWS[c[1]][1]
            if abs(c[3] * ts - tt) <= 2e-3 and abs(c[2] - dm) <= max(2 * ddm_row, 1.0):
                if best is None or c[5] > best[0][5]:
                    best = (c, rdm, rt0, rs, nbin, jdm, js)
        if best is None:
            continue
        found += 1
        c, rdm, rt0, rs, nbin, jdm, js = best
        errs.append(abs(rdm - dm))
        errs_search.append(abs(c[2] - dm))
        # arrival at the cutout's reference frequency, in cutout samples
        t_true = nbin / 2 + tt / ts - c[3] + KDM * (dm - c[2]) * (1 / fref ** 2 - 1 / fmax ** 2) / ts
        terrs.append(abs(rt0 - t_true))
    if args.htr_dump:
        with open(args.htr_dump, "w") as fh:
            json.dump({"truth": [list(t) for t in truth],
                       "candidates": [list(c) + [rdm, rt0, rs, nbin, jdm, js]
                                      for (c, rdm, rt0, rs, nbin, jdm, js) in res]}, fh)
    for (c, rdm, rt0, rs, nbin, jdm, js) in res:
        jclose.append(abs(rdm - jdm) <= 0.1)
        jtie.append(abs(rdm - jdm) <= 0.1 or rs >= js * (1 - 1e-4))
    HTR_CUTS.clear()
    roof = None
    kernels = None
    if kt_all is not None:
        kt, infos, nb_all = kt_all
        alg = {"stage1": 0, "stage2": 0, "boxcar": 0}
        for i in infos:
            for k, v in algorithmic_bytes(i, C, nb_all).items():
                alg[k] += v
        alg["transpose"] = 2 * C * ndump * nb_all
        dom = max(("stage1", "stage2", "boxcar", "transpose"), key=lambda k: kt[k][0])
        if kt[dom][0] > 0:
            achieved = alg[dom] / (kt[dom][0] / 1e3)
            roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": None,
                    "launch_ms": round(kt[dom][0] / kt[dom][1], 4)}
        kernels = {k: {"ms_per_beam": round(v[0], 3), "launches": v[1]} for k, v in kt.items()}
    line = {
        "metric": "high-time-resolution combined run: DM-trial\u00b7samples/s, then gradient refinement",
        "value": round(value, 1), "unit": "DM-trial\u00b7samples/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(t_search * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic (generated on the device, resident in HBM)",
        "config": {"workload": f"{nbeams} beams x {secs:g} s, 4096 ch 1000-1500 MHz, tsamp 16 us, 0.5 s blocks "
                               f"(+{nflush} flush), DM 0-5000 in rows {HTR_ROWS}, 26 widths, "
                               f"{group} beams resident at a time", "path": "batched",
                   "launch_shapes": shapes},
        "rtf": round(obs_s / t_search, 1), "roofline": roof, "kernels_one_beam": kernels,
        "refinement": {
            "candidates": len(res), "bursts_injected": len(truth), "bursts_found": found,
            "ms": round(t_ref * 1e3, 1),
            "dm_error_vs_truth_median": round(float(np.median(errs)), 4) if errs else None,
            "search_dm_error_vs_truth_median": round(float(np.median(errs_search)), 4) if errs else None,
            "t0_error_samples_median": round(float(np.median(terrs)), 3) if terrs else None,
            "dm_within_0p1_of_judge": round(float(np.mean(jclose)), 4) if jclose else None,
            "dm_within_0p1_or_snr_tie": round(float(np.mean(jtie)), 4) if jtie else None,
            "snr_evaluations_per_cutout_max": int(nev), "judge_evaluations_per_cutout_max": int(njev)},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


HTR_NBIN = [512, 512, 1024, 2048]       # cutout length per row: the DM window's residual sweep fits
HTR_DMHALF = [1.0, 1.0, 2.4, 6.0]       # refinement DM window per row (> half the row's DM step)
assert len(HTR_DMHALF) == len(HTR_ROWS) and all(h > r[1] / 2 for h, r in zip(HTR_DMHALF, HTR_ROWS))
HTR_CUTS: list = []


def _lib_point_dtype():
    from transientx_amd import _lib

    return _lib.POINT_DTYPE


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seconds", type=float, default=300.0, help="observation length")
    ap.add_argument("--cpu-blocks", type=int, default=12,
                    help="timed ready blocks per row for the CPU baseline (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--kernel-events", choices=["timed", "separate", "off"], default="timed")
    ap.add_argument("--pipeline", choices=["direct", "graph"], default="graph",
                    help="search workload: a launch per kernel, or one hipGraph launch per block")
    ap.add_argument("--workload", choices=["search", "refine", "batch", "htr", "rfi"], default="search",
                    help="search: the production single-beam search (default); refine: the "
                         "candidate refinement grid on a batch of cutouts; batch: 64 uneven beams; "
                         "htr: the high-time-resolution combined run (search, then gradient refinement)")
    ap.add_argument("--htr-beams", type=int, default=16)
    ap.add_argument("--htr-group", type=int, default=4, help="beams resident (and searched together) at a time")
    ap.add_argument("--htr-seconds", type=float, default=30.0)
    ap.add_argument("--htr-dump", default="", help="write the injected bursts and the refined candidates (JSON)")
    ap.add_argument("--cutouts", type=int, default=4096)
    ap.add_argument("--beams", type=int, default=64)
    ap.add_argument("--batch-scale", type=float, default=1.0,
                    help="batch workload: beam lengths U[60, 600] s times this")
    ap.add_argument("--refine-batch", type=int, default=512)
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the plans' cost-model launch shapes instead of timing the alternatives")
    ap.add_argument("--exact-order", action="store_true",
                    help="the headline search with the boxcar in exact order (maps stored), e.g. to profile it")
    ap.add_argument("--maps", choices=["on", "off"], default="off",
                    help="keep the full S/N and width maps on the device (off: points only)")
    ap.add_argument("--no-legs", dest="legs", action="store_false",
                    help="skip the maps, host-fed and front-end legs")
    ap.add_argument("--no-error", dest="error", action="store_false",
                    help="skip the FP64 error figures")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB float32 streaming copy (y = x * 1) five times; "
                         "its FETCH_SIZE/WRITE_SIZE counters calibrate the traffic figures")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload in ("refine", "batch", "htr", "rfi"):
        if args.gpus != 1:
            print("bench.py measures one GPU", file=sys.stderr)
            sys.exit(2)
        {"refine": refine_bench, "batch": batch_bench, "htr": htr_bench, "rfi": rfi_bench}[args.workload](args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this build "
              "(every box has one MI355X; beams or ddplan rows would shard one per GPU).",
              file=sys.stderr)
        sys.exit(2)

    import torch

    torch.cuda.init()
    obs = build_workload(args.seconds)
    tsamp = obs.tsamp
    ndump = int(2.0 / tsamp)
    nblocks = obs.nsamp // ndump
    vwn = width_ladder_of(tsamp)
    thre = 7.0
    maps = args.maps == "on" or args.exact_order

    from transientx_amd.synth import gpu_generate

    shapes_u8 = autotune_rows(obs, ndump, vwn, "u8", args)
    probe = search_setup(obs, ndump, vwn, "u8", maps, shapes_u8)
    infos = [p.info for p in probe[1]]
    probe[0].close()
    nflush = max(i.offset // ndump + 1 for i in infos)
    log("rows: " + "; ".join(f"ndm {i.ndm} nsub {i.nsub} group {i.group} ndmc {i.ndmc} "
                              f"offset {i.offset}" for i in infos))
    t0 = time.time()
    data, flush = gpu_generate(obs, flush_blocks=nflush, ndump=ndump)
    torch.cuda.synchronize()
    log(f"generated {tuple(data.shape)} u8 + {nflush} flush blocks in {time.time() - t0:.1f} s")
    blocks = [data[b * ndump:(b + 1) * ndump] for b in range(nblocks)]
    fblocks = [flush[b * ndump:(b + 1) * ndump] for b in range(nflush)]
    units = sum(i.ndm for i in infos) * ndump * nblocks
    obs_s = nblocks * ndump * tsamp
    # the end-of-file flush: row r runs offset/ndump + 1 blocks of it
    # (src/singlepulse/dedisperse_fil.cpp:345-354), nflush the longest
    nleft = [i.offset // ndump + 1 for i in infos]

    # the headline: 8-bit blocks resident in HBM, one graph launch per block
    head = search_leg(obs, ndump, vwn, "u8", maps, shapes_u8, blocks + fblocks, thre, args,
                      nleft=nleft,
                      pipeline=args.pipeline, events=args.kernel_events, exact=args.exact_order)
    ms_step = head["ms_per_step"]
    value = units / (ms_step / 1e3)
    kt = head["kernels"]
    log(f"step {ms_step:.2f} ms, {value:.4g} DM-trial*samples/s, RTF {obs_s / (ms_step / 1e3):.1f}, "
        f"points per row {head['points']}")
    log("kernel ms/launches over timed steps: " + json.dumps(kt))

    # roofline of the dominant kernel
    alg = {"stage1": 0, "stage2": 0, "boxcar": 0}
    nb_all = nblocks + nflush
    for i, nl in zip(infos, nleft):
        for k, v in algorithmic_bytes(i, obs.nchans, nblocks + nl, 1, maps).items():
            alg[k] += v * args.steps
    # K0 reads the 8-bit block and writes the ring: 1 + 1 bytes per sample
    alg["transpose"] = 2 * obs.nchans * ndump * nb_all * args.steps
    dom = max(("stage1", "stage2", "boxcar", "transpose"), key=lambda k: kt[k][0])
    roof = None
    if kt[dom][0] > 0:
        achieved = alg[dom] / (kt[dom][0] / 1e3)
        roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved / 1e9, 1),
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                "traffic": load_traffic(dom),
                "launch_ms": round(kt[dom][0] / kt[dom][1], 4),
                "alg_bytes_per_launch": int(alg[dom] / kt[dom][1])}
    kernels = {k: {"ms_per_step": round(v[0] / args.steps, 3), "launches": v[1],
                   "GB_per_s_alg": (round(alg[k] / (v[0] / 1e3) / 1e9, 1) if v[0] else None)}
               for k, v in kt.items()}
    # LDS read bytes per launch against the measured rate of the read form
    # each kernel uses: stage 1 reads 8 bytes per 4 8-bit adds, stage 2 one
    # float per add; the boxcar
    # with maps two floats per output and width (the points-only boxcar
    # skips most of them, so no such figure is given for it)
    nbox = len(vwn)
    lds_b = {"stage1": sum(2.0 * i.ndmc * obs.nchans * ndump for i in infos),
             "stage2": sum(4.0 * i.ndm * i.nsub * i.nsamples for i in infos)}
    if maps:
        lds_b["boxcar"] = sum(8.0 * i.ndm * i.nsamples * nbox for i in infos)
    roof_lds = {k: {"achieved": round(v * nb_all * args.steps / (kt[k][0] / 1e3) / 1e12, 2),
                    "peak": LDS_PEAK[k] / 1e12, "unit": "TB/s",
                    "frac": round(v * nb_all * args.steps / (kt[k][0] / 1e3) / LDS_PEAK[k], 3)}
                for k, v in lds_b.items() if kt[k][0] > 0}

    legs = {}
    if args.legs:
        # the same search keeping the full S/N and width maps on the device
        other = search_leg(obs, ndump, vwn, "u8", not maps, shapes_u8, blocks + fblocks, thre, args,
                           nleft=nleft,
                           pipeline=args.pipeline, events="off")
        legs["maps_on" if not maps else "maps_off"] = {
            "ms_per_step": other["ms_per_step"], "rtf": round(obs_s / (other["ms_per_step"] / 1e3), 1),
            "value": round(units / (other["ms_per_step"] / 1e3), 1), "points_per_row": other["points"]}
        # host-fed: every block from pinned host memory, pushed asynchronously
        # (double-buffered staging on a copy stream, so one block crosses PCIe
        # while the previous one is searched); the step includes every H2D.
        # Direct launches: with one graph launch per block the copies ran
        # 36 GB/s against 55 GB/s, not overlapping the graphs
        hb = torch.empty((nblocks + nflush) * ndump * obs.nchans, dtype=torch.uint8, pin_memory=True)
        hb.view(-1, obs.nchans)[: nblocks * ndump].copy_(data[: nblocks * ndump])
        hb.view(-1, obs.nchans)[nblocks * ndump:].copy_(flush[: nflush * ndump])
        hblocks = [hb.view(-1, obs.nchans)[b * ndump:(b + 1) * ndump] for b in range(nblocks + nflush)]
        hf = search_leg(obs, ndump, vwn, "u8", maps, shapes_u8, hblocks, thre, args, nleft=nleft,
                        pipeline="direct",
                        events="off", host=True)
        h2d = (nblocks + nflush) * ndump * obs.nchans
        legs["host_fed"] = {"ms_per_step": hf["ms_per_step"], "rtf": round(obs_s / (hf["ms_per_step"] / 1e3), 1),
                            "value": round(units / (hf["ms_per_step"] / 1e3), 1), "pipeline": "direct",
                            "h2d_GB_per_step": round(h2d / 1e9, 2),
                            "h2d_GB_per_s": round(h2d / (hf["ms_per_step"] / 1e3) / 1e9, 1),
                            "points_equal_headline": hf["points"] == head["points"]}
        del hb, hblocks
        # the equalized front end: raw 8-bit blocks through the device front
        # end (td 1, fd 1, statistics of the first block) into the f32 search
        from transientx_amd import frontend

        shapes_f32 = autotune_rows(obs, ndump, vwn, "f32", args)
        off, gain = frontend.stats(blocks[0], 1, 1)
        fe = search_leg(obs, ndump, vwn, "f32", maps, shapes_f32, blocks + fblocks, thre, args,
                        nleft=nleft,
                        pipeline=args.pipeline, events="separate",
                        frontend=(off.cpu().numpy(), gain.cpu().numpy()))
        fe_k = fe["kernels"]
        fe_ms_block = fe_k["transpose"][0] / max(fe_k["transpose"][1], 1)
        legs["frontend"] = {"ms_per_step": fe["ms_per_step"], "rtf": round(obs_s / (fe["ms_per_step"] / 1e3), 1),
                            "value": round(units / (fe["ms_per_step"] / 1e3), 1),
                            "kernel_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in fe_k.items()},
                            "k_frontend_u8_ms_per_block": round(fe_ms_block, 4),
                            "k_frontend_u8_hbm_frac": round(5.0 * obs.nchans * ndump / (fe_ms_block / 1e3)
                                                            / HBM_PEAK, 4) if fe_ms_block else None,
                            "launch_shapes": shapes_f32}
        # the boxcar in exact_order mode (maps bit-identical to the oracle's
        # restatement of Boxcar::run), same blocks and pipeline; its
        # sequential chains use few CUs, and the graph's row branches overlap
        # them with the other rows' dedispersion
        ex = search_leg(obs, ndump, vwn, "u8", True, shapes_u8, blocks + fblocks, thre, args,
                        nleft=nleft,
                        pipeline=args.pipeline, events="separate", exact=True)
        legs["exact_order"] = {"ms_per_step": ex["ms_per_step"], "rtf": round(obs_s / (ex["ms_per_step"] / 1e3), 1),
                               "value": round(units / (ex["ms_per_step"] / 1e3), 1),
                               "pipeline": args.pipeline,
                               "kernel_ms_per_step": {k: round(v[0] / args.steps, 3)
                                                      for k, v in ex["kernels"].items()},
                               "points_per_row": ex["points"]}
        # the drop-in used unchanged: every ready block's S/N and width maps
        # copied to the host for Cluster::run's scan (414 MB per block over
        # the three rows)
        mh = search_leg(obs, ndump, vwn, "u8", True, shapes_u8, blocks + fblocks, thre, args,
                        nleft=nleft, pipeline="direct", events="off", maps_to_host=True)
        map_bytes = sum(8 * i.ndm * i.nsamples for i in infos)
        legs["maps_to_host"] = {"ms_per_step": mh["ms_per_step"], "rtf": round(obs_s / (mh["ms_per_step"] / 1e3), 1),
                                "value": round(units / (mh["ms_per_step"] / 1e3), 1),
                                "d2h_MB_per_block": round(map_bytes / 1e6, 1),
                                "points_equal_headline": mh["points"] == head["points"]}
        # batched vs plain on the batching workload at a tenth of its beam
        # lengths (the full lengths: --workload batch)
        legs["batch"] = batch_compare(batch_specs(64, 0.1), graphs=False, host=True)
        legs["batch"]["workload"] = "64 beams, 1024/2048/4096 ch, 6-60 s (a tenth of 60-600 s)"
        log("legs: " + json.dumps(legs))

    error = None
    if args.error:
        error = fp64_errors(obs, data, ndump, vwn, thre, shapes_u8)
        log("error vs FP64: " + json.dumps(error))

    cpu = None
    if args.cpu_blocks > 0:
        nthreads = args.cpu_threads or host_threads()
        rate, tcpu = cpu_baseline(obs, data, vwn, thre, nthreads, args.cpu_blocks)
        cpu = {"value": round(rate, 1), "unit": "DM-trial·samples/s", "cores": nthreads,
               "kind": "port", "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
               "sample": f"{args.cpu_blocks} ready 2 s blocks per ddplan row of the same "
                         f"workload ({tcpu:.1f} s wall on {nthreads} threads: every CPU this process "
                         f"may use), oracle streaming search (stage 1 + stage 2 + AVX2-variant "
                         f"boxcar + threshold), OpenMP"}
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "DM-trial·samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (generated on the device, resident in HBM)",
        "config": {
            "workload": "production single beam: 4096 ch 856-1712 MHz, tsamp 76.56 us, "
                        f"{obs_s:.1f} s in {nblocks} blocks of {ndump} (+{'/'.join(map(str, nleft))} flush per row), "
                        "overlap 0.1, ddplan 0-500/0.5/64sub, 500-1500/2/64sub, "
                        "1500-3000/5/128sub, 33 boxcar widths, thre 7",
            "path": "plain",
            "pipeline": args.pipeline,
            "maps": "stored" if maps else "points only (S/N maps not materialized; rows recomputed on demand)",
            "boxcar": "exact order (the reference's sequential FP32 passes)" if args.exact_order else "default",
            "launch_shapes": shapes_u8,
            "rows": [{"ndm": i.ndm, "nsub": i.nsub, "group": i.group, "ndmc": i.ndmc,
                      "offset": i.offset} for i in infos],
        },
        "rtf": round(obs_s / (ms_step / 1e3), 1),
        "points_per_row": head["points"],
        "kernels": kernels,
        "roofline": roof,
        "roofline_lds": roof_lds,
        "batched_over_plain": legs["batch"]["batched_over_plain"] if legs else None,
        "legs": legs or None,
        "error": error,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)

if __name__ == "__main__":
    main()
