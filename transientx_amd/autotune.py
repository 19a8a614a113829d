"""Launch-shape autotuner for one ddplan row on the device.

The plan's built-in cost model picks the stage-1 tile (coarse DMs per
workgroup, output quads per thread) and the stage-2 tile (fine DMs per wave,
samples per lane, LDS per workgroup).  The autotuner times the
alternatives on synthetic 8-bit blocks of the row's real shape, stage by
stage (the two stages are independent), and returns the fastest; shapes
never change results, only speed.  Results are cached per (device, row
shape) in a JSON file ($TX_AUTOTUNE_CACHE or ~/.cache/transientx_amd/autotune.json).
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

from . import _lib
from .search import Beam, Plan

S1_CANDIDATES = [(16, 2), (16, 1), (8, 2), (8, 1), (4, 2)]
# float rows take 4x the LDS of 8-bit rows (and 8 coarse DMs at most): narrower tiles
S1_CANDIDATES_F32 = [(8, 2), (8, 1), (4, 2), (4, 1), (2, 2), (2, 1)]
# per-workgroup stage-2 LDS budgets: one, two, three or four workgroups per
# 160 KB CU (clamped to the device's per-workgroup maximum)
S2_LDS = (163840, 81920, 53248, 40960)


def _cache_path() -> Path:
    p = os.environ.get("TX_AUTOTUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "transientx_amd" / "autotune.json"


def _key(nchans, freqs, tsamp, ndump, overlap, dms, ddm, ndm, nsub, device_name):
    f = np.asarray(freqs, np.float64)
    return (f"{device_name}|{nchans}|{f.min():.6f}|{f.max():.6f}|{tsamp:.9g}|{ndump}|{overlap}|"
            f"{dms}|{ddm}|{ndm}|{nsub}")


def _time_stage(nchans, freqs, tsamp, ndump, row, vwn, overlap, shapes, data, stage, nblocks, dtype="u8"):
    dms, ddm, ndm, nsub = row
    beam = Beam(nchans, ndump, dtype=dtype, zero_point=128)
    try:
        plan = Plan(beam, freqs=freqs, tsamp=tsamp, dms=dms, ddm=ddm, ndm=ndm, vwn=vwn, overlap=overlap,
                    nsub=nsub, store_maps=False, shapes=shapes)
    except _lib.TxError:
        beam.close()
        return None
    try:
        for b in range(nblocks + 1):
            if b == 1:  # the first block pays first-use costs
                beam.sync()
                beam.timing()
                beam.set_timing(True)
            beam.push(data[b % len(data)])
            plan.dedisp_run()
        beam.set_timing(False)
        t = beam.timing()[stage]
        return t[0] / max(t[1], 1)
    finally:
        beam.close()


def tune_row(nchans, freqs, tsamp, ndump, row, vwn, overlap=0.0, nblocks=8, use_cache=True,
             verbose=False, dtype="u8"):
    """Fastest launch shapes for a row (dms, ddm, ndm, nsub) of 8-bit
    (dtype "u8") or float (dtype "f32", a different stage-1 kernel) data:
    a dict of PlanDesc overrides (s1_ncc, s1_q, s2_shape, s2_lds)."""
    import torch

    # the marketing name can read empty (e.g. under a profiler); the ISA and
    # CU count identify the device kind
    props = torch.cuda.get_device_properties(0)
    name = f"{getattr(props, 'gcnArchName', '')}/{props.multi_processor_count}CU"
    key = _key(nchans, freqs, tsamp, ndump, overlap, *row, name) + ("" if dtype == "u8" else f"|{dtype}")
    path = _cache_path()
    cache = {}
    if use_cache and path.exists():
        try:
            cache = json.loads(path.read_text())
        except ValueError:
            cache = {}
        if key in cache:
            return dict(cache[key]["shapes"])
    g = torch.Generator(device="cuda")
    g.manual_seed(12345)
    data = [(torch.randn((ndump, nchans), generator=g, device="cuda") * 12 + 128).round().clamp(0, 255)
            .to(torch.uint8) for _ in range(2)]
    if dtype != "u8":
        data = [d.float() - 128.0 for d in data]
    res = {}
    best1, t1 = {}, None
    for ncc, q in (S1_CANDIDATES if dtype == "u8" else S1_CANDIDATES_F32):
        t = _time_stage(nchans, freqs, tsamp, ndump, row, vwn, overlap, {"s1_ncc": ncc, "s1_q": q}, data,
                        "stage1", nblocks, dtype)
        res[f"s1 {ncc},{q}"] = t
        if t is not None and (t1 is None or t < t1):
            best1, t1 = {"s1_ncc": ncc, "s1_q": q}, t
    nshape = _lib.lib().tx_stage2_shapes(None, 0)
    best2, t2 = {}, None
    for sh in range(1, nshape + 1):
        for lds in S2_LDS:
            t = _time_stage(nchans, freqs, tsamp, ndump, row, vwn, overlap, {"s2_shape": sh, "s2_lds": lds},
                            data, "stage2", nblocks, dtype)
            res[f"s2 {sh},{lds}"] = t
            if t is not None and (t2 is None or t < t2):
                best2, t2 = {"s2_shape": sh, "s2_lds": lds}, t
    shapes = {**best1, **best2}
    if verbose:
        print(json.dumps({"row": list(row), "times_ms": res, "best": shapes}), flush=True)
    if use_cache:
        cache[key] = {"shapes": shapes, "stage1_ms": t1, "stage2_ms": t2}
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            path.write_text(json.dumps(cache, indent=1))
        except OSError:
            pass
    return shapes
