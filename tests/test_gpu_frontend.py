"""The front end on the device for raw 8-bit blocks (tx_frontend_u8,
tx_frontend_stats, tx_beam_set_frontend): the declared Downsample + Equalize
that SinglePulse::run applies before dedispersion
(src/singlepulse/singlepulse.cpp:289-291).  Outputs and statistics are
bit-identical to the oracle; a beam fed raw blocks through its front end
gives the same dedispersed series and points as an f32 beam fed the
oracle's front-end output, through direct, host-staged and graph pushes."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("nsi,nchi,td,fd", [(1000, 64, 1, 1), (1003, 130, 2, 1), (517, 96, 3, 4),
                                           (256, 2048, 1, 2), (77, 5, 8, 2)])
def test_frontend_and_stats_bit_identical(nsi, nchi, td, fd):
    import torch

    from transientx_amd import frontend

    rng = np.random.default_rng(nsi + nchi)
    x = rng.normal(128, 20, size=(nsi, nchi)).clip(0, 255).round().astype(np.uint8)
    x[:, 0] = 9  # a constant channel (gain 1)
    say("case frontend nsi", nsi, "nchi", nchi, "td", td, "fd", fd)
    g = torch.from_numpy(x).cuda()
    off, gain = frontend.stats(g, td, fd)
    ro, rg = oracle.frontend_stats(x, td, fd)
    assert np.array_equal(off.cpu().numpy().view(np.uint32), ro.view(np.uint32))
    assert np.array_equal(gain.cpu().numpy().view(np.uint32), rg.view(np.uint32))
    mask = (rng.random(nchi // fd) > 0.1).astype(np.uint8)
    out = frontend.apply(g, td, fd, off, gain, chan_mask=mask).cpu().numpy()
    ref = oracle.frontend_u8(x, td, fd, ro, rg, mask)
    assert out.shape == ref.shape
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("td,fd", [(2, 2), (1, 1)])
def test_beam_frontend_equals_f32_beam(td, fd):
    """Also device blocks off a 4-byte boundary, through direct and graph
    pushes (the kernel reads bytes, so any address serves), and the front
    end in stage 1's load (an 8-bit ring from the first block on, whose
    samples before the stream start must read 0)."""
    import torch

    from transientx_amd import Beam, Plan, width_ladder
    from transientx_amd.cutout import group_freqs
    from transientx_amd.synth import Observation, Pulse, generate

    nchi, ndump, nb = 128, 96, 14
    nch = nchi // fd
    obs = Observation(nchans=nchi, fbot=1200, ftop=1500, tsamp=1e-3 / td, nsamp=ndump * td * nb, sd=10, seed=5,
                      pulses=[Pulse(dm=120.0, t0=0.6, fwhm=5e-3, snr=60)])
    raw = generate(obs)
    freqs = group_freqs(obs.freqs, fd)
    tsamp = obs.tsamp * td
    off, gain = oracle.frontend_stats(raw[: ndump * td], td, fd)
    mask = np.ones(nch, np.uint8)
    mask[[3, 40]] = 0
    vwn = width_ladder(tsamp, tsamp, 0.03, 0.1)
    rows = [(0.0, 2.0, 48, 8), (90.0, 3.0, 32, 16)]
    blocks_raw = [raw[b * ndump * td:(b + 1) * ndump * td] for b in range(nb)]
    blocks_fe = [oracle.frontend_u8(r, td, fd, off, gain) for r in blocks_raw]

    def run(mode):
        beam = Beam(nch, ndump, dtype="f32", chan_mask=mask)
        plans = [Plan(beam, freqs=freqs, tsamp=tsamp, dms=d, ddm=dd, ndm=n, vwn=vwn, overlap=0.1, nsub=ns,
                      store_maps=False) for d, dd, n, ns in rows]
        try:
            if mode != "f32":
                beam.set_frontend(nchi, td, fd, off, gain, in_load=mode.endswith("in_load"))
            for p in plans:
                p.session_begin(1 << 16)
            if mode.startswith("graph"):
                beam.build_graph(7.0)
            series = []
            keep = []  # device blocks stay allocated while the beam's stream may read them
            for b in range(nb):
                if mode == "f32":
                    blk = torch.from_numpy(blocks_fe[b]).cuda()
                elif mode.startswith("host"):
                    blk = blocks_raw[b]
                elif mode.endswith("unaligned"):
                    raw_b = np.ascontiguousarray(blocks_raw[b])
                    store = torch.empty(raw_b.size + 4, dtype=torch.uint8, device="cuda")
                    blk = store[1:1 + raw_b.size].view(raw_b.shape)
                    blk.copy_(torch.from_numpy(raw_b))
                    assert blk.data_ptr() % 4 == 1
                else:
                    blk = torch.from_numpy(np.ascontiguousarray(blocks_raw[b])).cuda()
                keep.append(blk)
                if mode.startswith("graph"):
                    beam.graph_push(blk)
                    continue
                beam.push(blk)
                for p in plans:
                    p.dedisp_run()
                    p.boxcar_run(7.0)
                series.append(plans[1].timdata(5))
            out = []
            for p in plans:
                pts = p.session_end()
                out.append(pts[np.lexsort((pts["isamp"], pts["idm"], pts["block"]))])
            beam.sync()
            return out, series
        finally:
            beam.close()

    say("case beam frontend td", td, "fd", fd, "mode f32")
    ref, ref_series = run("f32")
    assert sum(len(a) for a in ref) > 0
    for mode in ("device", "host", "graph", "device_unaligned", "graph_unaligned", "device_in_load",
                 "host_in_load", "graph_in_load"):
        say("case beam frontend mode", mode)
        got, series = run(mode)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)
        for a, b in zip(ref_series, series):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    say("points", [len(a) for a in ref])


def test_frontend_argument_errors():
    import torch

    from transientx_amd import Beam
    from transientx_amd._lib import TxError

    say("case frontend errors")
    u8 = Beam(64, 32)
    try:
        with pytest.raises(TxError):
            u8.set_frontend(128, 1, 2, np.zeros(64), np.ones(64))  # not an f32 beam
    finally:
        u8.close()
    b = Beam(64, 32, dtype="f32")
    try:
        with pytest.raises(TxError):
            b.set_frontend(100, 1, 2, np.zeros(64), np.ones(64))  # 100/2 != 64
        with pytest.raises(TxError):
            b.set_frontend(128 * 4, 8, 4, np.zeros(64), np.ones(64))  # 512/4 != 64
        with pytest.raises(TxError):
            b.set_frontend(64 * 257, 1, 257, np.zeros(64), np.ones(64))  # td*fd > 256
        b.set_frontend(128, 3, 2, np.zeros(64), np.ones(64))
        with pytest.raises(TxError):
            b.push(torch.zeros((32, 128), dtype=torch.uint8, device="cuda"))  # needs 32*3 raw samples
        b.push(torch.zeros((96, 128), dtype=torch.uint8, device="cuda"))
        b.set_frontend(0, 1, 1, None, None)  # off again: f32 blocks of 64 channels
        b.push(torch.zeros((32, 64), dtype=torch.float32, device="cuda"))
        b.sync()
    finally:
        b.close()
    # in the stage-1 load with td = fd = 1 and no RFI filters at the first
    # push: the ring keeps the raw 8-bit samples and stage 1 applies the
    # front end, so the front end can no longer be switched off, resampled
    # or followed by RFI filters
    b = Beam(64, 32, dtype="f32")
    try:
        b.set_frontend(64, 1, 1, np.zeros(64), np.ones(64), in_load=True)
        b.push(torch.zeros((32, 64), dtype=torch.uint8, device="cuda"))
        with pytest.raises(TxError):
            b.set_frontend(64, 1, 1, np.zeros(64), np.ones(64), in_load=False)  # after the first push
        with pytest.raises(TxError):
            b.set_frontend(0, 1, 1, None, None)
        with pytest.raises(TxError):
            b.set_frontend(128, 1, 2, np.zeros(64), np.ones(64))
        with pytest.raises(TxError):
            b.set_rfi([("zdot",)])
        b.set_frontend(64, 1, 1, np.ones(64), np.ones(64))  # new offsets and gains
        b.push(torch.zeros((32, 64), dtype=torch.uint8, device="cuda"))
        b.sync()
    finally:
        b.close()


@pytest.mark.parametrize("td,fd,source", [(1, 1, "device"), (2, 2, "host"), (1, 1, "f32")])
def test_beam_rfi_chain_matches_oracle(td, fd, source):
    """The front end's RFI filters on the search path (tx_beam_set_rfi): raw
    8-bit blocks through the device front end, or f32 blocks with masked
    channels, then zdot and kadaneF in the DataBuffer layout, then the
    search.  Every block's dedispersed series are bit-exact against the
    oracle's front end, Candidate::zdot and Candidate::kadaneF of each block
    followed by the oracle's dedispersion."""
    import torch

    from transientx_amd import Beam, Plan, width_ladder
    from transientx_amd.cutout import group_freqs
    from transientx_amd.synth import Observation, Pulse, generate

    nchi, ndump, nb = 128, 256, 8
    nch = nchi // fd
    obs = Observation(nchans=nchi, fbot=1200, ftop=1500, tsamp=1e-3 / td, nsamp=ndump * td * nb, sd=10, seed=15,
                      pulses=[Pulse(dm=80.0, t0=0.9, fwhm=5e-3, snr=50)], rfi_frac=0.03, rfi_amp=4.0,
                      impulses=3)
    raw = generate(obs)
    freqs = group_freqs(obs.freqs, fd)
    tsamp = obs.tsamp * td
    off, gain = oracle.frontend_stats(raw[: ndump * td], td, fd)
    mask = np.ones(nch, np.uint8)
    mask[[2, 33]] = 0
    kw = oracle.kadane_wnlimit(4, 1e-2, tsamp, 2)
    ops = [("zdot",), ("kadaneF", 2, 2, 4.0, kw)]
    blocks_raw = [raw[b * ndump * td:(b + 1) * ndump * td] for b in range(nb)]
    # the oracle: front end, then each filter on the block in the Candidate layout
    clean = []
    for r in blocks_raw:
        x = oracle.frontend_u8(r, td, fd, off, gain, mask)
        c = np.ascontiguousarray(x.T)
        c = oracle.zdot(c)
        c = oracle.kadanef(c, 2, 2, 4.0, kw)
        clean.append(np.ascontiguousarray(c.T))
    xo = np.concatenate(clean)
    op = oracle.Plan(freqs=freqs, tsamp=tsamp, ndump=ndump, dms=0.0, ddm=2.0, ndm=48, overlap=0.1, nsub=8)
    vwn = width_ladder(tsamp, tsamp, 0.03, 0.1)
    say("case rfi chain td", td, "fd", fd, "source", source)
    beam = Beam(nch, ndump, dtype="f32", chan_mask=mask)
    try:
        plan = Plan(beam, freqs=freqs, tsamp=tsamp, dms=0.0, ddm=2.0, ndm=48, vwn=vwn, overlap=0.1, nsub=8,
                    store_maps=False)
        if source != "f32":
            beam.set_frontend(nchi, td, fd, off, gain)
        beam.set_rfi(ops)
        keep = []
        for b in range(nb):
            if source == "device":
                blk = torch.from_numpy(np.ascontiguousarray(blocks_raw[b])).cuda()
            elif source == "host":
                blk = np.ascontiguousarray(blocks_raw[b])
            else:
                # the f32 block of the front end without the mask: the chain zeroes masked channels
                blk = torch.from_numpy(oracle.frontend_u8(blocks_raw[b], td, fd, off, gain)).cuda()
            keep.append(blk)
            beam.push(blk)
            plan.dedisp_run()
            tim = np.stack([plan.timdata(i) for i in range(48)])
            ref = op.dedisperse(xo, op.window_start(b + 1))
            assert np.array_equal(tim.view(np.uint32), ref.view(np.uint32)), f"block {b + 1}"
        with pytest.raises(RuntimeError):
            plan.session_begin(1 << 10)
            beam.build_graph(7.0)  # graph pipelines refuse a beam with filters
    finally:
        beam.close()
