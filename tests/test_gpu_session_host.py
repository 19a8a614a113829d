"""Session lists in the caller's pinned host memory (tx_session_begin_host):
the boxcar stores each thresholded point straight into host memory, so the
points of a whole search reach the host while it runs.  The lists must be
the device lists' points exactly, through direct and graph pushes, across
sessions that reuse the same host buffer, and a list that fills up must
stop at its capacity (the bytes after it untouched) and report the
overflow."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def sorted_points(p):
    return p[np.lexsort((p["isamp"], p["idm"], p["block"]))]


def test_host_session_lists_equal_device_lists():
    import torch

    from transientx_amd import Beam, Plan, width_ladder
    from transientx_amd._lib import POINT_EX_DTYPE, TxError
    from transientx_amd.synth import Observation, Pulse, generate

    nch, ndump, nb = 128, 512, 8
    obs = Observation(nchans=nch, fbot=1200.0, ftop=1500.0, tsamp=1e-3, nsamp=ndump * nb, sd=10, seed=41,
                      impulses=2, pulses=[Pulse(dm=60.0, t0=1.5, fwhm=5e-3, snr=40),
                                          Pulse(dm=140.0, t0=2.9, fwhm=9e-3, snr=30)])
    x = torch.from_numpy(generate(obs)).cuda()
    blocks = [x[k * ndump:(k + 1) * ndump] for k in range(nb)]
    vwn = width_ladder(1e-3, 1e-3, 0.05, 0.1)
    rows = [(0.0, 2.0, 60, 8), (100.0, 4.0, 40, 16)]
    cap = 1 << 16
    rec = POINT_EX_DTYPE.itemsize
    hbufs = [torch.empty(cap * rec, dtype=torch.uint8, pin_memory=True) for _ in rows]

    def run(graph, host, thre=4.0, cap_=cap, repeat=1):
        beam = Beam(nch, ndump)
        try:
            plans = [Plan(beam, freqs=obs.freqs, tsamp=1e-3, dms=d, ddm=dd, ndm=n, vwn=vwn, overlap=0.1,
                          nsub=ns, store_maps=False) for d, dd, n, ns in rows]
            outs = []
            for _ in range(repeat):
                beam.reset()
                for p, h in zip(plans, hbufs):
                    p.session_begin(cap_, host=h if host else None)
                if graph:
                    beam.build_graph(thre)
                for blk in blocks:
                    if graph:
                        beam.graph_push(blk)
                    else:
                        beam.push(blk)
                        for p in plans:
                            p.dedisp_run()
                            p.boxcar_run(thre)
                outs.append([sorted_points(p.session_end()) for p in plans])
            return outs
        finally:
            beam.close()

    print("case host session direct/graph", flush=True)
    ref = run(False, False)[0]
    assert all(len(a) > 100 for a in ref)
    for graph in (False, True):
        got = run(graph, True, repeat=2)
        for g in got:  # the second session reuses the host lists
            for a, b in zip(ref, g):
                assert np.array_equal(a, b)
    print("points", [len(a) for a in ref], flush=True)

    # a list that fills up: the kernels stop at its capacity and the bytes
    # after it keep their sentinel
    small = 50
    for h in hbufs:
        h.fill_(0xA5)
    print("case host session overflow", flush=True)
    with pytest.raises(TxError):
        run(False, True, cap_=small)
    for h in hbufs:
        tail = h.numpy()[small * rec:]
        assert (tail == 0xA5).all()
    # a pageable numpy buffer is refused before anything runs
    beam = Beam(nch, ndump)
    try:
        p = Plan(beam, freqs=obs.freqs, tsamp=1e-3, dms=0.0, ddm=2.0, ndm=60, vwn=vwn, overlap=0.1, nsub=8,
                 store_maps=False)
        with pytest.raises(ValueError):
            p.session_begin(cap, host=torch.empty(cap * rec, dtype=torch.uint8))
    finally:
        beam.close()
