"""GPU batching: beams of uneven length and channel count searched together
(one stream per beam, rounds of one block per live beam) give exactly the
candidate lists of beam-by-beam runs, with blocks resident on the device or
pushed from pinned host memory (copies waited for, or queued)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batched_equals_plain():
    import torch

    from transientx_amd import width_ladder
    from transientx_amd.batch import BeamSearch, BeamSpec, collect, run_batched, run_plain
    from transientx_amd.synth import Observation, Pulse, generate

    specs = []
    for b, (C, nb) in enumerate([(64, 7), (128, 3), (256, 5), (64, 2), (128, 6), (256, 4)]):
        ndump = 512
        obs = Observation(nchans=C, fbot=1200, ftop=1500, tsamp=1e-3, nsamp=ndump * nb, sd=10, seed=30 + b,
                          pulses=[Pulse(dm=20.0 + 9 * b, t0=0.8 * nb * 0.512, fwhm=5e-3, snr=35)])
        x = torch.from_numpy(generate(obs)).cuda()
        rows = [(0.0, 2.0, 40, 8)] if C < 128 else [(0.0, 2.0, 40, 8), (80.0, 4.0, 30, 16)]
        specs.append(BeamSpec(freqs=obs.freqs, tsamp=1e-3, ndump=ndump, nblocks=nb, rows=rows,
                              vwn=width_ladder(1e-3, 1e-3, 0.05, 0.1), overlap=0.25,
                              blocks=[x[k * ndump:(k + 1) * ndump] for k in range(nb)],
                              host_blocks=[x[k * ndump:(k + 1) * ndump].cpu().pin_memory() for k in range(nb)]))
    print("beams", [(len(s.freqs), s.nblocks) for s in specs], flush=True)
    searches = [BeamSearch(s, cap=1 << 18) for s in specs]
    try:
        run_plain(searches, 7.0)
        plain = collect(searches)
        run_batched(searches, 7.0)
        batched = collect(searches)
        run_plain(searches[::-1], 7.0)
        again = collect(searches)
        # host-fed: synchronous copies (the drop-in's push) and queued copies
        run_plain(searches, 7.0, "host")
        hsync = collect(searches)
        run_batched(searches, 7.0, "host_async")
        hasync = collect(searches)
    finally:
        for s in searches:
            s.close()
    total = 0
    for pa, pb, pc, pd, pe in zip(plain, batched, again, hsync, hasync):
        for a, b, c, d, e in zip(pa, pb, pc, pd, pe):
            assert np.array_equal(a, b) and np.array_equal(a, c)
            assert np.array_equal(a, d) and np.array_equal(a, e)
            total += len(a)
    assert total > 0
