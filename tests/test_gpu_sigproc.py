"""8-bit SIGPROC ingest on the search path: a filterbank split over two
files, read block by block by the reader transientx_fil uses
(src/singlepulse/dedisperse_fil.cpp:278-283 --jump, :324 read_data) and
pushed into a beam (8-bit search, and the f32 search through the device
front end), gives the same dedispersed series and points as the same
samples pushed from memory."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("mode", ["u8", "frontend", "u8_pinned"])
def test_filterbank_ingest_equals_memory_blocks(tmp_path, mode):
    from transientx_amd import Beam, Plan, sigproc, width_ladder
    from transientx_amd.synth import Observation, Pulse, generate

    nch, ndump, nb = 256, 512, 9
    tsamp = 2e-4
    obs = Observation(nchans=nch, fbot=1100.0, ftop=1500.0, tsamp=tsamp, nsamp=ndump * nb + 300, sd=12, seed=11,
                      pulses=[Pulse(dm=150.0, t0=0.35, fwhm=3e-3, snr=40), Pulse(dm=320.0, t0=0.7, fwhm=6e-3,
                                                                                 snr=30)])
    raw = generate(obs)  # [nsamp][nchans] uint8, channel 0 the highest frequency
    # SIGPROC channel order: fch1 the first channel, foff the step
    fch1, foff = float(obs.freqs[0]), float(obs.freqs[1] - obs.freqs[0])
    cut = 2000
    paths = [tmp_path / "a.fil", tmp_path / "b.fil"]
    sigproc.write(paths[0], raw[:cut], tsamp=tsamp, fch1=fch1, foff=foff, tstart=59000.0, source_name="T")
    sigproc.write(paths[1], raw[cut:], tsamp=tsamp, fch1=fch1, foff=foff, tstart=59000.0 + cut * tsamp / 86400.0)
    jump = (0.01, 0.0)  # 50 samples left unread at the start
    skip = int(jump[0] / tsamp)
    vwn = width_ladder(tsamp, tsamp, 0.02, 0.1)
    rows = [(0.0, 2.0, 100, 16), (200.0, 4.0, 50, 16)]
    off, gain = oracle.frontend_stats(raw[skip:skip + ndump], 1, 1)
    say("case sigproc ingest", mode)

    def run(source):
        import torch

        f32 = mode == "frontend"
        beam = Beam(nch, ndump, dtype="f32" if f32 else "u8", zero_point=0 if f32 else 128)
        reader = None
        try:
            plans = [Plan(beam, freqs=obs.freqs, tsamp=tsamp, dms=d, ddm=dd, ndm=n, vwn=vwn, overlap=0.1, nsub=ns,
                          store_maps=False) for d, dd, n, ns in rows]
            if f32:
                beam.set_frontend(nch, 1, 1, off, gain)
            for p in plans:
                p.session_begin(1 << 16)
            if source == "file":
                reader = sigproc.FilReader(paths, jump=jump)
                assert reader.nsamples == len(raw) and np.allclose(reader.freqs, obs.freqs)
                blocks = reader.blocks(ndump, pinned=mode == "u8_pinned")
            else:
                n = (len(raw) - skip) // ndump
                blocks = (raw[skip + b * ndump:skip + (b + 1) * ndump] for b in range(n))
            series, keep, nblk = [], [], 0
            for blk in blocks:
                if mode == "u8_pinned" and source == "file":
                    keep.append(blk)  # unchanged until the next sync
                    beam.push_async(blk)
                else:
                    beam.push(np.ascontiguousarray(blk))
                for p in plans:
                    p.dedisp_run()
                    p.boxcar_run(7.0)
                series.append(plans[0].timdata(7))
                nblk += 1
            beam.sync()
            torch.cuda.synchronize()
            out = []
            for p in plans:
                pts = p.session_end()
                out.append(pts[np.lexsort((pts["isamp"], pts["idm"], pts["block"]))])
            return nblk, out, series
        finally:
            if reader is not None:
                reader.close()
            beam.close()

    n_ref, ref, ref_series = run("memory")
    n_got, got, series = run("file")
    assert n_ref == n_got == 9 and sum(len(a) for a in ref) > 0
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    for a, b in zip(ref_series, series):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    say("points", [len(a) for a in ref])
