"""GPU end to end: thresholded points -> clusters -> .cands lines, against
the oracle path (oracle S/N maps, oracle threshold scan, oracle cluster,
oracle fl/fh from the oracle's own sub-band series).

With exact_order the S/N maps are bit-identical to the oracle's, so the
points, the candidate tuples and every character of the .cands lines must
be identical too."""
import numpy as np
import pytest

import oracle
from transientx_amd import Beam, Plan, host, width_ladder
from transientx_amd.synth import Observation, Pulse, generate

pytestmark = pytest.mark.gpu

THRE = 7.0


def say(*a):
    print(*a, flush=True)


def expected_lines(cands, info, fl, fh, tstart, rootname, beam, plan_id, fname, num0):
    """CandPlot::plot's .cands text (src/singlepulse/candplot.cpp:32-110,420-433)."""
    order = sorted(range(len(cands)), key=lambda k: -cands[k][3])  # stable
    out, num = [], num0
    for iidx, k in enumerate(order):
        idm, isamp, wn, S = cands[k]
        mjd = tstart + (-info.offset + info.counter - info.ndump + int(isamp)) * info.tsamp / 86400.
        bmjd = tstart + (-info.offset + info.counter - info.ndump) * info.tsamp / 86400.
        png = f"{rootname}_{bmjd:.10f}_{beam}_{iidx + 1:02d}_{plan_id:02d}.png"
        num += 1
        out.append(f"{beam}\t{num}\t{mjd:.15f}\t{info.dms + int(idm) * info.ddm:.2f}\t"
                   f"{int(wn) * info.tsamp * 1000:.2f}\t{float(S):.1f}\t{float(fl[k]):.0f}\t"
                   f"{float(fh[k]):.0f}\t{png}\t{plan_id:02d}\t{fname}\n")
    return "".join(out), num


@pytest.mark.parametrize("minpts,drop", [(5, False), (0, True)])
def test_cands_lines_match_oracle_path(minpts, drop):
    C, ndump, nblocks = 64, 512, 7
    obs = Observation(nchans=C, fbot=1200, ftop=1500, tsamp=1e-3, nsamp=ndump * nblocks, sd=10,
                      seed=11, pulses=[Pulse(dm=37.0, t0=0.9, fwhm=6e-3, snr=40),
                                       Pulse(dm=61.0, t0=2.1, fwhm=2e-3, snr=25),
                                       Pulse(dm=20.0, t0=2.9, fwhm=20e-3, snr=30)])
    x = generate(obs)
    xo = x.astype(np.float32) - 128
    vwn = width_ladder(1e-3, 1e-3, 0.05, 0.1)
    kw = dict(freqs=obs.freqs, tsamp=obs.tsamp, dms=0.0, ddm=2.0, ndm=40, vwn=vwn, overlap=0.25,
              nsub=8, exact_order=True)
    op = oracle.Plan(freqs=obs.freqs, tsamp=obs.tsamp, ndump=ndump, dms=0.0, ddm=2.0, ndm=40,
                     overlap=0.25, nsub=8)
    fsub = np.asarray(obs.freqs, np.float64).reshape(8, C // 8).mean(axis=1)
    fcnt = np.full(8, C // 8, np.float64)
    say("case C", C, "ndump", ndump, "minpts", minpts, "drop", drop)
    beam = Beam(C, ndump)
    counter = host.CandCounter()
    num = 0
    nlines = 0
    try:
        plan = Plan(beam, **kw)
        assert np.allclose(plan.tables()["fsub"], fsub, rtol=0, atol=1e-9)
        for b in range(1, nblocks + 1):
            beam.push(x[(b - 1) * ndump:b * ndump])
            plan.dedisp_run()
            if not plan.boxcar_run(THRE):
                continue
            info = plan.info
            ckw = dict(dms=info.dms, ddm=info.ddm, fmax=info.fmax, fmin=info.fmin,
                       tsamp=info.tsamp, radius=1e-3, maxncand=100, minpts=minpts, drop=drop,
                       maxwn=info.maxwn)
            pts = plan.points()
            cands = host.cluster(pts, **ckw)
            # the oracle path
            tim, sub = op.dedisperse(xo, op.window_start(b), with_sub=True)
            rS, rW = oracle.boxcar(tim, vwn, variant=0, use_fma=True)
            idm, isamp, s, wn = oracle.threshold(rS, rW, THRE)
            assert np.array_equal(pts["idm"], idm) and np.array_equal(pts["isamp"], isamp)
            sel, _ = oracle.cluster(idm, isamp, s, wn, **ckw)
            want = [(idm[i], isamp[i], wn[i], s[i]) for i in sel]
            got = [(c["idm"], c["isamp"], c["wn"], c["S"]) for c in cands]
            assert got == want, f"block {b}"
            fl = np.zeros(len(want))
            fh = np.zeros(len(want))
            for k, (i, t, w, _) in enumerate(want):
                fl[k], fh[k] = oracle.cand_flfh(sub[i], fcnt, fsub, int(t), int(w))
            gfl, gfh = host.cand_flfh(plan, cands)
            assert np.array_equal(gfl, fl.astype(np.float32)) and np.array_equal(gfh, fh.astype(np.float32))
            text = host.cands_lines(plan, cands, tstart=60000.5, rootname="J0000-00",
                                    beam=host.beam_name(1), plan_id=1, fname="fake.fil",
                                    counter=counter)
            etext, num = expected_lines(want, info, fl, fh, 60000.5, "J0000-00", "cfbf00001", 1,
                                        "fake.fil", num)
            assert text == etext, f"block {b}"
            assert counter.value.value == num
            nlines += text.count("\n")
    finally:
        beam.close()
    assert nlines >= 2  # the bright pulses came out as candidates
