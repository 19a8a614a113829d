"""GPU: the C++ drop-in classes (include/transientx_gpu.hpp), driven like
SinglePulse by tests/cpp/dropin_search, against the oracle.  The program gets
its width ladder through Boxcar::run (plan created with nbox = 0), pushes
float32 host blocks, and copies the full maps to Boxcar::mxS / mxwn; with
exact_order the maps are bit-identical to the oracle's restatement of the
AVX2 Boxcar::match."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from transientx_amd import _lib, width_ladder
from transientx_amd.synth import Observation, Pulse, generate

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parent / "cpp" / "dropin_search"


def read_blocks(path):
    raw = open(path, "rb").read()
    pos, out = 0, []
    while pos < len(raw):
        b, counter, ndm, N = np.frombuffer(raw, np.int64, 4, pos)
        pos += 32
        n = int(ndm * N)
        S = np.frombuffer(raw, np.float32, n, pos).reshape(ndm, N)
        pos += 4 * n
        W = np.frombuffer(raw, np.int32, n, pos).reshape(ndm, N)
        pos += 4 * n
        npts = int(np.frombuffer(raw, np.int64, 1, pos)[0])
        pos += 8
        pts = np.frombuffer(raw, _lib.POINT_DTYPE, npts, pos)
        pos += npts * _lib.POINT_DTYPE.itemsize
        out.append((int(b), int(counter), S, W, pts))
    return out


def test_dropin_classes_match_oracle(tmp_path):
    assert BIN.exists(), "tests/cpp/dropin_search is built by transientx_amd.build"
    C, ndump, nblocks, tsamp = 64, 512, 6, 1e-3
    obs = Observation(nchans=C, fbot=1200, ftop=1500, tsamp=tsamp, nsamp=ndump * nblocks, sd=10,
                      seed=5, pulses=[Pulse(dm=37.0, t0=1.9, fwhm=6e-3, snr=40)])
    x = generate(obs)
    rng = np.random.default_rng(9)
    xf = ((x.astype(np.float32) - 128) * np.float32(0.37)
          + rng.normal(0, 0.01, x.shape).astype(np.float32))
    xf.tofile(tmp_path / "data.f32")
    np.asarray(obs.freqs, np.float64).tofile(tmp_path / "freqs.f64")
    args = [str(BIN), str(tmp_path / "data.f32"), str(tmp_path / "freqs.f64"),
            str(tmp_path / "out.bin"), str(C), str(ndump), str(nblocks), repr(tsamp),
            "0", "2", "40", "0.25", "8", "1e-3", "0.05", "0.1", "7", "1"]
    print("run", " ".join(args[4:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    blocks = read_blocks(tmp_path / "out.bin")
    op = oracle.Plan(freqs=obs.freqs, tsamp=tsamp, ndump=ndump, dms=0.0, ddm=2.0, ndm=40,
                     overlap=0.25, nsub=8)
    info = op.info()
    ready = -(-(info.offset - info.noverlap + ndump) // ndump)
    assert [b[0] for b in blocks] == list(range(ready, nblocks + 1))
    vwn = width_ladder(tsamp, 1e-3, 0.05, 0.1)
    for b, counter, S, W, pts in blocks:
        assert counter == b * ndump
        tim = op.dedisperse(xf, op.window_start(b))
        rS, rW = oracle.boxcar(tim, vwn, variant=0, use_fma=True)
        assert np.array_equal(S.view(np.uint32), rS.view(np.uint32)), f"mxS block {b}"
        assert np.array_equal(W, rW), f"mxwn block {b}"
        idm, isamp, s, wn = oracle.threshold(rS, rW, 7.0)
        assert np.array_equal(pts["idm"], idm) and np.array_equal(pts["isamp"], isamp)
        assert np.array_equal(pts["S"], s) and np.array_equal(pts["wn"], wn)
    assert sum(len(b[4]) for b in blocks) > 0


def test_reference_call_sequence_with_savetim(tmp_path):
    """tests/cpp/dropin_calls runs the reference's own call sequence on the
    drop-in (prepare, preparedump, run / boxcar.run / cache / rundump per
    block, the driver's flush blocks and modifynblock at the end) with the
    pulsarx format: the dump file exists and its header's nblock counts the
    blocks written."""
    import struct

    binary = Path(__file__).resolve().parent / "cpp" / "dropin_calls"
    if not binary.exists():
        pytest.fail("tests/cpp/dropin_calls is not built (python -m transientx_amd.build)")
    root = tmp_path / "J0000-00"
    print("case dropin_calls", flush=True)
    r = subprocess.run([str(binary), str(root)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "dropin_calls ok" in r.stdout
    files = list(tmp_path.glob("J0000-00_DM*.dedisp"))
    assert len(files) == 1
    buf = files[0].read_bytes()
    key = b"nblock"
    at = buf.index(key) + len(key)
    nblock = struct.unpack_from("<i", buf, at)[0]
    hdr_end = buf.index(b"HEADER_END") + len(b"HEADER_END")
    assert nblock >= 1
    assert len(buf) - hdr_end == nblock * 8 * 256  # [ndm 8][ndump 256] int8 per block

