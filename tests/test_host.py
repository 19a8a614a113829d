"""Host side around the kernels (CPU only): ddplan/JSON rows, ndump,
clustering, kadane.  Cluster and kadane are checked against the oracle,
which is itself pinned to scipy connected components and brute force."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from transientx_amd import _lib, host

GOLD = Path(__file__).resolve().parent / "golden"


# ------------------------------------------------------------------ a2
def test_ddplan_text_rows_follow_the_file():
    text = (GOLD / "ddplan_example.txt").read_text()
    rows = host.parse_ddplan(text, dict(overlap=0.1, thre=7.5))
    # the file's five lines, in order (examples/ddplan.txt)
    want = [(1, 1, 0.0, 0.1, 3344), (2, 1, 334.4, 0.2, 1255), (4, 1, 585.2, 0.3, 1334),
            (8, 1, 985.4, 0.5, 1674), (16, 1, 1821.9, 1.0, 1179)]
    assert [(r.td, r.fd, r.dms, r.ddm, r.ndm) for r in rows] == want
    assert [r.id for r in rows] == [1, 2, 3, 4, 5]
    for r in rows:
        assert r.snrloss == np.float32(0.1) and r.maxw == np.float32(0.1)
        assert r.overlap == 0.1 and r.thre == np.float32(7.5)  # from the command line
        assert r.rfi == [] and r.zap == []


def test_ddplan_text_comments_blank_lines_and_rfi_accumulation():
    text = "\n".join([
        "# td fd dms ddm ndm snrloss maxw",
        "",
        "   1 1 0 1 100 0.1 0.02 kadaneF 8 4 zdot",
        "x 2 1 100 2 50 0.1 0.02",          # not starting with a digit: skipped
        "2\t1  100 2 50 0.2 0.05 zap 1100 1150 mask 4 2 bogus",
    ])
    rows = host.parse_ddplan(text, dict(rfi=[["zero"]]))
    assert len(rows) == 2
    assert (rows[0].td, rows[0].dms, rows[0].ndm) == (1, 0.0, 100)
    assert (rows[1].td, rows[1].dms, rows[1].ndm, rows[1].snrloss) == (2, 100.0, 50, np.float32(0.2))
    # one SinglePulse object is reused, so the items pile up row after row
    assert rows[0].rfi == [["zero"], ["kadaneF", "8", "4"], ["zdot"]]
    assert rows[1].rfi == rows[0].rfi + [["mask", "4", "2"]]
    assert rows[0].zap == [] and rows[1].zap == [(1100.0, 1150.0)]


def test_ddplan_without_rows_gives_the_command_line_row():
    rows = host.parse_ddplan("# nothing\n", dict(dms=5.0, ddm=0.5, ndm=1000))
    assert len(rows) == 1 and rows[0].id == 1
    assert (rows[0].dms, rows[0].ddm, rows[0].ndm) == (5.0, 0.5, 1000)


def test_ddplan_short_line_is_an_error():
    with pytest.raises(RuntimeError, match="ddplan line"):
        host.parse_ddplan("1 1 0 1 100 0.1\n")


def test_json_config_rows_follow_the_file():
    text = (GOLD / "transientx_config_example.json").read_text()
    cfg = json.loads(text)
    rows = host.parse_config(text)
    plan = cfg["ddplan"]
    assert len(rows) == len(plan) >= 2
    acc_rfi = []
    acc_zap = []
    for r, c in zip(rows, plan):
        dd, bx, cl, rf = c["subdedispersion"], c["boxcar"], c["clustering"], c["rfi"]
        assert (r.td, r.fd) == (c["downsample"]["td"], c["downsample"]["fd"])
        assert (r.dms, r.ddm, r.ndm, r.overlap) == (dd["dms"], dd["ddm"], dd["ndm"], dd["overlap"])
        assert r.minw == np.float32(bx["minw"]) and r.maxw == np.float32(bx["maxw"])
        assert r.snrloss == np.float32(bx["snrloss"]) and r.iqr == bx["iqr"]
        assert r.thre == np.float32(cl["thre"]) and r.radius == cl["radius"] / 1000.0
        assert (r.neighbors, r.maxncand, r.minpts, r.drop) == (
            cl["neighbors"], cl["maxncand"], cl["minpts"], cl["drop"])
        assert r.bswidth == np.float32(c["baseline"]["width"])
        assert r.widthlimit == rf["widthlimit"] and r.threKadaneF == np.float32(rf["threKadaneF"])
        for item in rf["rfilist"]:
            acc_rfi.append(list(item))
        for z in rf["zaplist"]:
            acc_zap.append((float(z[0]), float(z[-1])))
        assert r.rfi == acc_rfi and r.zap == acc_zap


def test_json_missing_key_is_an_error():
    cfg = json.loads((GOLD / "transientx_config_example.json").read_text())
    del cfg["ddplan"][0]["boxcar"]["maxw"]
    with pytest.raises(RuntimeError, match="boxcar.maxw"):
        host.parse_config(json.dumps(cfg))
    with pytest.raises(RuntimeError, match="JSON"):
        host.parse_config("{ not json")


def test_json_optional_nsub_and_group():
    cfg = json.loads((GOLD / "transientx_config_example.json").read_text())
    cfg["ddplan"][0]["subdedispersion"]["nsub"] = 32
    cfg["ddplan"][0]["subdedispersion"]["group"] = 4
    rows = host.parse_config(json.dumps(cfg))
    assert (rows[0].nsub, rows[0].group) == (32, 4)
    assert (rows[1].nsub, rows[1].group) == (32, 4)  # carried on, like every other field


@pytest.mark.parametrize("seglen,tsamp,tds", [
    (2.0, 98.304e-6, [1]), (2.0, 76.56e-6, [1, 1, 1]), (0.5, 16e-6, [1, 1, 1, 1]),
    (1.0, 64e-6, [1, 2, 4, 8, 16]), (2.0, 64e-6, [3, 4, 6]), (1.5, 1e-4, [5, 7]),
])
def test_ndump_is_whole_lcm_multiples(seglen, tsamp, tds):
    n = host.ndump(seglen, tsamp, tds)
    lcm = math.lcm(*tds)
    raw = int(np.float64(np.float32(seglen)) / tsamp)
    assert n % lcm == 0 and raw - lcm < n <= raw
    # the blocks of the 2 s / 76.56 us and 2 s / 98.304 us configurations
    if (seglen, tsamp) == (2.0, 76.56e-6):
        assert n == 26123
    if (seglen, tsamp) == (2.0, 98.304e-6):
        assert n == 20345


# ------------------------------------------------------------ a15/a16
def _points(rng, n, ndm, nsamp, blobs=6):
    """Thresholded-point-like sets: dense blobs around a few (DM, t)
    centres plus scattered points, in row-major (DM, sample) order."""
    idm, isamp = [], []
    for _ in range(blobs):
        c0, t0 = rng.integers(0, ndm), rng.integers(0, nsamp)
        k = rng.integers(3, max(4, n // blobs))
        idm += list(np.clip(c0 + rng.integers(-6, 7, k), 0, ndm - 1))
        isamp += list(np.clip(t0 + rng.integers(-30, 31, k), 0, nsamp - 1))
    m = max(0, n - len(idm))
    idm += list(rng.integers(0, ndm, m))
    isamp += list(rng.integers(0, nsamp, m))
    key = np.unique(np.array(idm, np.int64) * nsamp + np.array(isamp, np.int64))
    pts = np.zeros(key.size, _lib.POINT_DTYPE)
    pts["idm"] = key // nsamp
    pts["isamp"] = key % nsamp
    pts["S"] = (7 + rng.exponential(2.0, key.size)).astype(np.float32)
    pts["S"][rng.random(key.size) < 0.1] = np.float32(9.5)  # ties
    pts["wn"] = rng.choice([1, 2, 4, 8, 16, 32], key.size)
    return pts


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("minpts,drop,maxncand", [(0, False, 100), (5, False, 100),
                                                  (3, True, 100), (2, False, 4)])
def test_cluster_matches_oracle(ora, seed, minpts, drop, maxncand):
    rng = np.random.default_rng(seed)
    pts = _points(rng, 400 + 300 * seed, 200, 5000)
    kw = dict(dms=10.0, ddm=0.5, fmax=1500.0, fmin=1200.0, tsamp=98.304e-6, radius=1e-3)
    cands, labels = host.cluster(pts, maxncand=maxncand, minpts=minpts, drop=drop, maxwn=32,
                                 labels=True, **kw)
    want_idx, want_labels = ora.cluster(pts["idm"], pts["isamp"], pts["S"], pts["wn"],
                                        maxncand=maxncand, minpts=minpts, drop=drop, maxwn=32, **kw)
    assert np.array_equal(labels, want_labels)
    got = [(c["idm"], c["isamp"], c["wn"], c["S"]) for c in cands]
    want = [(pts["idm"][i], pts["isamp"][i], pts["wn"][i], pts["S"][i]) for i in want_idx]
    assert got == want
    for c in cands:
        assert c["npts"] == np.count_nonzero(labels == c["cluster"])


def test_cluster_empty_and_radius_floor(ora):
    e = np.zeros(0, _lib.POINT_DTYPE)
    assert host.cluster(e, dms=0, ddm=1, fmax=1500, fmin=1200, tsamp=1e-4, radius=1e-3).size == 0
    # two points 2 samples apart at one DM: a radius below 1.5 tsamp is raised to it,
    # which still separates them; 2.5 tsamp joins them
    p = np.zeros(2, _lib.POINT_DTYPE)
    p["isamp"] = [100, 102]
    p["S"] = [8.0, 9.0]
    kw = dict(dms=0, ddm=1, fmax=1500, fmin=1200, tsamp=1e-4, minpts=0)
    assert host.cluster(p, radius=0.0, **kw).size == 2
    c = host.cluster(p, radius=2.5e-4, **kw)
    assert c.size == 1 and c[0]["isamp"] == 102 and c[0]["npts"] == 2


def test_kadane_matches_brute_force(ora):
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 17, 64):
        for _ in range(20):
            a = rng.normal(0, 1, n).astype(np.float32)
            a[rng.random(n) < 0.2] = 0.0
            assert host.kadane(a) == ora.kadane(a)
            s, e = host.kadane(a)
            best = max(np.float32(a[i:j + 1].sum()) for i in range(n) for j in range(i, n))
            assert abs(np.float32(a[s:e + 1].sum()) - best) <= 1e-5 * (1 + abs(best))


def test_points_digest_is_order_free_and_sensitive():
    """batch.points_digest (the full-length batch bench compares candidate
    lists by it): equal for any order of the same points, different when a
    single field of one point changes, and additive over a split list."""
    from transientx_amd.batch import points_digest

    rng = np.random.default_rng(11)
    n = 5000
    a = np.zeros(n, _lib.POINT_EX_DTYPE)
    a["block"] = rng.integers(0, 300, n)
    a["idm"] = rng.integers(0, 1000, n)
    a["isamp"] = rng.integers(0, 28735, n)
    a["S"] = rng.uniform(7.0, 50.0, n).astype(np.float32)
    a["wn"] = rng.integers(1, 1128, n)
    c0, d0 = points_digest(a)
    assert c0 == n
    assert points_digest(a[rng.permutation(n)]) == (c0, d0)
    for f in ("block", "idm", "isamp", "wn"):
        b = a.copy()
        b[f][1234] += 1
        assert points_digest(b)[1] != d0, f
    b = a.copy()
    b["S"][77] = np.nextafter(b["S"][77], np.float32(np.inf))
    assert points_digest(b)[1] != d0
    # a list drained in parts sums to the whole
    _, d1 = points_digest(a[:1700])
    _, d2 = points_digest(a[1700:])
    assert (d1 + d2) % (1 << 64) == d0
    assert points_digest(a[:0]) == (0, 0)
