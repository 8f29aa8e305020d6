"""The EXACT pipeline's second seed round (DTW and Frechet): after the tight
filter, each query's 32 candidates of smallest bound are refined against the
first threshold, the k-th best of both rounds' values becomes the main
refine's threshold, the first round's seeds leave the candidates and join
the final top-k. The engine takes this path for batches of at least 2^40
cell-equivalents; TS_TWO_ROUND=1 forces it here on small batches, so it runs
against the oracle over random cases (bands, k up to 100 with k > N, a k per
query, graph replay), and TS_TWO_ROUND=0 keeps it off for the same cases."""
import numpy as np
import pytest

import cudaruntime1_amd as ts
import oracle
from helpers import say

pytestmark = pytest.mark.gpu

CODES = {"dtw": oracle.DTW, "frechet": oracle.FRECHET}


def cases(seed, n):
    rng = np.random.default_rng(seed)
    for t in range(n):
        meas = ["dtw", "frechet"][int(rng.integers(0, 2))]
        N = int(rng.choice([1, 5, 40, 300, 2500]))
        Q = int(rng.integers(1, 20))
        lo = int(rng.integers(1, 80))
        hi = lo + int(rng.integers(0, 400))
        dl = rng.integers(lo, hi + 1, N).astype(np.int32)
        ql = rng.integers(lo, hi + 1, Q).astype(np.int32)
        sigma = float(rng.choice([1e-3, 1e-2]))
        dseed, qseed = int(rng.integers(1, 1 << 30)), int(rng.integers(1, 1 << 30))
        k = int(rng.choice([1, 3, 10, 50, 100]))
        kw = {}
        if meas == "dtw" and rng.random() < 0.4:
            kw["band"] = int(rng.choice([2, 10, -10, -40]))
        graphs = bool(rng.random() < 0.4)
        multi = bool(rng.random() < 0.3)
        yield t, meas, dl, ql, sigma, dseed, qseed, k, kw, graphs, multi


@pytest.fixture(scope="module")
def engines():
    ts.init(0)
    return {False: ts.Engine(), True: ts.Engine(graphs=True)}


@pytest.mark.parametrize("force", ["1", "0"])
def test_two_round_knn_matches_oracle(engines, monkeypatch, force):
    monkeypatch.setenv("TS_TWO_ROUND", force)
    rng = np.random.default_rng(5)
    for t, meas, dl, ql, sigma, dseed, qseed, k, kw, graphs, multi in cases(77, 40):
        say("two-round knn", force, t, meas, len(dl), len(ql), int(dl.min()), int(dl.max()), k, kw, graphs, multi)
        dx, do = ts.gen_walks(dseed, dl, sigma)
        qx, qo = ts.gen_walks(qseed, ql, sigma)
        db, qs = ts.TrajSet(dx, do), ts.TrajSet(qx, qo)
        p = ts.params(meas, mode=ts.EXACT, k=k, **kw)
        if multi:
            kq = rng.integers(1, k + 1, len(ql)).astype(np.int32)
            for rep in range(3 if graphs else 1):
                idx, val, _ = engines[graphs].knn_multi_k(db, qs, p, kq)
                for i, kk in enumerate(kq):
                    prm = oracle.params(CODES[meas], ts.EXACT, k=int(kk), kc=int(kk), band=p.band)
                    want = oracle.knn(prm, dx, do, qx[qo[i]:qo[i + 1]], qo[i:i + 2] - qo[i])
                    assert np.array_equal(idx[i, :kk], want[0][0]), (t, i, rep)
                    assert np.array_equal(val[i, :kk].view(np.int32), want[1][0].view(np.int32)), (t, i, rep)
        else:
            prm = oracle.params(CODES[meas], ts.EXACT, k=k, kc=p.kc, band=p.band)
            want = oracle.knn(prm, dx, do, qx, qo)
            before = engines[graphs].graph_info()
            for rep in range(3 if graphs else 1):
                idx, val, _ = engines[graphs].knn(db, qs, p)
                assert np.array_equal(idx, want[0]), (t, meas, k, kw, graphs, rep)
                assert np.array_equal(val.view(np.int32), want[1].view(np.int32)), (t, meas, k, kw, graphs, rep)
            if graphs:
                # the second call records and the third replays (a recording
                # that fails runs eagerly, with the same results)
                after = engines[graphs].graph_info()
                assert after["captures"] > before["captures"] and after["replays"] > before["replays"], (t, after)


def test_two_round_runs_its_second_round(monkeypatch):
    """With the path forced, the DP statistics show the second round's launch
    label; without it, none; the results agree."""
    ts.init(0)
    dl = np.full(1000, 100, np.int32)  # 20 x 1000 pairs: past the all-pairs path for small problems
    dx, do = ts.gen_walks(3, dl, 1e-2)
    qx, qo = ts.gen_walks(4, np.full(20, 100, np.int32), 1e-2)
    db, qs = ts.TrajSet(dx, do), ts.TrajSet(qx, qo)
    p = ts.params("dtw", mode=ts.EXACT, k=5)
    for force, expect in (("1", True), ("0", False)):
        monkeypatch.setenv("TS_TWO_ROUND", force)
        e = ts.Engine(timing=True)
        say("two-round label", force)
        got = e.knn(db, qs, p)
        names = [n for n in e.kernel_stats()]
        assert any(n.startswith("dtw_wavefront_seed2") for n in names) == expect, names
        prm = oracle.params(oracle.DTW, ts.EXACT, k=5, kc=p.kc)
        want = oracle.knn(prm, dx, do, qx, qo)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1].view(np.int32), want[1].view(np.int32))
