"""Randomised k-NN parity: cases drawn from a fixed seed over every measure,
both modes, database sizes 1..2500 (k > N included), query counts up to 23,
length ranges 1..360, k in {1, 3, 10, 50, 100}, DTW bands (absolute and
percentage), LCSS epsilon and windows, and graph replay on and off (replayed
three times). Each case's top-k indices and values equal the oracle's bit for
bit."""
import numpy as np
import pytest

import cudaruntime1_amd as ts
import oracle
from helpers import say

pytestmark = pytest.mark.gpu

NAMES = ["dtw", "lcss", "frechet", "hausdorff", "euclidean"]
CODES = {"dtw": oracle.DTW, "lcss": oracle.LCSS, "frechet": oracle.FRECHET, "hausdorff": oracle.HAUSDORFF,
         "euclidean": oracle.EUCLIDEAN}


def cases(seed, n):
    rng = np.random.default_rng(seed)
    for t in range(n):
        meas = NAMES[rng.integers(0, 5)]
        mode = int(rng.integers(0, 2))
        N = int(rng.choice([1, 2, 7, 30, 300, 2500]))
        Q = int(rng.integers(1, 24))
        lo = int(rng.integers(1, 60))
        hi = lo + int(rng.integers(0, 300))
        if meas == "euclidean":
            fix = int(rng.integers(1, 100))
            dl = np.where(rng.random(N) < 0.7, fix, rng.integers(lo, hi + 1, N)).astype(np.int32)
            ql = np.full(Q, fix, np.int32)
        else:
            dl = rng.integers(lo, hi + 1, N).astype(np.int32)
            ql = rng.integers(lo, hi + 1, Q).astype(np.int32)
        sigma = float(rng.choice([1e-3, 1e-2]))
        dseed, qseed = int(rng.integers(1, 1 << 30)), int(rng.integers(1, 1 << 30))
        k = int(rng.choice([1, 3, 10, 50, 100]))
        kw = {}
        if meas == "dtw" and rng.random() < 0.5:
            kw["band"] = int(rng.choice([1, 5, 20, -5, -10, -50]))
        if meas == "lcss":
            kw["epsilon"] = float(rng.choice([1e-3, 5e-3, 2e-2]))
            if rng.random() < 0.5:
                kw["delta"] = int(rng.choice([0, 3, 16, 64]))
        graphs = bool(rng.random() < 0.5)
        if mode == ts.REF:
            kw["kc"] = max(k, int(rng.choice([10, 20, 64])))
        yield t, meas, mode, dl, ql, sigma, dseed, qseed, k, kw, graphs


@pytest.fixture(scope="module")
def engines():
    ts.init(0)
    return {False: ts.Engine(), True: ts.Engine(graphs=True)}


@pytest.mark.parametrize("seed", [11, 12])
def test_random_knn_matches_oracle(engines, seed):
    for t, meas, mode, dl, ql, sigma, dseed, qseed, k, kw, graphs in cases(seed, 60):
        say("random knn", seed, t, meas, mode, len(dl), len(ql), int(dl.min()), int(dl.max()), k, kw, graphs)
        dx, do = ts.gen_walks(dseed, dl, sigma)
        qx, qo = ts.gen_walks(qseed, ql, sigma)
        p = ts.params(meas, mode=mode, k=k, **kw)
        prm = oracle.params(CODES[meas], mode, k=k, kc=p.kc, r=p.r, band=p.band, delta=p.delta, epsilon=p.epsilon)
        db, qs = ts.TrajSet(dx, do), ts.TrajSet(qx, qo)
        want = oracle.knn(prm, dx, do, qx, qo)
        for rep in range(3 if graphs else 1):
            idx, val, _ = engines[graphs].knn(db, qs, p)
            assert np.array_equal(idx, want[0]) and np.array_equal(val.view(np.int32), want[1].view(np.int32)), \
                (t, meas, mode, len(dl), len(ql), k, kw, graphs, rep)
