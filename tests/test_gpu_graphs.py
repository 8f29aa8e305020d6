"""hipGraph replay of batched k-NN: every replay equals the eager engine's
result, with eager work and synchronisations between replays, for the REF,
EXACT and all-pairs pipelines; in-place reloads are seen by replays; an
EXACT replay whose candidates outgrow the recorded capacity grows it and
repeats the call with identical results."""
import numpy as np
import pytest

import cudaruntime1_amd as ts
from helpers import planted, say, uniform_lengths, walks

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def work():
    ts.init(0)
    L = uniform_lengths(51, 3000, 32, 160)
    db_xy, db_off = walks(51, L, 2e-3)
    q_xy, q_off = planted(db_xy, db_off, np.arange(0, 3000, 150), L[0:3000:150], 1e-3, 52)
    return ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off), db_xy, db_off, q_xy, q_off


def same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.int32), b[1].view(np.int32))


@pytest.mark.parametrize("measure,mode", [("dtw", ts.EXACT), ("lcss", ts.EXACT), ("frechet", ts.REF),
                                          ("hausdorff", ts.EXACT)])
def test_replays_equal_eager(work, measure, mode):
    say("graphs", measure, mode)
    db, qs = work[0], work[1]
    p = ts.params(measure, mode=mode, k=10, epsilon=2e-3)
    eager = ts.Engine()
    g = ts.Engine(graphs=True)
    want = eager.knn(db, qs, p)
    for i in range(5):
        got = g.knn(db, qs, p)
        assert same(got, want), (measure, mode, i)
        # eager work on another engine between replays
        eager.knn(db, qs, ts.params("dtw", mode=ts.EXACT, k=3))
    info = g.graph_info()
    assert info["captures"] >= 1 and info["replays"] >= 3, info


def test_all_pairs_replays(work):
    say("graphs all-pairs")
    db_xy, db_off = work[2], work[3]
    small = ts.TrajSet(db_xy[:db_off[500]], db_off[:501])
    q = ts.TrajSet(work[4][:work[5][2]], work[5][:3])
    p = ts.params("dtw", mode=ts.EXACT, k=10)
    want = ts.Engine().knn(small, q, p)
    g = ts.Engine(graphs=True)
    for _ in range(4):
        assert same(g.knn(small, q, p), want)
    assert g.graph_info()["replays"] >= 2


def test_reload_and_capacity_growth():
    """Replays read the reloaded query set; a reload that raises the number
    of candidates past the recorded capacity triggers one retry. DB: 3000
    walks of 96 points whose first 600 are copies of walk 0. Queries A are
    exact copies of 40 unique DB walks (k=1: one candidate each, the copy
    itself, bound 0); queries B are 40 copies of walk 0 (600 candidates
    each, every duplicate has bound 0)."""
    say("graphs reload")
    lens = np.full(3000, 96, np.int32)
    db_xy, db_off = walks(55, lens, 2e-3)
    t0 = db_xy[0:96].copy()
    for j in range(1, 600):
        db_xy[db_off[j]:db_off[j + 1]] = t0
    db = ts.TrajSet(db_xy, db_off)
    q_off = ts.offsets_of([96] * 40)
    a_xy = np.concatenate([db_xy[db_off[j]:db_off[j + 1]] for j in range(1000, 1040)])
    b_xy = np.concatenate([t0] * 40)
    qs = ts.TrajSet(a_xy, q_off)
    p = ts.params("dtw", mode=ts.EXACT, k=1)
    eager = ts.Engine()
    g = ts.Engine(graphs=True)
    want_a = eager.knn(db, qs, p)
    assert np.array_equal(want_a[0][:, 0], np.arange(1000, 1040))
    for _ in range(3):
        assert same(g.knn(db, qs, p), want_a)
    assert g.graph_info()["replays"] >= 1
    qs.reload(b_xy, q_off)
    want_b = eager.knn(db, ts.TrajSet(b_xy, q_off), p)
    assert np.array_equal(want_b[0][:, 0], np.zeros(40))
    r0 = g.knn(db, qs, p)
    assert same(r0, want_b)
    assert r0[2]["overflow_retries"] >= 1, r0[2]
    for _ in range(3):
        assert same(g.knn(db, qs, p), want_b)
