"""The DB's length-bucketed tile view, which the REF filter of DTW, Frechet
and LCSS reads: scores bit for bit against the oracle's per-pair left fold
(dtw.cpp:45-58, frechet.cpp:46-62, lcss.cpp:46-63) for lengths spread over
three decades (tiles of very unequal lengths, one-point trajectories, a
partial last tile, a partial group of queries), after the set is reloaded
and after it is appended to (the tiles are rebuilt from the new contents),
with coordinates small enough to take the square root's slow path, and
through replayed HIP graphs with the database reloaded in place between
replays."""
import numpy as np
import pytest

import cudaruntime1_amd as ts
import oracle
from tests.helpers import say, walks

pytestmark = pytest.mark.gpu
MEASURES = ["dtw", "frechet", "lcss"]


@pytest.fixture(scope="module")
def eng():
    ts.init(0)
    e = ts.Engine()
    yield e
    e.close()


def check(eng, db, qs, db_xy, db_off, q_xy, q_off, what):
    for measure in MEASURES:
        p = ts.params(measure, epsilon=3e-3)
        got = eng.filter_ref(db, qs, p)
        want = oracle.filter_ref(oracle.params(p.measure, epsilon=p.epsilon), db_xy, db_off, q_xy, q_off)
        bad = np.argwhere(got.view(np.int32) != want.view(np.int32))
        assert len(bad) == 0, (what, measure, [(int(a), int(b), got[a, b], want[a, b]) for a, b in bad[:5]])


def test_tiles_uneven_lengths_reload_append(eng):
    L = ts.gen_lengths(81, 333, 1, 3000, True)
    L[:4] = [1, 3000, 2, 1]
    db_xy, db_off = walks(81, L, 2e-3)
    qL = ts.gen_lengths(82, 13, 1, 2000, True)
    qL[0] = 1
    q_xy, q_off = walks(82, qL, 2e-3)
    db, qs = ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off)
    try:
        say("tiles first build", db.count)
        check(eng, db, qs, db_xy, db_off, q_xy, q_off, "first")
        # reload with other contents and another shape
        L2 = ts.gen_lengths(83, 200, 40, 900, True)
        db2_xy, db2_off = walks(83, L2, 2e-3)
        db.reload(db2_xy, db2_off)
        say("tiles after reload", db.count)
        check(eng, db, qs, db2_xy, db2_off, q_xy, q_off, "reload")
        # append: the new trajectories get the next indices
        L3 = ts.gen_lengths(84, 77, 1, 1500, True)
        db3_xy, db3_off = walks(84, L3, 2e-3)
        db.append(db3_xy, db3_off)
        all_xy = np.concatenate([db2_xy, db3_xy])
        all_off = np.concatenate([db2_off, db2_off[-1] + db3_off[1:]])
        say("tiles after append", db.count)
        check(eng, db, qs, all_xy, all_off, q_xy, q_off, "append")
    finally:
        db.close()
        qs.close()


def test_tiles_tiny_coordinates(eng):
    """Coordinates near 1e-12: squared box distances below 2^-96, the
    square root's slow path."""
    L = ts.gen_lengths(85, 150, 1, 400, True)
    db_xy, db_off = walks(85, L, 2e-2)
    q_xy, q_off = walks(86, ts.gen_lengths(86, 9, 1, 300, True), 2e-2)
    db_xy = (db_xy.astype(np.float64) * 1e-12).astype(np.float32)
    q_xy = (q_xy.astype(np.float64) * 1e-12).astype(np.float32)
    db, qs = ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off)
    try:
        say("tiles tiny coordinates")
        for measure in ("dtw", "frechet"):
            p = ts.params(measure)
            got = eng.filter_ref(db, qs, p)
            want = oracle.filter_ref(oracle.params(p.measure), db_xy, db_off, q_xy, q_off)
            assert np.array_equal(got.view(np.int32), want.view(np.int32)), measure
    finally:
        db.close()
        qs.close()


def test_ref_knn_graph_replays_see_reloaded_tiles():
    """REF k-NN replayed from a graph: the database is reloaded in place with
    new walks of the same shape between calls, so the replay reads tiles
    rebuilt in the same buffers; every call equals the oracle."""
    ts.init(0)
    e = ts.Engine(graphs=True)
    L = ts.gen_lengths(87, 600, 16, 300, True)
    q_xy, q_off = walks(88, ts.gen_lengths(88, 12, 16, 300, True), 2e-3)
    qs = ts.TrajSet(q_xy, q_off)
    db = None
    try:
        for call in range(5):
            db_xy, db_off = walks(900 + call, L, 2e-3)
            if db is None:
                db = ts.TrajSet(db_xy, db_off)
            else:
                db.reload(db_xy, db_off)
            for measure in MEASURES:
                p = ts.params(measure, mode=ts.REF, k=5, epsilon=3e-3)
                say("ref graph call", call, measure)
                gi, gv, _ = e.knn(db, qs, p)
                prm = oracle.params(p.measure, oracle.REF, k=5, epsilon=p.epsilon)
                wi, wv, _ = oracle.knn(prm, db_xy, db_off, q_xy, q_off)
                assert np.array_equal(gi, wi) and np.array_equal(gv.view(np.int32), wv.view(np.int32)), (call, measure)
        g = e.graph_info()
        assert g["captures"] >= 1 and g["replays"] >= 1, g
    finally:
        if db is not None:
            db.close()
        qs.close()
        e.close()
