"""Serving: continuous batching over a resident DB. Each request's result is
the same alone (a single-query ts_knn) and in a mixed stream of measures,
lengths and k; admission refuses what cannot fit; cancellation works."""
import sys
import threading

import numpy as np
import pytest

import cudaruntime1_amd as ts
from tests.helpers import planted, uniform_lengths, walks

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def world():
    ts.init(0)
    L = ts.gen_lengths(41, 4000, 16, 200, True)
    db_xy, db_off = walks(41, L)
    db = ts.TrajSet(db_xy, db_off)
    qL = ts.gen_lengths(42, 120, 16, 300, True)
    picks = np.arange(5, 4000, 33)[:120]
    q_xy, q_off = planted(db_xy, db_off, picks, qL, 5e-4, 42)
    mix = [("dtw", ts.EXACT, 10), ("lcss", ts.EXACT, 1), ("frechet", ts.EXACT, 100), ("hausdorff", ts.EXACT, 10),
           ("dtw", ts.REF, 10), ("lcss", ts.REF, 10)]
    reqs = []
    for i in range(120):
        m, mode, k = mix[i % len(mix)]
        reqs.append((q_xy[q_off[i]:q_off[i + 1]], ts.params(m, mode=mode, k=k, epsilon=2e-3)))
    return db, reqs


def alone(eng, db, q, p):
    s = ts.TrajSet(q, np.array([0, len(q)], np.int64))
    i, v, _ = eng.knn(db, s, p)
    return i[0], v[0]


def test_results_same_alone_and_in_mix(world):
    db, reqs = world
    eng = ts.Engine()
    want = [alone(eng, db, q, p) for q, p in reqs]
    say("server mix", len(reqs))
    srv = ts.Server(db, streams=4, max_batch=64)
    got = [None] * len(reqs)

    def client(lo, hi):
        tickets = [(i, srv.submit(reqs[i][0], reqs[i][1])) for i in range(lo, hi)]
        for i, t in tickets:
            got[i] = srv.wait(t, reqs[i][1].k, timeout_ms=60000)

    threads = [threading.Thread(target=client, args=(j * 30, (j + 1) * 30)) for j in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for i, ((wi, wv), g) in enumerate(zip(want, got)):
        assert g is not None, i
        assert np.array_equal(g[0], wi) and np.array_equal(g[1].view(np.int32), wv.view(np.int32)), i
        assert g[2] > 0
    c = srv.counters()
    say("counters", c)
    assert c["completed"] == len(reqs) and c["refused"] == 0 and c["batches"] < len(reqs)
    srv.close()


def test_admission_refuses_what_cannot_fit(world):
    db, reqs = world
    say("admission")
    srv = ts.Server(db, streams=1, fake_free_mem=50_000, mem_reserve=0)
    q, p = reqs[0]
    with pytest.raises(ts.TrajsimError) as e:
        srv.submit(q, p)  # EXACT workspace for a 4000-trajectory DB exceeds 50 kB
    assert e.value.status == 6
    assert srv.counters()["refused"] == 1
    srv.close()
    srv = ts.Server(db, streams=1, max_queue=2, batch_wait_us=100000)
    accepted, refused = 0, 0
    for i in range(40):
        try:
            srv.submit(reqs[i][0], reqs[i][1])
            accepted += 1
        except ts.TrajsimError as err:
            assert err.status == 6
            refused += 1
    assert refused > 0 and accepted >= 2
    srv.close()


def test_cancel_queued_request(world):
    db, reqs = world
    say("cancel")
    srv = ts.Server(db, streams=1, max_batch=1)
    tickets = [srv.submit(q, p) for q, p in reqs[:20]]
    srv.cancel(tickets[-1])
    with pytest.raises(ts.TrajsimError) as e:
        srv.wait(tickets[-1], reqs[19][1].k, timeout_ms=60000)
    assert e.value.status in (8,)
    for t, (q, p) in zip(tickets[:-1], reqs[:19]):
        srv.wait(t, p.k, timeout_ms=60000)
    assert srv.counters()["cancelled"] >= 1
    srv.close()
