"""Oracle parity of served results at the serving workload: a database of
2,000,000 random walks (lengths log-uniform in [16, 512], DB seed 0xDB5EED)
resident in HBM; requests with query lengths log-uniform in [16, 1024]
(query seed 0x9E3779B9), measure mix DTW 40% / LCSS 20% / Frechet 20% /
Hausdorff 20% and k uniform in {1, 10, 100} (arrival seed 0xA11CE), EXACT
top-k, LCSS eps 2e-3.

128 requests arrive as a Poisson stream at 2,000 per second into the
continuous-batching server; every served top-k must equal the oracle's
bit for bit (indices and values, ties by index)."""
import threading
import time

import numpy as np
import pytest

import cudaruntime1_amd as ts
import oracle
from tests.helpers import say

pytestmark = pytest.mark.gpu
DB_SEED, Q_SEED, ARR_SEED = 0xDB5EED, 0x9E3779B9, 0xA11CE
N_DB, N_REQ, QPS = 2_000_000, 128, 2000.0
NAMES = ["dtw", "lcss", "frechet", "hausdorff"]
ORACLE_MEASURE = {0: oracle.DTW, 1: oracle.LCSS, 2: oracle.FRECHET, 3: oracle.HAUSDORFF}


def test_c5_served_results_match_oracle():
    ts.init(0)
    db_len = ts.gen_lengths(DB_SEED, N_DB, 16, 512, True)
    db_xy, db_off = ts.gen_walks(DB_SEED, db_len)
    q_len = ts.gen_lengths(Q_SEED, N_REQ, 16, 1024, True)
    q_xy, q_off = ts.gen_walks(Q_SEED, q_len)
    rng = np.random.default_rng(ARR_SEED)
    mix = rng.choice(4, size=N_REQ, p=[0.4, 0.2, 0.2, 0.2])
    ks = rng.choice([1, 10, 100], size=N_REQ)
    arrivals = np.cumsum(rng.exponential(1.0 / QPS, size=N_REQ))
    prm = {(m, k): ts.params(NAMES[m], mode=ts.EXACT, k=int(k), epsilon=2e-3) for m in range(4) for k in (1, 10, 100)}
    db = ts.TrajSet(db_xy, db_off)
    say("c5 serve", N_REQ, "requests at", QPS, "QPS over", N_DB)
    srv = ts.Server(db, streams=4, max_batch=256, batch_wait_us=500)
    tickets = [None] * N_REQ
    got = [None] * N_REQ

    def waiter():
        for i in range(N_REQ):
            while tickets[i] is None:
                time.sleep(0.0002)
            got[i] = srv.wait(tickets[i], int(ks[i]), timeout_ms=120000)

    th = threading.Thread(target=waiter)
    th.start()
    start = time.perf_counter()
    try:
        for i in range(N_REQ):
            while time.perf_counter() - start < arrivals[i]:
                pass
            tickets[i] = srv.submit(q_xy[q_off[i]:q_off[i + 1]], prm[(int(mix[i]), int(ks[i]))])
    finally:
        th.join()
        c = srv.counters()
        srv.close()
        db.close()
    say("c5 counters", c)
    assert c["completed"] == N_REQ and c["refused"] == 0
    bad = []
    for (m, k), p in prm.items():
        sel = [i for i in range(N_REQ) if mix[i] == m and ks[i] == k]
        if not sel:
            continue
        sub = [q_xy[q_off[i]:q_off[i + 1]] for i in sel]
        say("c5 oracle", NAMES[m], "k", k, len(sel), "requests")
        op = oracle.params(ORACLE_MEASURE[m], oracle.EXACT, k=int(k), epsilon=2e-3)
        wi, wv, _ = oracle.knn(op, db_xy, db_off, np.concatenate(sub), ts.offsets_of([len(t) for t in sub]))
        for r, i in enumerate(sel):
            gi, gv, lat = got[i]
            if not (np.array_equal(gi, wi[r]) and np.array_equal(gv.view(np.int32), wv[r].view(np.int32))):
                bad.append((i, NAMES[m], int(k), int(q_len[i])))
            assert lat > 0
    assert not bad, bad[:5]
