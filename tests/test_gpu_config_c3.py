"""Oracle parity at the uneven-length batching workload: 4,096 query random
walks (lengths log-uniform in [16, 2048]) against a resident database of
1,000,000 random walks (lengths log-uniform in [16, 1024]), DTW and discrete
Frechet, EXACT top-10 (DB seed 0xDB5EED, query seed 0x9E3779B9).

The batched call runs over all 4,096 queries at once, as the benchmark does;
its rows for a sample of queries are then compared bit for bit with the
oracle: the longest queries (up to 2,048 points: the longest pairs and most
refines), the shortest (16 points) and a few from the middle. The plain
method's answers for the sample must equal them too."""
import numpy as np
import pytest

import cudaruntime1_amd as ts
import oracle
from tests.helpers import say

pytestmark = pytest.mark.gpu
DB_SEED, Q_SEED = 0xDB5EED, 0x9E3779B9
N_DB, N_Q = 1_000_000, 4096


@pytest.fixture(scope="module")
def work():
    ts.init(0)
    db_len = ts.gen_lengths(DB_SEED, N_DB, 16, 1024, True)
    db_xy, db_off = ts.gen_walks(DB_SEED, db_len)
    q_len = ts.gen_lengths(Q_SEED, N_Q, 16, 2048, True)
    q_xy, q_off = ts.gen_walks(Q_SEED, q_len)
    order = np.argsort(q_len, kind="stable")
    w = dict(db=ts.TrajSet(db_xy, db_off), qs=ts.TrajSet(q_xy, q_off), db_xy=db_xy, db_off=db_off, q_xy=q_xy,
             q_off=q_off, q_len=q_len, order=order, eng=ts.Engine(), full={})
    yield w
    for key in ("db", "qs", "eng"):
        w[key].close()


def sample(w, group):
    o = w["order"]
    if group == "longest":
        return o[-4:]
    if group == "shortest":
        return o[:4]
    return o[[len(o) // 4, len(o) // 2, 3 * len(o) // 4]]


def sub_queries(w, sel):
    parts = [w["q_xy"][w["q_off"][i]:w["q_off"][i + 1]] for i in sel]
    return np.concatenate(parts), ts.offsets_of([len(p) for p in parts])


def batched(w, measure):
    """The whole 4,096-query batch, once per measure."""
    if measure not in w["full"]:
        say("c3 batched", measure, N_Q, "x", N_DB)
        p = ts.params(measure, mode=ts.EXACT, k=10)
        w["full"][measure] = w["eng"].knn(w["db"], w["qs"], p)[:2]
    return w["full"][measure]


@pytest.mark.parametrize("group", ["longest", "shortest", "middle"])
@pytest.mark.parametrize("measure", ["dtw", "frechet"])
def test_c3_batched_rows_match_oracle(work, measure, group):
    gi, gv = batched(work, measure)
    sel = sample(work, group)
    say("c3 oracle", measure, group, "lengths", work["q_len"][sel].tolist())
    s_xy, s_off = sub_queries(work, sel)
    p = ts.params(measure, mode=ts.EXACT, k=10)
    wi, wv, _ = oracle.knn(oracle.params(p.measure, oracle.EXACT, k=10), work["db_xy"], work["db_off"], s_xy, s_off)
    assert np.array_equal(gi[sel], wi), (measure, group, gi[sel], wi)
    assert np.array_equal(gv[sel].view(np.int32), wv.view(np.int32)), (measure, group, gv[sel], wv)
    # the plain method on the same queries
    sq = ts.TrajSet(s_xy, s_off)
    try:
        pi, pv, _ = work["eng"].knn_plain(work["db"], sq, p)
    finally:
        sq.close()
    assert np.array_equal(pi, wi) and np.array_equal(pv.view(np.int32), wv.view(np.int32)), (measure, group, "plain")
