"""Oracle parity at the long-trajectory workload: random walks of lengths
uniform in [2048, 8192], DTW with a Sakoe-Chiba band of 10% of the longer
length (w = ceil(0.1 max(n, m)) per pair) and LCSS with eps 2e-3 and an
index window of 64, EXACT top-10 (DB seed 0xDB5EED, query seed 0x9E3779B9).

The database is the workload's first 1,000 trajectories (the full 20,000
would keep the oracle busy for minutes) and the queries its first 8, with
one trajectory of each set set to the longest length, 8,192, and one to the
shortest, 2,048. Pair values cover the 8,192 x 8,192 pair under every
measure's options, banded and unbanded; the k-NN is checked batched and
plain."""
import numpy as np
import pytest

import cudaruntime1_amd as ts
import oracle
from tests.helpers import say

pytestmark = pytest.mark.gpu
DB_SEED, Q_SEED = 0xDB5EED, 0x9E3779B9
N_DB, N_Q = 1000, 8


@pytest.fixture(scope="module")
def work():
    ts.init(0)
    db_len = ts.gen_lengths(DB_SEED, N_DB, 2048, 8192)
    db_len[0], db_len[1] = 8192, 2048
    db_xy, db_off = ts.gen_walks(DB_SEED, db_len)
    q_len = ts.gen_lengths(Q_SEED, N_Q, 2048, 8192)
    q_len[0], q_len[1] = 8192, 2048
    q_xy, q_off = ts.gen_walks(Q_SEED, q_len)
    w = dict(db=ts.TrajSet(db_xy, db_off), qs=ts.TrajSet(q_xy, q_off), db_xy=db_xy, db_off=db_off, q_xy=q_xy,
             q_off=q_off, eng=ts.Engine())
    yield w
    for key in ("db", "qs", "eng"):
        w[key].close()


OPTIONS = [("dtw", dict(band=-10)), ("dtw", dict(band=0)), ("lcss", dict(epsilon=2e-3, delta=64)),
           ("lcss", dict(epsilon=2e-3)), ("frechet", {})]


@pytest.mark.parametrize("measure,kw", OPTIONS)
def test_c4_longest_pairs_bit_exact(work, measure, kw):
    """(query, DB) pairs 8192 x 8192, 8192 x 2048, 2048 x 8192, 2048 x 2048
    and two of mixed length, exact values against the oracle."""
    qi = np.array([0, 0, 1, 1, 2, 5], np.int32)
    di = np.array([0, 1, 0, 1, 7, 3], np.int32)
    say("c4 pairs", measure, kw)
    p = ts.params(measure, **kw)
    got = work["eng"].pair_values(work["db"], work["qs"], p, qi, di)
    prm = oracle.params(p.measure, band=p.band, delta=p.delta, epsilon=p.epsilon)
    want = oracle.pair_values(prm, work["db_xy"], work["db_off"], work["q_xy"], work["q_off"], qi, di)
    assert np.array_equal(got.view(np.int32), want.view(np.int32)), (got, want)


@pytest.mark.parametrize("measure,kw", [OPTIONS[0], OPTIONS[2]])
def test_c4_exact_top10_matches_oracle(work, measure, kw):
    p = ts.params(measure, mode=ts.EXACT, k=10, **kw)
    say("c4 knn", measure, kw)
    prm = oracle.params(p.measure, oracle.EXACT, k=10, band=p.band, delta=p.delta, epsilon=p.epsilon)
    wi, wv, _ = oracle.knn(prm, work["db_xy"], work["db_off"], work["q_xy"], work["q_off"])
    for fn in (work["eng"].knn, work["eng"].knn_plain):
        gi, gv, _ = fn(work["db"], work["qs"], p)
        assert np.array_equal(gi, wi), (fn.__name__, gi, wi)
        assert np.array_equal(gv.view(np.int32), wv.view(np.int32)), (fn.__name__, gv, wv)
