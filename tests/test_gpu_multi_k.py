"""ts_knn_multi_k: one batch in which every query asks for its own k. Row i's
first k_i entries must equal the oracle's top-k_i for that query (EXACT:
the true top-k_i over the whole database; REF: the best k_i of the query's
kc filter candidates), bit for bit, and its later slots hold index -1; for
every measure, with ties (planted near-duplicates) and k up to 100 > some
queries' candidate counts."""
import numpy as np
import pytest

import cudaruntime1_amd as ts
import oracle
from tests.helpers import planted, say, walks

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world():
    ts.init(0)
    L = ts.gen_lengths(91, 1500, 16, 240, True)
    db_xy, db_off = walks(91, L)
    qL = ts.gen_lengths(92, 36, 16, 300, True)
    q_xy, q_off = planted(db_xy, db_off, np.arange(7, 1500, 41)[:36], qL, 5e-4, 93)
    e_xy, e_off = walks(94, np.full(800, 48, np.int32))
    eq_xy, eq_off = walks(95, np.full(20, 48, np.int32))
    w = dict(db=ts.TrajSet(db_xy, db_off), qs=ts.TrajSet(q_xy, q_off), db_xy=db_xy, db_off=db_off, q_xy=q_xy,
             q_off=q_off, e_db=ts.TrajSet(e_xy, e_off), e_qs=ts.TrajSet(eq_xy, eq_off), e_db_xy=e_xy, e_db_off=e_off,
             e_q_xy=eq_xy, e_q_off=eq_off, eng=ts.Engine())
    yield w
    for k in ("db", "qs", "e_db", "e_qs", "eng"):
        w[k].close()


@pytest.mark.parametrize("mode", [ts.EXACT, ts.REF])
@pytest.mark.parametrize("measure,kw", [("dtw", {}), ("dtw", dict(band=-10)), ("frechet", {}), ("hausdorff", {}),
                                        ("lcss", dict(epsilon=2e-3)), ("lcss", dict(epsilon=2e-3, delta=16)),
                                        ("euclidean", {})])
def test_multi_k_rows_match_oracle_at_their_own_k(world, measure, kw, mode):
    if mode == ts.REF and ("band" in kw or "delta" in kw):
        pytest.skip("bands and windows are EXACT-mode extensions")
    eu = measure == "euclidean"
    db, qs = (world["e_db"], world["e_qs"]) if eu else (world["db"], world["qs"])
    dxy, doff, qxy, qoff = ((world["e_db_xy"], world["e_db_off"], world["e_q_xy"], world["e_q_off"]) if eu else
                            (world["db_xy"], world["db_off"], world["q_xy"], world["q_off"]))
    rng = np.random.default_rng(len(measure) * 7 + mode)
    kq = rng.choice([1, 10, 100], size=qs.count).astype(np.int32)
    kq[0] = 100
    p = ts.params(measure, mode=mode, k=100, kc=100 if mode == ts.REF else None, **kw)
    say("multi_k", measure, kw, "mode", mode)
    gi, gv, _ = world["eng"].knn_multi_k(db, qs, p, kq)
    for k in (1, 10, 100):
        rows = np.nonzero(kq == k)[0]
        if len(rows) == 0:
            continue
        s_xy = np.concatenate([qxy[qoff[i]:qoff[i + 1]] for i in rows])
        s_off = ts.offsets_of([qoff[i + 1] - qoff[i] for i in rows])
        prm = oracle.params(p.measure, mode, k=k, kc=p.kc, band=p.band, delta=p.delta, epsilon=p.epsilon)
        wi, wv, _ = oracle.knn(prm, dxy, doff, s_xy, s_off)
        assert np.array_equal(gi[rows, :k], wi), (measure, kw, mode, k)
        assert np.array_equal(gv[rows, :k].view(np.int32), wv.view(np.int32)), (measure, kw, mode, k)
        assert (gi[rows, k:] == -1).all()
