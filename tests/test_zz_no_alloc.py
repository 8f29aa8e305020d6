"""ts_pair_values_noalloc (the allocation-free variant of ts_pair_values,
all device buffers in the library's static scratch array) against
ts_pair_values, bit for bit, through the C ABI: two calls in a row, the
second on fewer pairs than the first, for every measure and DP option, over
pairs that fill every DP length bucket; sizes above the maximum are refused.
Collected last (the name sorts after every other test module), so it runs
both after the whole suite and alone in a fresh process."""
import numpy as np
import pytest

import cudaruntime1_amd as ts
from tests.helpers import say, walks

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world():
    ts.init(0)
    db_len = ts.gen_lengths(71, 300, 1, 1500, True)
    q_len = ts.gen_lengths(72, 60, 1, 2100, True)
    db_len[:3] = [1, 1500, 1024]
    q_len[:3] = [2100, 1, 256]
    db_xy, db_off = walks(71, db_len, 5e-3)
    q_xy, q_off = walks(72, q_len, 5e-3)
    e_db_xy, e_db_off = walks(73, np.full(200, 64, np.int32), 5e-3)
    e_q_xy, e_q_off = walks(74, np.full(40, 64, np.int32), 5e-3)
    w = dict(db=ts.TrajSet(db_xy, db_off), qs=ts.TrajSet(q_xy, q_off), e_db=ts.TrajSet(e_db_xy, e_db_off),
             e_qs=ts.TrajSet(e_q_xy, e_q_off), eng=ts.Engine())
    yield w
    for k in ("db", "qs", "e_db", "e_qs", "eng"):
        w[k].close()


def pairs(seed, n, nq, nd):
    rng = np.random.default_rng(seed)
    return rng.integers(0, nq, n).astype(np.int32), rng.integers(0, nd, n).astype(np.int32)


CASES = [("dtw", {}), ("dtw", dict(band=-10)), ("dtw", dict(band=7)), ("frechet", {}),
         ("lcss", dict(epsilon=1e-2)), ("lcss", dict(epsilon=1e-2, delta=20)), ("hausdorff", {}), ("euclidean", {})]


@pytest.mark.parametrize("measure,kw", CASES)
def test_noalloc_equals_pair_values_two_calls(world, measure, kw):
    eng = world["eng"]
    db, qs = (world["e_db"], world["e_qs"]) if measure == "euclidean" else (world["db"], world["qs"])
    p = ts.params(measure, **kw)
    for call, n in enumerate((ts.NOALLOC_MAX_PAIRS, 777)):
        qi, di = pairs(1000 * call + sum(map(ord, measure + str(kw))), n, qs.count, db.count)
        say("noalloc", measure, kw, "call", call, "pairs", n)
        want = eng.pair_values(db, qs, p, qi, di)
        got = eng.pair_values_noalloc(db, qs, p, qi, di)
        bad = np.nonzero(got.view(np.int32) != want.view(np.int32))[0]
        assert len(bad) == 0, (measure, kw, call, [(int(qi[b]), int(di[b]), got[b], want[b]) for b in bad[:5]])


def test_noalloc_refuses_more_than_its_maximum(world):
    n = ts.NOALLOC_MAX_PAIRS + 1
    qi, di = np.zeros(n, np.int32), np.zeros(n, np.int32)
    with pytest.raises(ts.TrajsimError) as e:
        world["eng"].pair_values_noalloc(world["db"], world["qs"], ts.params("dtw"), qi, di)
    assert e.value.status == 1 and "TS_NOALLOC_MAX_PAIRS" in str(e.value)
    # a refused call leaves the next one intact
    got = world["eng"].pair_values_noalloc(world["db"], world["qs"], ts.params("dtw"), qi[:5], di[:5])
    want = world["eng"].pair_values(world["db"], world["qs"], ts.params("dtw"), qi[:5], di[:5])
    assert np.array_equal(got.view(np.int32), want.view(np.int32))
