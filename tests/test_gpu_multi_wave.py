"""The multi-wave DP shapes (one pair over the 2 or 4 waves of a block, 4 or
8 cells per lane step, the rows between strips handed over in LDS): forced
on every bucket, on pairs up to 8192 x 8192, unbanded, with DTW bands and
LCSS windows, and with coordinates whose squared distances fall below 2^-96;
values equal the oracle's bit for bit, every launch runs the forced shape
(none falls back to another). Then an EXACT top-10 of long queries with the
multi-wave shapes on every bucket."""
import numpy as np
import pytest

import cudaruntime1_amd as ts
import oracle
from tests.helpers import say, walks

pytestmark = pytest.mark.gpu

MW = [(W, U) for (W, U) in ts.dp_shapes() if W > 64]
CASES = [(0, "dtw", {}), (1, "dtw", {"band": -10}), (1, "dtw", {"band": 300}), (2, "frechet", {}),
         (3, "lcss", {"epsilon": 2e-3}), (4, "lcss", {"epsilon": 2e-3, "delta": 64}),
         (4, "lcss", {"epsilon": 2e-3, "delta": 0})]


@pytest.fixture(scope="module")
def long_pairs():
    ts.init(0)
    db_xy, db_off = walks(61, [8192, 7000, 4100, 1500, 300, 65, 64, 1])
    q_xy, q_off = walks(62, [8192, 5000, 2049, 100, 17])
    nq, nd = len(q_off) - 1, len(db_off) - 1
    qi, di = np.meshgrid(np.arange(nq), np.arange(nd), indexing="ij")
    return dict(db=ts.TrajSet(db_xy, db_off), qs=ts.TrajSet(q_xy, q_off), db_xy=db_xy, db_off=db_off, q_xy=q_xy,
                q_off=q_off, qi=qi.ravel(), di=di.ravel())


def mw_engine(variant, W, U, buckets=None, waves=4):
    e = ts.Engine(timing=True)
    for b in (range(len(ts.DP_BUCKET_CAPS)) if buckets is None else buckets):
        e.set_dp_shape(variant, b, W, U, waves)
    return e


def ran_shape(e, W, U):
    return any(f"/w{W}u{U}" in name for name in e.kernel_stats())


def test_multi_wave_shapes_are_built():
    assert set(MW) == {(128, 4), (128, 8), (256, 4), (256, 8)}


@pytest.mark.parametrize("variant,measure,kw", CASES)
def test_multi_wave_long_pairs_bit_exact(long_pairs, variant, measure, kw):
    s = long_pairs
    p = ts.params(measure, **kw)
    prm = oracle.params(p.measure, band=p.band, delta=p.delta, epsilon=p.epsilon)
    want = oracle.pair_values(prm, s["db_xy"], s["db_off"], s["q_xy"], s["q_off"], s["qi"], s["di"])
    for (W, U) in MW:
        say("multi-wave pairs", measure, kw, W, U)
        e = mw_engine(variant, W, U)
        got = e.pair_values(s["db"], s["qs"], p, s["qi"], s["di"])
        assert e.dp_fallbacks() == 0 and ran_shape(e, W, U), (W, U, e.kernel_stats())
        bad = np.nonzero(got.view(np.int32) != want.view(np.int32))[0]
        assert len(bad) == 0, (W, U, [(int(s["qs"].lengths[s["qi"][b]]), int(s["db"].lengths[s["di"][b]]),
                                       got[b], want[b]) for b in bad[:8]])
        e.close()


def lattice_below(seed, lengths):
    """Coordinates +-(2^-25 - m 2^-49), m in 1..5 (all below 2^-25), against
    +-(2^-25 + m 2^-48) or 0: squared distances down to 2^-98."""
    rng = np.random.default_rng(seed)
    off = ts.offsets_of(lengths)
    n = int(off[-1])
    sign = rng.choice([-1.0, 1.0], size=(n, 2))
    a = sign * (2.0 ** -25 + rng.integers(0, 6, size=(n, 2)) * 2.0 ** -48)
    a[rng.random((n, 2)) < 0.2] = 0.0
    b = sign * (2.0 ** -25 - rng.integers(1, 6, size=(n, 2)) * 2.0 ** -49)
    return a.astype(np.float32), b.astype(np.float32), off


@pytest.mark.parametrize("band", [None, -10])
def test_multi_wave_tiny_coordinates_bit_exact(band):
    ts.init(0)
    a_xy, _, a_off = lattice_below(63, [6000, 700])
    _, b_xy, b_off = lattice_below(64, [6500, 800])
    qs, db = ts.TrajSet(b_xy, b_off), ts.TrajSet(a_xy, a_off)
    assert qs.flags() == ts.SET_TINY_COORD
    qi, di = np.meshgrid(np.arange(2), np.arange(2), indexing="ij")
    qi, di = qi.ravel(), di.ravel()
    p = ts.params("dtw", band=band)
    want = oracle.pair_values(oracle.params(p.measure, band=p.band), a_xy, a_off, b_xy, b_off, qi, di)
    for (W, U) in MW:
        say("multi-wave tiny", band, W, U)
        e = mw_engine(1 if band else 0, W, U)
        got = e.pair_values(db, qs, p, qi, di)
        assert e.dp_fallbacks() == 0 and ran_shape(e, W, U)
        assert np.array_equal(got.view(np.int32), want.view(np.int32)), (W, U, got, want)
        e.close()


@pytest.mark.parametrize("variant,measure,kw", [c for c in CASES if c[2].get("delta") != 0])
def test_multi_wave_exact_knn_matches_oracle(variant, measure, kw):
    """EXACT top-10 of 4 queries (8192, 3000, 500 and 40 points) over 240 DB
    trajectories of 16 to 1500 points and 3 of 8000, every bucket run by a
    multi-wave shape (the seeding pass and the later refine alike; LCSS with
    delta 64 visits 129 columns a row, the fourth bucket)."""
    ts.init(0)
    L = np.concatenate([ts.gen_lengths(65, 240, 16, 1500, True), [8000, 8000, 8000]]).astype(np.int32)
    db_xy, db_off = walks(66, L)
    q_xy, q_off = walks(67, [8192, 3000, 500, 40])
    db, qs = ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off)
    p = ts.params(measure, mode=ts.EXACT, k=10, **kw)
    want = oracle.knn(oracle.params(p.measure, p.mode, k=10, band=p.band, delta=p.delta, epsilon=p.epsilon),
                      db_xy, db_off, q_xy, q_off)
    for (W, U) in [(128, 8), (256, 4)]:
        say("multi-wave knn", measure, kw, W, U)
        e = mw_engine(variant, W, U)
        gi, gv, _ = e.knn(db, qs, p)
        assert e.dp_fallbacks() == 0 and ran_shape(e, W, U), (W, U, e.kernel_stats())
        assert np.array_equal(gi, want[0]) and np.array_equal(gv.view(np.int32), want[1].view(np.int32)), (W, U)
        e.close()
