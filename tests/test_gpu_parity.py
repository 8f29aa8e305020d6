"""GPU parity: every HIP kernel path of libtrajsim.so against the CPU oracle
on the same seeded inputs. Integer/index results and f32 values are compared
bit for bit (the oracle restates the reference's f32 arithmetic; the kernels
keep its operation order, no FMA contraction, correctly rounded sqrtf)."""
import ctypes
import os
import sys

import numpy as np
import pytest

import cudaruntime1_amd as ts
import oracle
from tests.helpers import planted, uniform_lengths, walks

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

LENGTHS = [1, 2, 3, 15, 16, 17, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 257]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    ts.init(0)
    say("device", ts.device_name())
    e = ts.Engine()
    yield e
    e.close()


@pytest.fixture(scope="module")
def shapes():
    db_xy, db_off = walks(21, LENGTHS, 2e-2)
    q_xy, q_off = walks(22, LENGTHS[::-1], 2e-2)
    db = ts.TrajSet(db_xy, db_off)
    qs = ts.TrajSet(q_xy, q_off)
    n = len(LENGTHS)
    qi, di = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    return dict(db=db, qs=qs, db_xy=db_xy, db_off=db_off, q_xy=q_xy, q_off=q_off, qi=qi.ravel(), di=di.ravel())


def check_pairs(eng, s, measure, **kw):
    p = ts.params(measure, **kw)
    prm = oracle.params(p.measure, p.mode, k=p.k, kc=p.kc, r=p.r, band=p.band, delta=p.delta, epsilon=p.epsilon)
    got = eng.pair_values(s["db"], s["qs"], p, s["qi"], s["di"])
    want = oracle.pair_values(prm, s["db_xy"], s["db_off"], s["q_xy"], s["q_off"], s["qi"], s["di"])
    bad = np.nonzero(got.view(np.int32) != want.view(np.int32))[0]
    assert len(bad) == 0, [(int(s["qs"].lengths[s["qi"][b]]), int(s["db"].lengths[s["di"][b]]), got[b], want[b])
                           for b in bad[:8]]


@pytest.mark.parametrize("measure", ["dtw", "frechet", "lcss", "hausdorff"])
def test_pair_values_bit_exact_all_shapes(eng, shapes, measure):
    say("pairs", measure, len(shapes["qi"]))
    check_pairs(eng, shapes, measure, epsilon=3e-2)


@pytest.mark.parametrize("scale", [1e-22, 1e-12, 1.0, 1e17])
def test_dtw_extreme_scales_bit_exact(eng, scale):
    """DTW point distances across the f32 range: squared distances below
    2^-96 (1e-22: every wave takes plain sqrtf), around it (1e-12: waves of
    both kinds), ordinary, and near the top (1e17), with zero distances on
    the self pairs (the set is its own database)."""
    xy, off = walks(31, LENGTHS, 2e-2)
    xy = (xy.astype(np.float64) * scale).astype(np.float32)
    t = ts.TrajSet(xy, off)
    n = len(LENGTHS)
    qi, di = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    s = dict(db=t, qs=t, db_xy=xy, db_off=off, q_xy=xy, q_off=off, qi=qi.ravel(), di=di.ravel())
    say("dtw scale", scale)
    check_pairs(eng, s, "dtw")
    check_pairs(dtw_engine_at(16), s, "dtw")


@pytest.mark.parametrize("scale", [3e18, 1e20])
def test_dtw_huge_coordinates_bit_exact(eng, scale):
    """Coordinates near 2^62 (3e18: unflagged, squared distances finite,
    DP sums that may overflow) and above it (1e20: flagged, squared
    distances of +inf, which the DP's fast square root does not take; the
    flagged path's plain sqrtf does)."""
    xy, off = walks(32, LENGTHS, 2e-2)
    xy = ((xy.astype(np.float64) - 0.5) * scale).astype(np.float32)
    t = ts.TrajSet(xy, off)
    assert t.flags() == (ts.SET_HUGE_COORD if np.abs(xy).max() >= 2.0 ** 62 else 0)
    n = len(LENGTHS)
    qi, di = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    s = dict(db=t, qs=t, db_xy=xy, db_off=off, q_xy=xy, q_off=off, qi=qi.ravel(), di=di.ravel())
    say("dtw huge scale", scale)
    check_pairs(eng, s, "dtw")
    check_pairs(dtw_engine_at(16), s, "dtw")
    check_pairs(eng, s, "dtw", band=-10)


def dtw_engine_at(waves):
    """An engine whose unbanded DTW launches run `waves` waves per CU (each
    bucket's tuned shape): at 12 or more they take the step loop split by
    the tiny-coordinate flag."""
    e = ts.Engine()
    for b in range(len(ts.DP_BUCKET_CAPS)):
        W, U, _ = e.get_dp_shape(0, b, table=1)
        e.set_dp_shape(0, b, W, U, waves)
    return e


def lattice(seed, lengths, below=False):
    """Coordinates 0 or +-(2^-25 + m 2^-48), m < 6: no tiny coordinate
    (0 < |c| < 2^-25), and coordinates 2^-48 apart, so squared distances of
    exactly 2^-96. below: +-(2^-25 - m 2^-49), m in 1..5, all tiny, whose
    squared distances to the first kind go down to 2^-98."""
    rng = np.random.default_rng(seed)
    off = ts.offsets_of(lengths)
    n = int(off[-1])
    sign = rng.choice([-1.0, 1.0], size=(n, 2))
    if below:
        xy = sign * (2.0 ** -25 - rng.integers(1, 6, size=(n, 2)) * 2.0 ** -49)
    else:
        xy = sign * (2.0 ** -25 + rng.integers(0, 6, size=(n, 2)) * 2.0 ** -48)
        xy[rng.random((n, 2)) < 0.2] = 0.0
    return xy.astype(np.float32), off


def pair_case(q_xy, q_off, db_xy, db_off):
    nq, nd = len(q_off) - 1, len(db_off) - 1
    qi, di = np.meshgrid(np.arange(nq), np.arange(nd), indexing="ij")
    return dict(db=ts.TrajSet(db_xy, db_off), qs=ts.TrajSet(q_xy, q_off), db_xy=db_xy, db_off=db_off, q_xy=q_xy,
                q_off=q_off, qi=qi.ravel(), di=di.ravel())


def min_sq_distance(a, b):
    d = ((a[:, None, :].astype(np.float64) - b[None, :, :].astype(np.float64)) ** 2).sum(-1)
    return d[d > 0].min()


@pytest.mark.parametrize("band", [None, 3, -10])
def test_dtw_at_the_tiny_threshold_bit_exact(eng, band):
    """Sets without a tiny coordinate skip the DP's test for square roots
    of values below 2^-96: on this lattice squared distances reach 2^-96
    itself (and 0), the smallest values that path then has to round."""
    q_xy, q_off = lattice(41, LENGTHS)
    db_xy, db_off = lattice(42, LENGTHS[::-1])
    assert min_sq_distance(q_xy, db_xy) == 2.0 ** -96
    s = pair_case(q_xy, q_off, db_xy, db_off)
    assert s["qs"].flags() == 0 and s["db"].flags() == 0
    say("dtw lattice band", band)
    check_pairs(eng, s, "dtw", band=band)
    check_pairs(dtw_engine_at(16), s, "dtw", band=band)
    p = ts.params("dtw", mode=ts.EXACT, k=5, band=band)
    prm = oracle.params(p.measure, p.mode, k=5, kc=p.kc, r=p.r, band=p.band)
    got = eng.knn(s["db"], s["qs"], p)
    want = oracle.knn(prm, db_xy, db_off, q_xy, q_off)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1].view(np.int32), want[1].view(np.int32))


@pytest.mark.parametrize("tiny_side", ["query", "database"])
def test_dtw_one_set_with_tiny_coordinates_bit_exact(eng, tiny_side):
    """Square roots of values below 2^-96 between a set without tiny
    coordinates and one with them: either set's flag alone switches the
    DP to the test."""
    a_xy, a_off = lattice(43, LENGTHS)
    b_xy, b_off = lattice(44, LENGTHS[::-1], below=True)
    assert min_sq_distance(a_xy, b_xy) < 2.0 ** -96
    if tiny_side == "query":
        s = pair_case(b_xy, b_off, a_xy, a_off)
        assert s["qs"].flags() == ts.SET_TINY_COORD and s["db"].flags() == 0
    else:
        s = pair_case(a_xy, a_off, b_xy, b_off)
        assert s["qs"].flags() == 0 and s["db"].flags() == ts.SET_TINY_COORD
    say("dtw tiny", tiny_side)
    check_pairs(eng, s, "dtw")
    check_pairs(dtw_engine_at(16), s, "dtw")
    check_pairs(eng, s, "dtw", band=-10)


def test_dtw_long_pairs_with_tiny_coordinates_bit_exact(eng):
    """The same on pairs long enough for the DP variants that read global
    memory, and for banded pairs past the lean kernels' LDS budget."""
    a_xy, a_off = lattice(47, [6000, 700])
    b_xy, b_off = lattice(48, [6500, 800], below=True)
    s = pair_case(b_xy, b_off, a_xy, a_off)
    say("dtw tiny long")
    check_pairs(eng, s, "dtw")
    check_pairs(dtw_engine_at(16), s, "dtw")
    check_pairs(eng, s, "dtw", band=-10)


def test_set_flags_follow_the_points(eng):
    """The tiny-coordinate flag is found at load, kept by appends and
    recomputed by a reload."""
    xy, off = walks(45, LENGTHS, 2e-2)
    xy = (xy + 2.0).astype(np.float32)  # every coordinate well above 2^-25
    s = ts.TrajSet(xy, off)
    say("set flags")
    assert s.flags() == 0
    s.append(xy[:off[2]], off[:3])
    assert s.flags() == 0
    t_xy, t_off = lattice(46, [5, 9], below=True)
    s.append(t_xy, t_off)
    assert s.flags() == ts.SET_TINY_COORD
    s.reload(xy, off)
    assert s.flags() == 0


@pytest.mark.parametrize("band", [1, 3, 10, 40, -1, -10, -55])
def test_dtw_band_bit_exact(eng, shapes, band):
    say("dtw band", band)
    check_pairs(eng, shapes, "dtw", band=band)


@pytest.mark.parametrize("delta", [0, 5, 16, 64])
def test_lcss_window_bit_exact(eng, shapes, delta):
    say("lcss delta", delta)
    check_pairs(eng, shapes, "lcss", delta=delta, epsilon=3e-2)


@pytest.mark.parametrize("variant,measure,kw", [(0, "dtw", {}), (1, "dtw", {"band": 3}), (1, "dtw", {"band": -20}),
                                                (2, "frechet", {}), (3, "lcss", {"epsilon": 3e-2}),
                                                (4, "lcss", {"epsilon": 3e-2, "delta": 5})])
@pytest.mark.parametrize("waves", [4, 16])
def test_every_dp_shape_bit_exact(shapes, variant, measure, kw, waves):
    """Each lane-group shape the autotuner can choose gives the oracle's
    values bit for bit, in every bucket; at 16 waves per CU unbanded DTW
    runs the step loop split by the tiny-coordinate flag."""
    e = ts.Engine()
    for (W, U) in ts.dp_shapes():
        say("dp shape", measure, kw, W, U, waves)
        for b in range(len(ts.DP_BUCKET_CAPS)):
            e.set_dp_shape(variant, b, W, U, waves)
        check_pairs(e, shapes, measure, **kw)


def test_euclidean_pairs_bit_exact(eng):
    say("euclidean pairs")
    xy, off = walks(23, [128] * 40 + [7] * 10)
    s = ts.TrajSet(xy, off)
    qi = np.repeat(np.arange(50), 50)
    di = np.tile(np.arange(50), 50)
    keep = s.lengths[qi] == s.lengths[di]
    qi, di = qi[keep], di[keep]
    p = ts.params("euclidean")
    got = eng.pair_values(s, s, p, qi, di)
    want = oracle.pair_values(oracle.params(oracle.EUCLIDEAN), xy, off, xy, off, qi, di)
    assert np.array_equal(got.view(np.int32), want.view(np.int32))


def test_long_pairs_bit_exact(eng):
    say("long pairs")
    db_xy, db_off = walks(24, [2048, 3000, 700])
    q_xy, q_off = walks(25, [2500, 900])
    db, qs = ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off)
    qi, di = np.array([0, 0, 1, 1, 0]), np.array([0, 1, 0, 2, 2])
    for measure, kw in [("dtw", {}), ("dtw", {"band": 205}), ("dtw", {"band": -10}), ("frechet", {}),
                        ("lcss", {"epsilon": 2e-3, "delta": 64}), ("lcss", {"epsilon": 2e-3, "delta": 0}),
                        ("lcss", {"epsilon": 2e-3, "delta": 1000})]:
        p = ts.params(measure, **kw)
        prm = oracle.params(p.measure, band=p.band, delta=p.delta, epsilon=p.epsilon)
        got = eng.pair_values(db, qs, p, qi, di)
        want = oracle.pair_values(prm, db_xy, db_off, q_xy, q_off, qi, di)
        assert np.array_equal(got.view(np.int32), want.view(np.int32)), (measure, got, want)
    # unbanded DTW at 16 waves per CU: the split step loop, trajectories read from global memory
    p = ts.params("dtw")
    got = dtw_engine_at(16).pair_values(db, qs, p, qi, di)
    want = oracle.pair_values(oracle.params(p.measure), db_xy, db_off, q_xy, q_off, qi, di)
    assert np.array_equal(got.view(np.int32), want.view(np.int32)), (got, want)


@pytest.mark.parametrize("variant,measure,kw", [(1, "dtw", {"band": -10}), (1, "dtw", {"band": 205}),
                                                (4, "lcss", {"epsilon": 2e-3, "delta": 64}),
                                                (4, "lcss", {"epsilon": 2e-3, "delta": 1000})])
def test_long_ranged_pairs_every_shape_bit_exact(variant, measure, kw):
    """Bands and windows on long pairs with every lane-group shape forced.
    With 16-lane groups, a band 601 columns wide on 3000-point rows needs a
    1024-float halo ring and a 376-float strip table per group, 22.4 KB for
    the wave's 4 groups, more than the 20 KiB a wave of the lean banded kernel
    may take, so those launches run k_dp; so do 16- and 32-lane groups with a
    window of 1000 (a 2500-float halo row). The other shapes run k_dp_lean_r
    with both trajectories read from global memory."""
    db_xy, db_off = walks(24, [2048, 3000, 700])
    q_xy, q_off = walks(25, [2500, 900])
    db, qs = ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off)
    qi, di = np.array([0, 0, 1, 1, 0]), np.array([0, 1, 0, 2, 2])
    p = ts.params(measure, **kw)
    prm = oracle.params(p.measure, band=p.band, delta=p.delta, epsilon=p.epsilon)
    want = oracle.pair_values(prm, db_xy, db_off, q_xy, q_off, qi, di)
    e = ts.Engine()
    for (W, U) in ts.dp_shapes():
        say("long ranged pairs", measure, kw, W, U)
        for b in range(len(ts.DP_BUCKET_CAPS)):
            e.set_dp_shape(variant, b, W, U, 4)
        got = e.pair_values(db, qs, p, qi, di)
        assert np.array_equal(got.view(np.int32), want.view(np.int32)), (W, U, got, want)


def test_long_rows_short_columns_bit_exact(eng):
    """Rows too long to stage beside the columns (the unbanded DP then reads
    row points from global memory), and rows and columns both too long."""
    say("long rows")
    db_xy, db_off = walks(28, [200, 150, 3000, 1100])
    q_xy, q_off = walks(29, [2500, 250, 40])
    db, qs = ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off)
    qi = np.array([0, 0, 1, 1, 2, 0, 1])
    di = np.array([0, 1, 2, 3, 2, 3, 0])
    for measure in ("dtw", "frechet", "lcss"):
        p = ts.params(measure, epsilon=2e-3)
        prm = oracle.params(p.measure, epsilon=p.epsilon)
        got = eng.pair_values(db, qs, p, qi, di)
        want = oracle.pair_values(prm, db_xy, db_off, q_xy, q_off, qi, di)
        assert np.array_equal(got.view(np.int32), want.view(np.int32)), (measure, got, want)


@pytest.mark.parametrize("measure", ["dtw", "frechet", "lcss", "hausdorff", "euclidean"])
def test_filter_ref_bit_exact(eng, measure):
    say("filter_ref", measure)
    L = uniform_lengths(26, 700, 32, 128)
    if measure == "euclidean":
        L[::3] = 64
    db_xy, db_off = walks(26, L)
    q_xy, q_off = planted(db_xy, db_off, np.arange(0, 700, 70), L[:10] if measure != "euclidean" else [64] * 10,
                          1e-3, 2)
    db, qs = ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off)
    p = ts.params(measure, epsilon=2e-3)
    got = eng.filter_ref(db, qs, p)
    if measure in ("dtw", "frechet", "lcss"):
        want = oracle.filter_ref(oracle.params(p.measure, epsilon=p.epsilon), db_xy, db_off, q_xy, q_off)
        assert np.array_equal(got.view(np.int32), want.view(np.int32))
    else:
        for qi in range(10):
            mq = oracle.mbr(q_xy[q_off[qi]:q_off[qi + 1]])
            for j in range(0, 700, 7):
                c = db_xy[db_off[j]:db_off[j + 1]]
                if measure == "euclidean" and len(c) != q_off[qi + 1] - q_off[qi]:
                    assert np.isinf(got[qi, j])
                else:
                    assert got[qi, j] == oracle.box_box_d2(mq, oracle.mbr(c))


def test_topk_rows_order_and_ties(eng):
    say("topk rows")
    rng = np.random.default_rng(27)
    for rows, cols, k in [(3, 5, 5), (7, 1000, 10), (4, 5000, 100), (2, 70000, 37), (5, 64, 1)]:
        s = np.round(rng.random((rows, cols)) * 50).astype(np.float32)  # many ties
        for desc in (False, True):
            idx, val = eng.topk_rows(s, k, desc)
            for r in range(rows):
                order = np.lexsort((np.arange(cols), -s[r] if desc else s[r]))[:k]
                assert np.array_equal(idx[r, :min(k, cols)], order), (rows, cols, k, desc)
                assert np.array_equal(val[r, :min(k, cols)], s[r, order])


def knn_case(measure, n_db=1500, n_q=12, lo=32, hi=128, seed=30):
    L = uniform_lengths(seed, n_db, lo, hi)
    if measure == "euclidean":
        L[::2] = 96
    db_xy, db_off = walks(seed, L)
    picks = np.arange(1, n_db, n_db // n_q)[:n_q]
    qL = L[picks] if measure != "euclidean" else np.full(n_q, 96)
    if measure == "euclidean":
        picks = picks - (picks % 2)
    q_xy, q_off = planted(db_xy, db_off, picks, qL, 5e-4, seed)
    return db_xy, db_off, q_xy, q_off


@pytest.mark.parametrize("measure", ["dtw", "frechet", "lcss", "hausdorff", "euclidean"])
@pytest.mark.parametrize("mode", [ts.REF, ts.EXACT])
def test_knn_batched_and_plain_match_oracle(eng, measure, mode):
    say("knn", measure, "mode", mode)
    db_xy, db_off, q_xy, q_off = knn_case(measure)
    db, qs = ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off)
    k = 10 if mode == ts.EXACT else 5
    p = ts.params(measure, mode=mode, k=k, epsilon=2e-3)
    prm = oracle.params(p.measure, p.mode, k=p.k, kc=p.kc, r=p.r, band=p.band, delta=p.delta, epsilon=p.epsilon)
    wi, wv, _ = oracle.knn(prm, db_xy, db_off, q_xy, q_off)
    for fn in (eng.knn, eng.knn_plain):
        gi, gv, st = fn(db, qs, p)
        assert np.array_equal(gi, wi), (fn.__name__, gi[:2], wi[:2])
        assert np.array_equal(gv.view(np.int32), wv.view(np.int32)), (fn.__name__, gv[:2], wv[:2])


@pytest.mark.parametrize("measure", ["dtw", "frechet", "lcss", "hausdorff"])
@pytest.mark.parametrize("n_db,n_q,k", [(300, 3, 10), (6, 2, 10), (1, 1, 3)])
def test_knn_small_all_pairs_match_oracle(eng, measure, n_db, n_q, k):
    """Small EXACT problems (every pair refined in one DP round), including
    k larger than the DB."""
    say("knn small", measure, n_db, n_q, k)
    L = uniform_lengths(33, n_db, 8, 200)
    db_xy, db_off = walks(33, L, 5e-3)
    q_xy, q_off = planted(db_xy, db_off, np.arange(n_q) % n_db, (L[np.arange(n_q) % n_db] * 3) // 4 + 1, 1e-3, 34)
    db, qs = ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off)
    p = ts.params(measure, mode=ts.EXACT, k=k, epsilon=2e-3)
    prm = oracle.params(p.measure, p.mode, k=p.k, epsilon=p.epsilon)
    wi, wv, _ = oracle.knn(prm, db_xy, db_off, q_xy, q_off)
    for fn in (eng.knn, eng.knn_plain):
        gi, gv, st = fn(db, qs, p)
        assert np.array_equal(gi, wi), (fn.__name__, gi, wi)
        assert np.array_equal(gv.view(np.int32), wv.view(np.int32)), (fn.__name__, gv, wv)


@pytest.mark.parametrize("band", [-10, 3, -50])
def test_knn_banded_dtw_match_oracle(eng, band):
    """EXACT DTW with a band over 400 walks (many candidates near the k-th
    value, pruned by the banded point-to-window bound) equals the oracle."""
    say("knn banded dtw", band)
    L = uniform_lengths(35, 400, 40, 300)
    db_xy, db_off = walks(35, L, 4e-3)
    q_xy, q_off = planted(db_xy, db_off, np.arange(0, 400, 20), (L[0:400:20] * 5) // 6 + 3, 2e-3, 36)
    db, qs = ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off)
    p = ts.params("dtw", mode=ts.EXACT, k=10, band=band)
    prm = oracle.params(p.measure, p.mode, k=p.k, band=p.band)
    wi, wv, _ = oracle.knn(prm, db_xy, db_off, q_xy, q_off)
    for fn in (eng.knn, eng.knn_plain):
        gi, gv, st = fn(db, qs, p)
        assert np.array_equal(gi, wi), (fn.__name__, gi, wi)
        assert np.array_equal(gv.view(np.int32), wv.view(np.int32)), (fn.__name__, gv, wv)


@pytest.mark.parametrize("delta", [0, 5, 64])
def test_knn_windowed_lcss_match_oracle(eng, delta):
    """EXACT LCSS with an index window over 400 walks (pruned by the window
    box count) equals the oracle, ties at similarity 0 included."""
    say("knn windowed lcss", delta)
    L = uniform_lengths(37, 400, 40, 300)
    db_xy, db_off = walks(37, L, 4e-3)
    q_xy, q_off = planted(db_xy, db_off, np.arange(0, 400, 20), L[0:400:20], 2e-3, 38)
    db, qs = ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off)
    p = ts.params("lcss", mode=ts.EXACT, k=10, epsilon=5e-3, delta=delta)
    prm = oracle.params(p.measure, p.mode, k=p.k, delta=p.delta, epsilon=p.epsilon)
    wi, wv, _ = oracle.knn(prm, db_xy, db_off, q_xy, q_off)
    for fn in (eng.knn, eng.knn_plain):
        gi, gv, st = fn(db, qs, p)
        assert np.array_equal(gi, wi), (fn.__name__, gi, wi)
        assert np.array_equal(gv.view(np.int32), wv.view(np.int32)), (fn.__name__, gv, wv)


@pytest.mark.parametrize("measure,kw", [("dtw", {"band": -10}), ("lcss", {"epsilon": 2e-3, "delta": 64})])
def test_knn_long_banded_match_oracle(eng, measure, kw):
    """Long-trajectory shape (lengths U[2048,4096]) with a per-pair 10% band
    or an LCSS window: EXACT top-5 through the banded multi-strip DP."""
    say("knn long", measure, kw)
    L = uniform_lengths(31, 60, 2048, 4096)
    db_xy, db_off = walks(31, L, 2e-3)
    q_xy, q_off = planted(db_xy, db_off, np.array([3, 40]), L[[3, 40]] // 2 + 17, 5e-4, 32)
    db, qs = ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off)
    p = ts.params(measure, mode=ts.EXACT, k=5, **kw)
    prm = oracle.params(p.measure, p.mode, k=p.k, band=p.band, delta=p.delta, epsilon=p.epsilon)
    wi, wv, _ = oracle.knn(prm, db_xy, db_off, q_xy, q_off)
    for fn in (eng.knn, eng.knn_plain):
        gi, gv, st = fn(db, qs, p)
        assert np.array_equal(gi, wi), (fn.__name__, gi, wi)
        assert np.array_equal(gv.view(np.int32), wv.view(np.int32)), (fn.__name__, gv, wv)


def test_compat_reproduces_reference_fixtures(eng):
    """The re-exported reference signatures on the GPU give exactly the
    h_results of the reference's own CPU build (tests/golden)."""
    lib = ctypes.CDLL(ts.LIB_PATH)
    fns = {oracle.DTW: lib._Z24launch_dtw_batch_gpu_knnPK5PointS1_PfiiiRf,
           oracle.FRECHET: lib._Z34launch_frechet_batch_gpu_wavefrontPK5PointS1_PfiiiRf,
           oracle.LCSS: lib._Z35launch_lcss_batch_gpu_wavefront_knnPK5PointS1_PfiiifRf}
    files = sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))
    assert files
    for f in files:
        say("compat", f)
        z = np.load(os.path.join(GOLDEN, f))
        num_t, n, m, meas = int(z["num_t"]), int(z["n"]), int(z["m"]), int(z["measure"])
        t1 = np.ascontiguousarray(z["t1"], np.float32)
        t2 = np.ascontiguousarray(z["t2"], np.float32)
        out = np.zeros(num_t, np.float32)
        ms = ctypes.c_float()
        P = ctypes.c_void_p
        if meas == oracle.LCSS:
            fns[meas](P(t1.ctypes.data), P(t2.ctypes.data), P(out.ctypes.data), num_t, n, m,
                      ctypes.c_float(float(z["eps"])), ctypes.byref(ms))
        else:
            fns[meas](P(t1.ctypes.data), P(t2.ctypes.data), P(out.ctypes.data), num_t, n, m, ctypes.byref(ms))
        assert np.array_equal(out.view(np.int32), z["h_results"].view(np.int32)), f


def test_tuning_table_lines_fill_their_tables(tmp_path):
    """Plain lines fill every table not yet given its own line; "lat", "mid"
    and "few" lines fill the small-, mid-size and few-pair tables; bad lines
    are refused."""
    e = ts.Engine()
    try:
        t = tmp_path / "table.txt"
        t.write_text("dtw 3 32 4 8\n"
                     "dtw 2 16 2 6 lat\n"
                     "dtw 2 64 4 12 mid\n"
                     "dtw 2 128 4 4 few\n"
                     "dtw 2 32 8 16  # throughput after lat, mid and few\n")
        e.load_tuning(str(t))
        for table in (0, 1, 3, 4):
            assert e.get_dp_shape(0, 3, table) == (32, 4, 8)
        assert e.get_dp_shape(0, 2, 0) == (32, 8, 16)
        assert e.get_dp_shape(0, 2, 1) == (16, 2, 6)
        assert e.get_dp_shape(0, 2, 3) == (64, 4, 12)
        assert e.get_dp_shape(0, 2, 4) == (128, 4, 4)
        e.set_dp_shape(0, 2, 64, 2, 4, table=3)
        assert e.get_dp_shape(0, 2, 3) == (64, 2, 4) and e.get_dp_shape(0, 2, 1) == (16, 2, 6)
        e.set_dp_shape(0, 2, 256, 8, 4, table=4)
        assert e.get_dp_shape(0, 2, 4) == (256, 8, 4) and e.get_dp_shape(0, 2, 3) == (64, 2, 4)
        bad = tmp_path / "bad.txt"
        bad.write_text("dtw 2 64 4 12 fast\n")
        with pytest.raises(ts.TrajsimError):
            e.load_tuning(str(bad))
        with pytest.raises(ts.TrajsimError):
            e.get_dp_shape(0, 2, 2)
    finally:
        e.close()


def test_pair_values_refuses_more_than_int32_pairs(eng, shapes):
    """A pair's value slot is a 32-bit index: 2^31 pairs in one call are
    refused before any array is read (the arrays here hold 4 entries)."""
    lib = ctypes.CDLL(ts.LIB_PATH)
    qi = np.zeros(4, np.int32)
    out = np.zeros(4, np.float32)
    p = ts.params("dtw")
    st = lib.ts_pair_values(ctypes.c_void_p(eng._h.value), ctypes.c_void_p(shapes["db"].handle.value),
                            ctypes.c_void_p(shapes["qs"].handle.value), ctypes.byref(p),
                            qi.ctypes.data_as(ctypes.c_void_p), qi.ctypes.data_as(ctypes.c_void_p),
                            ctypes.c_int64(2 ** 31), out.ctypes.data_as(ctypes.c_void_p))
    assert st == 4, st  # TS_ERR_UNSUPPORTED
