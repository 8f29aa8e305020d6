"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker.

Each function of the oracle is checked against something that does not
restate its own formula: the reference's own CPU path compiled from the
reference sources (oracle/_ref), the golden fixtures made from it
(tests/golden/), brute force over all warping paths / couplings /
subsequences, textbook known answers, closed forms and library routines
(numpy, scipy). A planted mistake (a dropped term, a swapped index) fails at
least one of them.
"""
import itertools
import os

import numpy as np
import pytest

import oracle
from tests.helpers import planted, traj, uniform_lengths, walks

F = np.float32
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def f32_dist(a, b):
    dx = F(a[0]) - F(b[0])
    dy = F(a[1]) - F(b[1])
    return np.sqrt(F(dx * dx + dy * dy))


def f64_dist(a, b):
    return float(np.hypot(float(a[0]) - float(b[0]), float(a[1]) - float(b[1])))


def paths(n, m):
    """All monotone warping paths from (0,0) to (n-1,m-1)."""
    def rec(i, j):
        if (i, j) == (n - 1, m - 1):
            yield [(i, j)]
            return
        for di, dj in ((1, 0), (0, 1), (1, 1)):
            a, b = i + di, j + dj
            if a < n and b < m:
                for rest in rec(a, b):
                    yield [(i, j)] + rest
    return list(rec(0, 0))


def rand_traj(rng, n, scale=1.0):
    return (rng.random((n, 2)) * scale).astype(np.float32)


# ------------------------------------------------------------- envelope
def test_envelope_is_windowed_min_max():
    rng = np.random.default_rng(1)
    for n, r in [(1, 5), (3, 5), (11, 5), (40, 5), (17, 0), (9, 2)]:
        q = rand_traj(rng, n)
        env = oracle.envelope(q, r)
        for i in range(n):
            w = q[max(0, i - r):min(n, i + r + 1)]
            assert np.array_equal(env[i], [w[:, 0].min(), w[:, 0].max(), w[:, 1].min(), w[:, 1].max()])


# ------------------------------------------------------------- filters
def box_dist_f32(e, p):
    dx = max(F(0), max(F(e[0] - p[0]), F(p[0] - e[1])))
    dy = max(F(0), max(F(e[2] - p[1]), F(p[1] - e[3])))
    return np.sqrt(F(dx * dx + dy * dy))


def test_filters_match_per_point_closed_form():
    rng = np.random.default_rng(2)
    for n, m in [(1, 1), (5, 3), (3, 5), (13, 29), (64, 17)]:
        q, c = rand_traj(rng, n), rand_traj(rng, m)
        env = oracle.envelope(q, 5)
        s, mx, cnt = F(0), F(0), 0
        eps = F(0.05)
        for v in range(m):
            e = env[v * n // m]
            d = box_dist_f32(e, c[v])
            s = F(s + d)
            mx = max(mx, d)
            cnt += int(c[v, 0] > F(e[0] - eps) and c[v, 0] < F(e[1] + eps) and
                       c[v, 1] > F(e[2] - eps) and c[v, 1] < F(e[3] + eps))
        assert oracle.lb_dtw(env, c) == s
        assert oracle.lb_frechet(env, c) == mx
        assert oracle.ub_lcss_count(env, c, eps) == cnt


# ------------------------------------------------------------- DTW
def test_dtw_brute_force_over_warping_paths():
    rng = np.random.default_rng(3)
    for n in range(1, 6):
        for m in range(1, 6):
            q, c = rand_traj(rng, n), rand_traj(rng, m)
            best = min(sum(f64_dist(q[i], c[j]) for i, j in p) for p in paths(n, m))
            assert abs(oracle.dtw(q, c, f64=True) - best) <= 1e-12 * max(1.0, best)
            assert abs(oracle.dtw(q, c) - best) <= 1e-5 * max(1e-3, best)


def test_dtw_band_brute_force():
    rng = np.random.default_rng(4)
    for n, m, w in [(5, 5, 1), (5, 5, 2), (4, 6, 1), (6, 4, 2), (6, 6, 6), (3, 5, 1)]:
        q, c = rand_traj(rng, n), rand_traj(rng, m)
        inb = lambda i, j: abs(i * m - j * n) <= w * max(n, m)
        cand = [p for p in paths(n, m) if all(inb(i, j) for i, j in p)]
        assert cand, "band keeps a path"
        best = min(sum(f64_dist(q[i], c[j]) for i, j in p) for p in cand)
        assert abs(oracle.dtw(q, c, band=w, f64=True) - best) <= 1e-12 * max(1.0, best)
    # a band at least as wide as the longer side is the unconstrained DTW
    q, c = rand_traj(rng, 30), rand_traj(rng, 21)
    assert oracle.dtw(q, c, band=30) == oracle.dtw(q, c)
    # a negative band is a percentage of the longer length, rounded up
    assert oracle.dtw(q, c, band=-100) == oracle.dtw(q, c)
    assert oracle.dtw(q, c, band=-10) == oracle.dtw(q, c, band=3)  # ceil(0.1 * 30)
    assert oracle.dtw(c, q, band=-34) == oracle.dtw(c, q, band=11)  # ceil(0.34 * 30)
    assert oracle.dtw(q, c, band=-1) == oracle.dtw(q, c, band=1)


def test_dtw_band_matches_masked_full_dp():
    """The banded oracle visits only in-band columns; it must equal the full
    f32 recurrence with out-of-band cells set to 1e20f, bit for bit."""
    rng = np.random.default_rng(41)
    F = np.float32
    for _ in range(60):
        n, m = int(rng.integers(1, 28)), int(rng.integers(1, 28))
        band = int(rng.choice([1, 2, 3, 7, 40, -1, -10, -34, -100]))
        q, c = rand_traj(rng, n), rand_traj(rng, m)
        w = band if band > 0 else (-band * max(n, m) + 99) // 100
        D = np.full((n + 1, m + 1), F(1e20), np.float32)
        D[0, 0] = 0
        for i in range(1, n + 1):
            for j in range(1, m + 1):
                if abs((i - 1) * m - (j - 1) * n) > w * max(n, m):
                    continue
                dx, dy = F(q[i - 1, 0] - c[j - 1, 0]), F(q[i - 1, 1] - c[j - 1, 1])
                d = np.sqrt(F(F(dx * dx) + F(dy * dy)))
                D[i, j] = F(d + min(D[i - 1, j], D[i, j - 1], D[i - 1, j - 1]))
        assert oracle.dtw(q, c, band=band) == D[n, m], (n, m, band)


def test_dtw_closed_forms_and_symmetry():
    rng = np.random.default_rng(5)
    a = rand_traj(rng, 37)
    assert oracle.dtw(a, a) == 0.0
    b = rand_traj(rng, 23)
    # n = 1: the single query point is matched to every DB point, in order
    s = F(0)
    for j in range(len(b)):
        s = F(s + f32_dist(a[0], b[j]))
    assert oracle.dtw(a[:1], b) == s
    # symmetry is bit-exact: the recurrence and the point cost are symmetric
    assert oracle.dtw(a, b) == oracle.dtw(b, a)
    assert oracle.dtw(a, b, band=3) == oracle.dtw(b, a, band=3)
    # equal lengths: the diagonal path is one warping path (upper bound)
    c = rand_traj(rng, 37)
    diag = sum(f64_dist(a[i], c[i]) for i in range(37))
    assert oracle.dtw(a, c, f64=True) <= diag + 1e-12


# ------------------------------------------------------------- Frechet
def test_frechet_brute_force_over_couplings_bit_exact():
    rng = np.random.default_rng(6)
    for n in range(1, 6):
        for m in range(1, 6):
            q, c = rand_traj(rng, n), rand_traj(rng, m)
            best = min(max(f32_dist(q[i], c[j]) for i, j in p) for p in paths(n, m))
            assert oracle.frechet(q, c) == best


def test_frechet_closed_forms():
    rng = np.random.default_rng(7)
    a, b = rand_traj(rng, 19), rand_traj(rng, 31)
    assert oracle.frechet(a, a) == 0.0
    assert oracle.frechet(a[:1], b) == max(f32_dist(a[0], p) for p in b)
    f = oracle.frechet(a, b)
    assert f >= max(f32_dist(a[0], b[0]), f32_dist(a[-1], b[-1]))
    assert f == oracle.frechet(b, a)


# ------------------------------------------------------------- LCSS
def lattice(word):
    return np.array([[float("ABCD".index(ch)), 0.0] for ch in word], np.float32)


def test_lcss_textbook_lcs_known_answer():
    # CLRS: LCS("ABCBDAB", "BDCABA") = 4. Points on a lattice of spacing 1
    # with eps 0.5 match exactly when the letters are equal.
    assert oracle.lcss(lattice("ABCBDAB"), lattice("BDCABA"), 0.5) == 4
    assert oracle.lcss(lattice("BDCABA"), lattice("ABCBDAB"), 0.5) == 4
    assert oracle.lcss(lattice("AAAA"), lattice("BBB"), 0.5) == 0
    assert oracle.lcss(lattice("ABCD"), lattice("ABCD"), 0.5) == 4


def lcss_brute(q, c, eps, delta):
    n, m = len(q), len(c)
    match = lambda i, j: (abs(F(q[i, 0] - c[j, 0])) < F(eps) and abs(F(q[i, 1] - c[j, 1])) < F(eps)
                          and (delta < 0 or abs(i - j) <= delta))
    best = 0
    for r in range(1, min(n, m) + 1):
        for sub in itertools.combinations(range(n), r):
            j = 0
            ok = True
            for i in sub:  # earliest feasible partner keeps the most room
                while j < m and not match(i, j):
                    j += 1
                if j == m:
                    ok = False
                    break
                j += 1
            if ok:
                best = r
    return best


def test_lcss_brute_force_with_and_without_window():
    rng = np.random.default_rng(8)
    for n, m in [(1, 4), (4, 1), (5, 5), (6, 4), (4, 7), (7, 6)]:
        q = (rng.integers(0, 3, (n, 2)) * 0.1).astype(np.float32)
        c = (rng.integers(0, 3, (m, 2)) * 0.1).astype(np.float32)
        for delta in (-1, 0, 1, 2, 10):
            assert oracle.lcss(q, c, 0.05, delta) == lcss_brute(q, c, 0.05, delta), (n, m, delta)


def test_lcss_bounds_and_symmetry():
    xy, off = walks(11, [50, 70])
    a, b = traj(xy, off, 0), traj(xy, off, 1)
    for eps, delta in [(2e-3, -1), (5e-2, 16), (1.0, -1)]:
        v = oracle.lcss(a, b, eps, delta)
        assert 0 <= v <= 50 and v == oracle.lcss(b, a, eps, delta)
    assert oracle.lcss(a, b, 10.0) == 50  # every pair matches
    assert oracle.lcss(a, a, 1e-6) == 50


# ------------------------------------------------------------- Hausdorff / Euclidean
def test_hausdorff_matches_numpy_and_scipy():
    from scipy.spatial.distance import directed_hausdorff
    rng = np.random.default_rng(9)
    for n, m in [(1, 1), (1, 9), (33, 12), (70, 130)]:
        q, c = rand_traj(rng, n), rand_traj(rng, m)
        d = q[:, None, :] - c[None, :, :]  # float32 broadcasting: same f32 ops
        d2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]
        h2 = max(d2.min(axis=1).max(), d2.min(axis=0).max())
        assert oracle.hausdorff(q, c) == np.sqrt(F(h2))
        ref = max(directed_hausdorff(q.astype(np.float64), c.astype(np.float64))[0],
                  directed_hausdorff(c.astype(np.float64), q.astype(np.float64))[0])
        assert abs(oracle.hausdorff(q, c, f64=True) - ref) <= 1e-12
    q = rand_traj(rng, 20)
    assert oracle.hausdorff(q, q) == 0.0


def test_euclidean_closed_form():
    rng = np.random.default_rng(10)
    for n in (1, 7, 128):
        q, c = rand_traj(rng, n), rand_traj(rng, n)
        ref = np.sqrt(((q.astype(np.float64) - c.astype(np.float64)) ** 2).sum())
        assert abs(float(oracle.euclidean(q, c)) - ref) <= 1e-7 * max(ref, 1e-30)


def test_mbr_and_box_distance():
    rng = np.random.default_rng(11)
    a, b = rand_traj(rng, 10), rand_traj(rng, 10) + 2.0
    ma, mb = oracle.mbr(a), oracle.mbr(b)
    assert np.array_equal(ma, [a[:, 0].min(), a[:, 0].max(), a[:, 1].min(), a[:, 1].max()])
    gx, gy = mb[0] - ma[1], mb[2] - ma[3]
    assert oracle.box_box_d2(ma, mb) == F(F(gx * gx) + F(gy * gy))
    assert oracle.box_box_d2(ma, ma) == 0.0


# ------------------------------------------------------------- EXACT bounds
@pytest.mark.parametrize("measure", [oracle.DTW, oracle.FRECHET, oracle.HAUSDORFF, oracle.LCSS, oracle.EUCLIDEAN])
def test_summary_bounds_are_admissible(measure):
    L = uniform_lengths(12, 60, 2, 40)
    if measure == oracle.EUCLIDEAN:
        L[:] = 24
    xy, off = walks(12, L, 2e-2)
    prm = oracle.params(measure, oracle.EXACT, epsilon=3e-2, band=0)
    qi, di = np.meshgrid(np.arange(60), np.arange(60), indexing="ij")
    exact = oracle.pair_values(prm, xy, off, xy, off, qi.ravel(), di.ravel()).reshape(60, 60)
    for i in range(60):
        for j in range(60):
            b = oracle.summary_bound(prm, traj(xy, off, i), traj(xy, off, j))
            if measure == oracle.LCSS:
                assert b >= exact[i, j]
            else:
                assert b <= exact[i, j]


# ------------------------------------------------------------- whole k-NN
@pytest.mark.parametrize("measure", [oracle.DTW, oracle.FRECHET, oracle.LCSS, oracle.HAUSDORFF, oracle.EUCLIDEAN])
def test_exact_knn_pruned_equals_brute_force(measure):
    L = uniform_lengths(13, 300, 8, 48)
    if measure == oracle.EUCLIDEAN:
        L[:] = 32
    db_xy, db_off = walks(13, L)
    # duplicates create exact ties that must break by index
    db_xy = db_xy.copy()
    db_xy[db_off[5]:db_off[6]] = db_xy[db_off[4]:db_off[5]] if L[4] == L[5] else db_xy[db_off[5]:db_off[6]]
    picks = np.arange(0, 300, 37)
    q_xy, q_off = planted(db_xy, db_off, picks, L[picks], 5e-4, 1)
    for k in (1, 10):
        prm = oracle.params(measure, oracle.EXACT, k=k, epsilon=2e-3)
        i1, v1, r1 = oracle.knn(prm, db_xy, db_off, q_xy, q_off, brute=False)
        i2, v2, r2 = oracle.knn(prm, db_xy, db_off, q_xy, q_off, brute=True)
        assert np.array_equal(i1, i2) and np.array_equal(v1, v2)
        if measure != oracle.LCSS:
            assert r1 < r2  # pruning did something
    # planted queries find their source trajectory
    prm = oracle.params(measure, oracle.EXACT, k=1, epsilon=2e-3)
    i1, _, _ = oracle.knn(prm, db_xy, db_off, q_xy, q_off)
    assert (i1[:, 0] == picks).mean() >= 0.8


@pytest.mark.skipif(not oracle.ref_available(), reason="reference sources not built here")
@pytest.mark.parametrize("measure", [oracle.DTW, oracle.FRECHET, oracle.LCSS])
def test_ref_mode_equals_reference_build(measure):
    """Oracle REF mode at the reference's shapes (Q == N, fixed lengths) is
    bit-identical to the reference's launch_*_batch_cpu."""
    for num_t, n, m, eps in [(40, 50, 70, 2e-3), (30, 64, 64, 5e-3), (12, 90, 33, 0.01), (5, 8, 8, 0.5)]:
        q_xy, q_off = walks(100 + n, [n] * num_t)
        db_xy, db_off = planted(q_xy, q_off, np.arange(num_t), [m] * num_t, 1e-3, 7)
        ref, _ = oracle.ref_batch(measure, q_xy, db_xy, num_t, n, m, eps)
        prm = oracle.params(measure, oracle.REF, k=1, kc=10, epsilon=eps)
        _, val, _ = oracle.knn(prm, db_xy, db_off, q_xy, q_off)
        assert np.array_equal(val[:, 0], ref), (measure, num_t, n, m)


def test_ref_mode_matches_golden_fixtures():
    files = sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))
    assert files, "golden fixtures missing"
    for f in files:
        z = np.load(os.path.join(GOLDEN, f))
        num_t, n, m = int(z["num_t"]), int(z["n"]), int(z["m"])
        q_off = np.arange(num_t + 1, dtype=np.int64) * n
        db_off = np.arange(num_t + 1, dtype=np.int64) * m
        prm = oracle.params(int(z["measure"]), oracle.REF, k=1, kc=10, epsilon=float(z["eps"]))
        _, val, _ = oracle.knn(prm, z["t2"], db_off, z["t1"], q_off)
        assert np.array_equal(val[:, 0], z["h_results"]), f


def test_f32_error_against_f64_oracle():
    """States the f32 error of each distance measure against the FP64 oracle
    on random-walk pairs (the same inputs in double arithmetic)."""
    L = uniform_lengths(14, 40, 64, 256)
    xy, off = walks(14, L)
    q_xy, q_off = planted(xy, off, np.arange(40), L[::-1], 1e-3, 3)
    qi = np.repeat(np.arange(40), 5)
    di = (qi + np.tile(np.arange(5), 40)) % 40
    for measure, bound in [(oracle.DTW, 1e-5), (oracle.FRECHET, 1e-6), (oracle.HAUSDORFF, 1e-6)]:
        prm = oracle.params(measure, oracle.EXACT)
        v32 = oracle.pair_values(prm, xy, off, q_xy, q_off, qi, di).astype(np.float64)
        v64 = oracle.pair_values(prm, xy, off, q_xy, q_off, qi, di, f64=True)
        rel = np.abs(v32 - v64) / np.maximum(v64, 1e-30)
        assert rel.max() < bound, (measure, rel.max())
