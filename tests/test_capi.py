"""CPU-side checks of the C ABI: the library loads, exports every function
that include/trajsim.h declares and the 12 reference signatures of
include/trajsim_compat.hpp, and the synthetic generator is deterministic.
No compute call touches a GPU here."""
import ctypes
import os
import re
import subprocess

import numpy as np

import cudaruntime1_amd as ts

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "trajsim.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ts_[a-z0-9_]+)\s*\(", src)))


def test_header_functions_are_exported():
    names = declared_functions()
    assert len(names) >= 25
    lib = ctypes.CDLL(ts.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(ts.exported_symbols())


COMPAT = ["launch_dtw_batch_gpu_knn", "launch_dtw_batch_cpu", "launch_frechet_batch_gpu_wavefront",
          "launch_frechet_batch_cpu", "launch_lcss_batch_gpu_wavefront_knn", "launch_lcss_batch_cpu",
          "launch_hausdorff_batch_cpu_rtree", "run_hausdorff_rtree_gpu_pipeline", "launch_hausdorff_rtree_gpu",
          "launch_euclidean_batch_cpu_rtree", "launch_euclidean_batch_gpu_rtree_exact",
          "execute_rtree_csr_kernel_on_gpu"]


def test_compat_signatures_are_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", ts.LIB_PATH], capture_output=True, text=True).stdout
    for name in COMPAT:
        assert re.search(r"\b%s\(Point const\*" % name, out), name
    # exact reference signature of one of them (dtw.cuh:6)
    assert "launch_dtw_batch_gpu_knn(Point const*, Point const*, float*, int, int, int, float&)" in out
    assert "launch_hausdorff_rtree_gpu(Point const*, int, int, Point const*, int, int, int const*, int, float*, float&)" in out


def test_generator_is_deterministic_and_shaped():
    L1 = ts.gen_lengths(7, 1000, 16, 1024, True)
    L2 = ts.gen_lengths(7, 1000, 16, 1024, True)
    assert np.array_equal(L1, L2) and L1.min() >= 16 and L1.max() <= 1024
    # log-uniform: median near sqrt(16*1025)
    assert 80 < np.median(L1) < 200
    Lu = ts.gen_lengths(8, 4000, 64, 256, False)
    assert Lu.min() == 64 and Lu.max() == 256 and abs(Lu.mean() - 160) < 4
    xy, off = ts.gen_walks(9, Lu[:50])
    xy2, _ = ts.gen_walks(9, Lu[:50])
    assert np.array_equal(xy, xy2)
    starts = xy[off[:-1]]
    assert (starts >= 0).all() and (starts < 1).all()
    steps = np.diff(xy[off[0]:off[1]].astype(np.float64), axis=0)
    assert 5e-4 < steps.std() < 2e-3


def test_params_defaults_match_reference_knobs():
    p = ts.params("dtw")
    assert (p.k, p.kc, p.r, p.band, p.delta) == (10, 10, 5, 0, -1) and abs(p.epsilon - 0.5) < 1e-7
    assert ts.params("euclidean").kc == 50


def test_init_reports_status():
    # no GPU in the build container: a clean status, never a crash
    rc = ts._lib.ts_init(0)
    assert rc in (0, 5)


def test_committed_tuning_table_is_well_formed():
    """Every line of the DP tuning table beside the library parses as the
    engine's loader reads it: "<variant> <bucket> <lanes> <cells>
    <waves_per_cu> [lat|mid|few]" with a known variant, bucket and shape,
    waves per CU in 1..64, '#' comments, lines shorter than the loader's
    buffer; every (variant, bucket) has a line for each of the four tables."""
    path = os.path.join(os.path.dirname(ts.LIB_PATH), "dp_tuning.txt")
    shapes = {(16, 1), (16, 2), (16, 4), (32, 2), (32, 4), (64, 2), (64, 4), (32, 8), (64, 8), (128, 4), (128, 8),
              (256, 4), (256, 8)}
    seen = set()
    for n, line in enumerate(open(path), 1):
        assert len(line) < 1024, n
        body = line.split("#")[0].split()
        if not body:
            continue
        assert len(body) in (5, 6), (n, line)
        v, b, W, U, g = body[0], int(body[1]), int(body[2]), int(body[3]), int(body[4])
        assert v in ts.DP_VARIANTS and 0 <= b < len(ts.DP_BUCKET_CAPS), (n, line)
        assert (W, U) in shapes and 1 <= g <= 64, (n, line)
        table = body[5] if len(body) == 6 else ""
        assert table in ("", "lat", "mid", "few"), (n, line)
        seen.add((v, b, table))
    assert len(seen) == len(ts.DP_VARIANTS) * len(ts.DP_BUCKET_CAPS) * 4


def _cells_by_rows(measure, band, delta, n, m):
    """DP cells counted row by row: every cell, the rows' in-band column
    ranges of a DTW band (|r*C - j*R| <= w*R), or the window |r - j| <= delta."""
    R, C = max(n, m), min(n, m)
    if measure == "dtw" and band != 0:
        bw = (band if band > 0 else (-band * R + 99) // 100) * R
        tot = 0
        for r in range(R):
            lo = max(0, -((bw - r * C) // R))  # ceil((r*C - bw) / R), clamped
            hi = min(C - 1, (r * C + bw) // R)
            tot += max(0, hi - lo + 1)
        return tot
    if measure == "lcss" and delta >= 0:
        return sum(min(C - 1, r + delta) - max(0, r - delta) + 1 for r in range(min(R, C + delta)))
    return n * m


def test_dp_cells_closed_forms_match_row_counts():
    """ts_dp_cells (closed forms over the rows, as the engine's cell counters
    use them) equals the row-by-row count of in-band DTW cells and in-window
    LCSS cells, over random lengths, bands and windows and their edges."""
    rng = np.random.default_rng(11)
    cases = [(1, 1), (1, 7), (7, 1), (2, 2), (3, 1000), (1000, 3), (2048, 8192), (8192, 8191), (5000, 4999)]
    cases += [(int(a), int(b)) for a, b in rng.integers(1, 3000, (300, 2))]
    for n, m in cases:
        for band in (1, 2, 7, 100, -1, -10, -37, -100, -250):
            p = ts.params("dtw", band=band)
            assert ts.dp_cells(p, n, m) == _cells_by_rows("dtw", band, -1, n, m), (n, m, band)
        for delta in (0, 1, 16, 64, 5000):
            p = ts.params("lcss", delta=delta)
            assert ts.dp_cells(p, n, m) == _cells_by_rows("lcss", 0, delta, n, m), (n, m, delta)
        assert ts.dp_cells(ts.params("dtw"), n, m) == n * m
    assert ts.dp_cells(ts.params("dtw"), 0, 5) == -1
