"""Autotuner for the wavefront DP: per DP variant and length bucket, the
lane-group shape (lanes per pair, cells per lane step) and the persistent
grid (waves per CU) with the highest measured cell rate, and for refines of
few pairs (at most one per CU: 128 pairs), small refines (one round of
pairs, up to 4096) and mid-size refines (4096 pairs) the ones with the
lowest time.

    python -m cudaruntime1_amd.tune [--out PATH] [--quick] [--variants dtw,frechet] [--few-only]

For every (variant, bucket) it builds a set of random-walk pairs whose DP
row width lies in the bucket, runs the DP over them once per candidate
(shape, waves per CU) with HIP events around the launch, and keeps the
fastest. Every candidate's values must equal the default shape's bit for bit
(the shape changes the schedule, never the arithmetic); a mismatch aborts.
The table is written as text (default: dp_tuning.txt beside the library,
which engines read when they are created).
"""
import argparse
import os
import sys
import time

import numpy as np

import cudaruntime1_amd as ts

# Row widths per bucket the tuning pairs are drawn from, and the number of
# DP cells to time per candidate (about a millisecond or more of DP).
BUCKET_WIDTHS = [(16, 31), (32, 63), (64, 127), (128, 255), (256, 1023), (1024, 2048)]
TARGET_CELLS = 6e8
# at least four rounds of the widest grid (256 CUs x 16 waves, one pair per
# 64-lane group), so tails do not decide
MIN_PAIRS = 16384
WAVES = [4, 6, 8, 12, 16]
# pairs of the small-refine measurement (a seed round; at most kLatencyPairs)
NLAT = 1024
# pairs of the mid-size measurement (a few rounds of the grid: the main refine
# of a batch of tens of queries)
NMID = 4096
# pairs of the few-pair measurement (at most one per CU: the seeds of one
# query, a pair-value call on a handful of long pairs)
NFEW = 128


def variant_params(v, width):
    """Parameters of variant v for pairs of row width about `width`."""
    if v == "dtw":
        return ts.params("dtw", mode=ts.EXACT)
    if v == "frechet":
        return ts.params("frechet", mode=ts.EXACT)
    if v == "lcss":
        return ts.params("lcss", mode=ts.EXACT, epsilon=2e-3)
    if v == "dtw_band":
        return ts.params("dtw", mode=ts.EXACT, band=-10)
    return ts.params("lcss", mode=ts.EXACT, epsilon=2e-3, delta=max(0, (width - 1) // 2))


def pair_lengths(v, lo, hi, rng, n):
    """(n, m) per pair so that the visited row width is in [lo, hi]."""
    if v == "dtw_band":
        # width 2*ceil(R/10)+1 in [lo, hi] <=> R in [10*ceil((lo-1)/2)-9, 10*floor((hi-1)/2)];
        # C in [R/2, R] stays above the width
        r_lo = max(10 * ((lo - 1 + 1) // 2) - 9, 4)
        r_hi = 10 * ((hi - 1) // 2)
        R = rng.integers(r_lo, r_hi + 1, n)
        C = np.maximum((R * rng.uniform(0.5, 1.0, n)).astype(np.int64), 2)
    elif v == "lcss_window":
        # width 2*delta+1 (delta fixed per bucket): C, R well above it
        C = rng.integers(4 * hi, 8 * hi + 1, n)
        R = (C * rng.uniform(1.0, 1.5, n)).astype(np.int64)
    else:
        # widths log-uniform in the bucket and the longer side 1-4 times the
        # shorter, log-uniform: the shape of pairs drawn from log-uniform
        # length distributions, where short widths dominate a bucket
        C = np.clip(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), n)).astype(np.int64), lo, hi)
        R = (C * np.exp(rng.uniform(0.0, np.log(4.0), n))).astype(np.int64)
    swap = rng.random(n) < 0.5
    a, b = np.where(swap, C, R), np.where(swap, R, C)
    return a.astype(np.int32), b.astype(np.int32)


def tune(quick=False, log=print, only=None, few_only=False):
    ts.init(0)
    shapes = ts.dp_shapes()
    rng = np.random.default_rng(0x7E5)
    eng = ts.Engine(timing=True)
    lines = []
    results = {}
    variants = ts.DP_VARIANTS
    waves = [8] if quick else WAVES
    for vi, v in enumerate(variants):
        if only and v not in only:
            continue
        for b, (lo, hi) in enumerate(BUCKET_WIDTHS):
            width = (lo + hi) // 2
            p = variant_params(v, width)
            # one pair set per bucket
            probe_a, probe_b = pair_lengths(v, lo, hi, rng, 64)
            mean_cells = float(np.mean(probe_a.astype(np.float64) * probe_b))
            if v == "dtw_band":
                mean_cells *= 0.2
            elif v == "lcss_window":
                mean_cells = float(np.mean(np.maximum(probe_a, probe_b))) * (2 * p.delta + 1)
            npairs = int(np.clip(TARGET_CELLS / max(mean_cells, 1.0), MIN_PAIRS, 65536))
            na, nb = pair_lengths(v, lo, hi, rng, npairs)
            q_xy, q_off = ts.gen_walks(0x7E50 + 2 * b, na, 2e-3)
            d_xy, d_off = ts.gen_walks(0x7E51 + 2 * b, nb, 2e-3)
            qs, db = ts.TrajSet(q_xy, q_off), ts.TrajSet(d_xy, d_off)
            idx = np.arange(npairs, dtype=np.int32)
            # every pair in this set must fall in bucket b
            widths = np.array([dp_width(v, p, int(x), int(y)) for x, y in zip(na[:256], nb[:256])])
            assert widths.min() > (BUCKET_WIDTHS[b - 1][1] if b else 0) and widths.max() <= ts.DP_BUCKET_CAPS[b], \
                (v, b, widths.min(), widths.max())
            default = eng.get_dp_shape(vi, b, 0)
            default_lat = eng.get_dp_shape(vi, b, 1)
            default_mid = eng.get_dp_shape(vi, b, 3)
            default_few = eng.get_dp_shape(vi, b, 4)
            ref = eng.pair_values(db, qs, p, idx, idx)
            lat_idx = idx[:NLAT]
            mid_idx = idx[:NMID]
            few_idx = idx[:NFEW]
            best, best_lat, best_mid, best_few = None, None, None, None

            def dp_time(sel):
                eng.reset_stats()
                got = eng.pair_values(db, qs, p, sel, sel)
                ks = eng.kernel_stats()
                ms = sum(s["ms"] for k, s in ks.items() if "wavefront" in k)
                cells = sum(s["units"] for k, s in ks.items() if "wavefront" in k)
                if not np.array_equal(got.view(np.int32), ref[:len(sel)].view(np.int32)):
                    raise AssertionError("shape (%d, %d) changes %s values in bucket %d" % (W, U, v, b))
                return ms, cells

            for (W, U) in shapes:
                # no more than half of each row's chunks idle
                if (hi + U - 1) // U < W // 2:
                    continue
                for g in waves:
                    eng.set_dp_shape(vi, b, W, U, g)  # every table
                    f0 = eng.dp_fallbacks()
                    eng.pair_values(db, qs, p, idx[:256], idx[:256])  # warm-up
                    if eng.dp_fallbacks() != f0:
                        break  # this shape's halo does not fit the LDS for these pairs
                    fms, _ = dp_time(few_idx)  # a few pairs: the longest one's latency decides
                    if best_few is None or fms < best_few[0]:
                        best_few = (fms, W, U, g)
                    if few_only:
                        continue
                    ms, cells = dp_time(idx)
                    rate = cells / (ms * 1e-3)
                    if best is None or rate > best[0]:
                        best = (rate, W, U, g, ms)
                    lms, _ = dp_time(lat_idx)  # one small refine: its latency decides
                    if best_lat is None or lms < best_lat[0]:
                        best_lat = (lms, W, U, g)
                    mms, _ = dp_time(mid_idx)
                    if best_mid is None or mms < best_mid[0]:
                        best_mid = (mms, W, U, g)
            eng.set_dp_shape(vi, b, *default, table=0)
            eng.set_dp_shape(vi, b, *default_lat, table=1)
            eng.set_dp_shape(vi, b, *default_mid, table=3)
            eng.set_dp_shape(vi, b, *default_few, table=4)
            fms, W, U, g = best_few
            log("%-12s bucket %d: few pairs (%d): lanes %3d cells %d waves/CU %2d  %.3f ms" %
                (v, b, NFEW, W, U, g, fms))
            lines.append("%s %d %d %d %d few  # %d pairs in %.3f ms" % (v, b, W, U, g, NFEW, fms))
            if few_only:
                continue
            rate, W, U, g, ms = best
            results[(v, b)] = best
            log("%-12s bucket %d: lanes %2d cells %d waves/CU %2d  %.1f G cells/s (%d pairs, %.3f ms)%s" %
                (v, b, W, U, g, rate / 1e9, npairs, ms, "" if (W, U, g) != default else "  [default]"))
            lines.append("%s %d %d %d %d  # %.1f G cells/s" % (v, b, W, U, g, rate / 1e9))
            lms, W, U, g = best_lat
            log("%-12s bucket %d: small refines (%d pairs): lanes %2d cells %d waves/CU %2d  %.3f ms" %
                (v, b, NLAT, W, U, g, lms))
            lines.append("%s %d %d %d %d lat  # %d pairs in %.3f ms" % (v, b, W, U, g, NLAT, lms))
            mms, W, U, g = best_mid
            log("%-12s bucket %d: mid-size refines (%d pairs): lanes %2d cells %d waves/CU %2d  %.3f ms" %
                (v, b, NMID, W, U, g, mms))
            lines.append("%s %d %d %d %d mid  # %d pairs in %.3f ms" % (v, b, W, U, g, NMID, mms))
    return lines, results


def dp_width(v, p, n, m):
    c, r = min(n, m), max(n, m)
    if v == "dtw_band":
        w = p.band if p.band > 0 else (-p.band * r + 99) // 100
        return min(c, 2 * w + 1)
    if v == "lcss_window":
        return min(c, 2 * p.delta + 1)
    return c


def line_key(line):
    """(variant, bucket, table) of a table line, None for comments."""
    body = line.split("#")[0].split()
    if len(body) < 5:
        return None
    return body[0], int(body[1]), body[5] if len(body) > 5 else ""


def merge(old_lines, new_lines):
    """The old table with the new lines in place of the lines of the same
    (variant, bucket, table); an old line whose comment says it was set by
    hand stays."""
    new = {line_key(l): l for l in new_lines}
    out, seen = [], set()
    for l in old_lines:
        k = line_key(l)
        if k in new and "set by hand" not in l:
            out.append(new[k])
            seen.add(k)
        else:
            out.append(l)
            seen.add(k)
    out += [l for k, l in new.items() if k not in seen]
    return out


def main():
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dp_tuning.txt")
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=here)
    ap.add_argument("--quick", action="store_true", help="waves per CU fixed at 8")
    ap.add_argument("--variants", default="", help="comma-separated variants to tune; the table beside the library "
                                                   "keeps its other lines")
    ap.add_argument("--few-only", action="store_true", help="tune only the few-pair table; the table beside the "
                                                           "library keeps its other lines")
    args = ap.parse_args()
    only = [v for v in args.variants.split(",") if v]
    t0 = time.time()
    lines, _ = tune(args.quick, log=lambda *a: print(*a, file=sys.stderr, flush=True), only=only,
                    few_only=args.few_only)
    what = ", variants " + ",".join(only) if only else ""
    if args.few_only:
        what += ", few-pair lines"
    head = ["# DP tuning table written by cudaruntime1_amd/tune.py on %s (%.0f s%s)" %
            (ts.device_name().strip(), time.time() - t0, what),
            "# variant bucket lanes cells waves_per_cu"]
    if (only or args.few_only) and os.path.exists(here):
        old = [l.rstrip("\n") for l in open(here) if l.strip()]
        lines = merge([l for l in old if not l.startswith("# variant")], lines)
        head = head[:1] + ["# variant bucket lanes cells waves_per_cu"]
    with open(args.out, "w") as f:
        f.write("\n".join(head + lines) + "\n")
    print("wrote", args.out)


if __name__ == "__main__":
    main()
