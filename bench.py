"""Benchmark of the trajectory k-NN hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): 64 query
random walks against a resident database of 100,000 random walks, lengths
uniform in [64, 256], start U[0,1)^2, Gaussian steps sigma=1e-3 (DB seed
0xDB5EED, query seed 0x9E3779B9); DTW top-10 with the bounding-box/endpoint
lower-bound filter (EXACT mode: the true top-10 over the whole DB).

One step = one batched k-NN call (ts_knn) over the 64 queries. The line also
carries Frechet, the plain one-launch-sequence-per-query method, per-query
latency p50/p99, the dominant kernel's roofline (HIP events over the timed
region) and the CPU oracle timed on a bounded sample of the same workload on
this host's cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DTW/Fréchet pairs/s & cell-updates/s, batched vs plain; p50/p99 query latency"
DB_SEED, Q_SEED = 0xDB5EED, 0x9E3779B9
# MI355X published FP32 vector peak: an FMA counts two flops and packed FP32
# two lanes' worth per instruction.
FP32_PEAK_TFLOPS = 157.3
# The ceiling of code built without packed FP32 (disabled by build flag),
# counting one operation per VALU lane-instruction: 256 CUs x 64 lanes x 2.4 GHz.
VALU_ATTAINABLE_TOPS = 39.3
HBM_PEAK_GBS = 8000.0
# Source-level f32 operations per unit of work (DP cell or pair).
OPS = {"dtw_wavefront": 9.0, "frechet_wavefront": 9.0, "lcss_wavefront": 7.0, "dtw_wavefront_seed": 9.0,
       "frechet_wavefront_seed": 9.0, "lcss_wavefront_seed": 7.0, "bound_seed": 24.0, "bound_compact": 24.0}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_workload(ts, n_db, n_q, lo, hi):
    db_len = ts.gen_lengths(DB_SEED, n_db, lo, hi, False)
    db_xy, db_off = ts.gen_walks(DB_SEED, db_len)
    q_len = ts.gen_lengths(Q_SEED, n_q, lo, hi, False)
    q_xy, q_off = ts.gen_walks(Q_SEED, q_len)
    return db_xy, db_off, q_xy, q_off


def timed(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = fn()
    return (time.perf_counter() - t0) * 1e3 / steps, out


def ops_of(name):
    """Source-level ops per work unit of a kernel label ("dtw_wavefront/b3/w32u4"
    is a DP launch of bucket 3 with 32 lanes x 4 cells)."""
    return OPS.get(name.split("/")[0])


def roofline_of(kstats, pmc):
    """Dominant kernel (one launch label: measure, bucket, shape) by total
    time; achieved = algorithmic ops per launch / mean launch duration (HIP
    events on the stream the kernel runs on)."""
    name = max((k for k in kstats if ops_of(k)), key=lambda k: kstats[k]["ms"])
    s = kstats[name]
    per_launch_units = s["units"] / s["launches"]
    mean_ms = s["ms"] / s["launches"]
    achieved = per_launch_units * ops_of(name) / (mean_ms * 1e-3) / 1e12
    traffic = None
    if pmc and name in pmc:
        traffic = pmc[name].get("hbm_bytes_per_launch")
    return {"bound": "valu", "kernel": name, "achieved": round(achieved, 4), "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
            "units_per_launch": per_launch_units, "ms_per_launch": round(mean_ms, 4),
            "ops_per_unit": ops_of(name),
            "attainable": VALU_ATTAINABLE_TOPS, "frac_of_attainable": round(achieved / VALU_ATTAINABLE_TOPS, 4),
            "peak_note": "peak: MI355X's published FP32 vector figure (FMA = 2 flops, packed FP32). The DP "
                         "is bound by VALU issue, neither by HBM nor by MFMA (a min-plus recurrence); built "
                         "without packed FP32 and at one op per lane-instruction its ceiling is 'attainable' "
                         "(256 CUs x 64 lanes x 2.4 GHz)"}


def cpu_baseline(db_xy, db_off, q_xy, q_off, n_sample):
    import oracle
    prm = oracle.params(oracle.DTW, oracle.EXACT, k=10)
    sub_off = q_off[:n_sample + 1]
    sub_xy = q_xy[:sub_off[-1]]
    t0 = time.perf_counter()
    idx, val, refined = oracle.knn(prm, db_xy, db_off, sub_xy, sub_off)
    dt = time.perf_counter() - t0
    n_db = len(db_off) - 1
    return {"value": round(n_sample * n_db / dt, 1), "unit": "pairs/s", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": "%d of the 64 queries against the full %d-trajectory DB, EXACT DTW top-10 (oracle best-first "
                      "search with the same bounds), %.2f s" % (n_sample, n_db, dt)}, idx, val


def copy_calibration(nbytes=256 << 20, reps=3):
    """Plain device-to-device copies of known size through the HIP runtime's
    copy kernel: the reference point that scales rocprof's FETCH_SIZE and
    WRITE_SIZE counters to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log("copy calibration: %d x %d MiB device copies" % (reps, nbytes >> 20))


def c2_other_measures(ts, eng, db, qs, args):
    """The other measures of the 64 x 100k workload, batched EXACT top-10:
    LCSS (eps 2e-3, window 16), Hausdorff, and lock-step Euclidean on a
    separately generated DB of 100,000 walks of 128 points (Euclidean needs
    equal lengths) with 64 queries of 128 points."""
    Q, N = qs.count, db.count
    out = {}
    legs = [("lcss", ts.params("lcss", mode=ts.EXACT, k=10, epsilon=2e-3, delta=16), db, qs),
            ("hausdorff", ts.params("hausdorff", mode=ts.EXACT, k=10), db, qs)]
    e_db_xy, e_db_off = ts.gen_walks(DB_SEED + 1, np.full(N, 128, np.int32))
    e_q_xy, e_q_off = ts.gen_walks(Q_SEED + 1, np.full(Q, 128, np.int32))
    e_db, e_qs = ts.TrajSet(e_db_xy, e_db_off), ts.TrajSet(e_q_xy, e_q_off)
    legs.append(("euclidean", ts.params("euclidean", mode=ts.EXACT, k=10), e_db, e_qs))
    for name, p, d, q in legs:
        ms, (_, _, st) = timed(lambda: eng.knn(d, q, p), max(1, args.steps // 2), 2)
        out[name] = {"pairs_per_s": q.count * d.count / (ms * 1e-3), "ms_per_step": round(ms, 4),
                     "cells_per_step": st["cells"], "cell_updates_per_s": st["cells"] / (ms * 1e-3),
                     "refined_pairs_per_step": st["refined_pairs"]}
        log("c2", name, out[name])
    return out


def cpu_extra(db_xy, db_off, q_xy, q_off, n_one=4):
    """More CPU figures on the same workload: the oracle with one thread
    (the reference's CPU build ran single-threaded) on n_one queries, and the
    REF pipeline (envelope filter, top-10 candidates, exact refine) over all
    queries on every core."""
    import oracle
    n_db = len(db_off) - 1
    cores = oracle.num_threads()
    prm = oracle.params(oracle.DTW, oracle.EXACT, k=10)
    oracle.set_threads(1)
    try:
        t0 = time.perf_counter()
        oracle.knn(prm, db_xy, db_off, q_xy[:q_off[n_one]], q_off[:n_one + 1])
        dt1 = time.perf_counter() - t0
    finally:
        oracle.set_threads(cores)
    ref = oracle.params(oracle.DTW, oracle.REF, k=10)
    t0 = time.perf_counter()
    oracle.knn(ref, db_xy, db_off, q_xy, q_off)
    dtr = time.perf_counter() - t0
    nq = len(q_off) - 1
    return {"exact_1_thread_pairs_per_s": round(n_one * n_db / dt1, 1), "exact_1_thread_sample": n_one,
            "ref_pairs_per_s": round(nq * n_db / dtr, 1), "ref_cores": cores, "ref_seconds": round(dtr, 3)}


def run_c1(ts, args):
    """The reference's CPU-sized case: 1 query against 2,000 DB random walks,
    lengths uniform [32, 128]; DTW REF (r=5, 10 candidates) and EXACT top-10;
    the oracle on the host cores beside it."""
    import oracle
    db_xy, db_off, q_xy, q_off = build_workload(ts, 2000, 1, 32, 128)
    db, qs = ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off)
    eng = ts.Engine(graphs=not args.no_graphs)
    N = db.count
    out = {}
    for mode, name in ((ts.REF, "ref"), (ts.EXACT, "exact")):
        p = ts.params("dtw", mode=mode, k=10)
        ms, (gi, gv, st) = timed(lambda: eng.knn(db, qs, p), max(args.steps, 50), args.warmup)
        prm = oracle.params(oracle.DTW, mode, k=10)
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            oi, ov, _ = oracle.knn(prm, db_xy, db_off, q_xy, q_off)
        cms = (time.perf_counter() - t0) * 1e3 / reps
        out[name] = {"gpu_ms_per_query": round(ms, 4), "gpu_pairs_per_s": N / (ms * 1e-3),
                     "gpu_cell_updates_per_s": st["cells"] / (ms * 1e-3), "cells": st["cells"],
                     "cpu_ms_per_query": round(cms, 4), "cpu_pairs_per_s": N / (cms * 1e-3),
                     "topk_equal_oracle": bool(np.array_equal(gi, oi) and np.array_equal(gv, ov))}
        log("c1", name, out[name])
    line = {"metric": METRIC, "value": round(out["exact"]["gpu_pairs_per_s"], 1), "unit": "pairs/s", "n_gpus": 1,
            "steps": max(args.steps, 50), "warmup": args.warmup, "ms_per_step": out["exact"]["gpu_ms_per_query"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "1 query vs 2,000 DB random walks, lengths U[32,128], DTW top-10 (REF and EXACT)",
                       "queries": 1, "db_trajectories": N, "k": 10},
            "cpu_baseline": {"value": round(out["exact"]["cpu_pairs_per_s"], 1), "unit": "pairs/s",
                             "cores": 1, "kind": "port",
                             "sample": "the whole workload (one query: the oracle's loop over queries runs on one "
                                       "thread), EXACT DTW top-10, mean of 20 runs"},
            "extra": out}
    print(json.dumps(line), flush=True)


def run_c3(ts, args):
    """Uneven-length batching workload: 4,096 queries (log-uniform lengths
    [16, 2048]) against 1,000,000 DB random walks (log-uniform [16, 1024]),
    DTW and Frechet top-10. Batched = all queries in one ts_knn call; plain =
    one launch sequence per query (ts_knn_plain) on a sample of queries, whose
    top-10 must equal the batched ones."""
    t0 = time.time()
    db_len = ts.gen_lengths(DB_SEED, args.n_db, 16, 1024, True)
    db_xy, db_off = ts.gen_walks(DB_SEED, db_len)
    q_len = ts.gen_lengths(Q_SEED, args.n_q, 16, 2048, True)
    q_xy, q_off = ts.gen_walks(Q_SEED, q_len)
    db = ts.TrajSet(db_xy, db_off)
    qs = ts.TrajSet(q_xy, q_off)
    ns = min(args.plain_sample, args.n_q)
    sample = ts.TrajSet(q_xy[:q_off[ns]], q_off[:ns + 1])
    log("c3 workload ready in %.1fs: %d DB trajectories (%d points), %d queries" %
        (time.time() - t0, db.count, len(db_xy), qs.count))
    eng = ts.Engine(graphs=not args.no_graphs)
    Q, N = qs.count, db.count
    out = {}
    for meas in ("dtw", "frechet"):
        p = ts.params(meas, mode=ts.EXACT, k=10)
        eng.knn(db, sample, p)  # warm-up (workspace growth); batches this large launch eagerly
        t1 = time.perf_counter()
        idx, val, st = eng.knn(db, qs, p)
        tb = time.perf_counter() - t1
        t1 = time.perf_counter()
        pi, pv, pst = eng.knn_plain(db, sample, p)
        tp = time.perf_counter() - t1
        same = bool(np.array_equal(pi, idx[:ns]) and np.array_equal(pv, val[:ns]))
        out[meas] = {"batched_pairs_per_s": Q * N / tb, "batched_s": round(tb, 4), "graphs": eng.graph_info(),
                     "batched_cell_updates_per_s": st["cells"] / tb, "cells": st["cells"],
                     "plain_pairs_per_s": ns * N / tp, "plain_s_for_sample": round(tp, 4), "plain_sample": ns,
                     "batched_vs_plain": round((Q * N / tb) / (ns * N / tp), 2), "plain_topk_equals_batched": same,
                     "overflow_retries": st["overflow_retries"]}
        log(meas, out[meas])
    out["db_file"] = db_file_leg(ts, eng, db, sample)
    cpu = None
    if not args.no_cpu:
        import oracle
        nc = min(16, Q)
        prm = oracle.params(oracle.DTW, oracle.EXACT, k=10)
        t1 = time.perf_counter()
        oracle.knn(prm, db_xy, db_off, q_xy[:q_off[nc]], q_off[:nc + 1])
        dt = time.perf_counter() - t1
        cpu = {"value": round(nc * N / dt, 1), "unit": "pairs/s", "cores": min(nc, oracle.num_threads()),
               "kind": "port", "sample": "%d of the %d queries against the full DB, EXACT DTW top-10 (oracle "
                                         "best-first search, one thread per query), %.2f s" % (nc, Q, dt)}
        log("c3 cpu", cpu)
    line = {"metric": METRIC, "value": round(out["dtw"]["batched_pairs_per_s"], 1), "unit": "pairs/s", "n_gpus": 1,
            "steps": 1, "warmup": 1, "ms_per_step": round(out["dtw"]["batched_s"] * 1e3, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "uneven-length batching: %d queries log-uniform [16,2048] vs %d DB random walks "
                                   "log-uniform [16,1024], DTW/Frechet top-10, EXACT bounding-box filter" % (Q, N),
                       "queries": Q, "db_trajectories": N, "k": 10},
            "cpu_baseline": cpu, "extra": out}
    print(json.dumps(line), flush=True)


def db_file_leg(ts, eng, db, sample):
    """The DB as a tsdb file: save from HBM, load back (the file then sits in
    the host's page cache: a warm restart), and the loaded set's DTW top-10
    for the query sample equals the original's."""
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "db.tsdb")
        t1 = time.perf_counter()
        db.save(path)
        t_save = time.perf_counter() - t1
        nbytes = os.path.getsize(path)
        t1 = time.perf_counter()
        db2 = ts.TrajSet.load(path)
        t_load = time.perf_counter() - t1
    p = ts.params("dtw", mode=ts.EXACT, k=10)
    a, b = eng.knn(db, sample, p), eng.knn(db2, sample, p)
    same = bool(np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.int32), b[1].view(np.int32)))
    db2.close()
    res = {"bytes": nbytes, "save_s": round(t_save, 3), "load_s": round(t_load, 3),
           "load_GB_per_s": round(nbytes / t_load / 1e9, 2), "loaded_topk_equals": same,
           "note": "load = map, check, copy points to HBM, recompute and compare summaries; file in the page cache"}
    log("db file", res)
    return res


def run_c4(ts, args):
    """Long trajectories: 256 queries against 20,000 DB random walks, lengths
    uniform in [2048, 8192]; DTW with a Sakoe-Chiba band of
    w = ceil(0.1 * max(n, m)) per pair, and LCSS with eps = 2e-3 and an index
    window delta = 64; EXACT top-10. The figure is in-band DP cell updates per
    second (cells inside the band or window of every refined pair / time)."""
    t0 = time.time()
    db_len = ts.gen_lengths(DB_SEED, args.n_db, 2048, 8192, False)
    db_xy, db_off = ts.gen_walks(DB_SEED, db_len)
    q_len = ts.gen_lengths(Q_SEED, args.n_q, 2048, 8192, False)
    q_xy, q_off = ts.gen_walks(Q_SEED, q_len)
    db = ts.TrajSet(db_xy, db_off)
    qs = ts.TrajSet(q_xy, q_off)
    ns = min(args.plain_sample, args.n_q)
    sample = ts.TrajSet(q_xy[:q_off[ns]], q_off[:ns + 1])
    log("c4 workload ready in %.1fs: %d DB trajectories (%d points), %d queries" %
        (time.time() - t0, db.count, len(db_xy), qs.count))
    eng = ts.Engine()
    Q, N = qs.count, db.count
    out = {}
    legs = (("dtw", ts.params("dtw", mode=ts.EXACT, k=10, band=-10)),
            ("lcss", ts.params("lcss", mode=ts.EXACT, k=10, epsilon=2e-3, delta=64)))
    for meas, p in legs:
        eng.knn(db, sample, p)  # warm-up (workspace growth)
        eng.set_timing(True)
        eng.reset_stats()
        t1 = time.perf_counter()
        idx, val, st = eng.knn(db, qs, p)
        tb = time.perf_counter() - t1
        kst = eng.kernel_stats()
        eng.set_timing(False)
        wave = {k: v for k, v in kst.items() if "wavefront" in k}
        dp_ms = sum(v["ms"] for v in wave.values())
        t1 = time.perf_counter()
        pi, pv, pst = eng.knn_plain(db, sample, p)
        tp = time.perf_counter() - t1
        same = bool(np.array_equal(pi, idx[:ns]) and np.array_equal(pv, val[:ns]))
        r = {"batched_s": round(tb, 4), "pairs_per_s": Q * N / tb, "cells": st["cells"],
             "cell_updates_per_s": st["cells"] / tb, "dp_kernel_ms": round(dp_ms, 3),
             "dp_kernel_cell_updates_per_s": st["cells"] / (dp_ms * 1e-3) if dp_ms > 0 else None,
             "refined_pairs_per_query": st["refined_pairs"] / Q,
             "kernels": {k: {"launches": v["launches"], "ms": round(v["ms"], 3)} for k, v in kst.items()},
             "plain_s_for_sample": round(tp, 4), "plain_sample": ns, "plain_topk_equals_batched": same}
        if args.oracle_queries > 0:
            import oracle
            nq = min(args.oracle_queries, Q)
            prm = oracle.params(p.measure, oracle.EXACT, k=10, band=p.band, delta=p.delta, epsilon=p.epsilon)
            t1 = time.perf_counter()
            oi, ov, _ = oracle.knn(prm, db_xy, db_off, q_xy[:q_off[nq]], q_off[:nq + 1])
            r["oracle_check"] = {"queries": nq, "seconds": round(time.perf_counter() - t1, 2),
                                 "topk_equal": bool(np.array_equal(oi, idx[:nq]) and np.array_equal(ov, val[:nq]))}
        out[meas] = r
        log("c4", meas, {k: v for k, v in r.items() if k != "kernels"})
    cpu = None
    if not args.no_cpu:
        import oracle
        rng = np.random.default_rng(0xC4)
        npairs = 192
        qi = rng.integers(0, Q, npairs).astype(np.int32)
        di = rng.integers(0, N, npairs).astype(np.int32)
        cpu = {}
        for meas, p in legs:
            prm = oracle.params(p.measure, band=p.band, delta=p.delta, epsilon=p.epsilon)
            t1 = time.perf_counter()
            oracle.pair_values(prm, db_xy, db_off, q_xy, q_off, qi, di)
            dt = time.perf_counter() - t1
            cells = band_cells(p, q_len[qi], db_len[di])
            cpu[meas] = {"value": round(cells / dt, 1), "unit": "cell-updates/s", "seconds": round(dt, 3)}
        cpu = {"value": cpu["dtw"]["value"], "unit": "cell-updates/s", "cores": oracle.num_threads(),
               "kind": "port", "sample": "%d random (query, DB) pairs of the workload, exact values on every core "
                                         "(the banded DTW visits only in-band cells; the LCSS oracle fills the "
                                         "whole table)" % npairs, "lcss": cpu["lcss"]}
        log("c4 cpu", cpu)
    line = {"metric": METRIC, "value": round(out["dtw"]["cell_updates_per_s"], 1), "unit": "cell-updates/s",
            "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": round(out["dtw"]["batched_s"] * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "long trajectories: %d queries vs %d DB random walks, lengths U[2048,8192], "
                                   "DTW band ceil(0.1 max(n,m)) and LCSS eps 2e-3 delta 64, EXACT top-10" % (Q, N),
                       "queries": Q, "db_trajectories": N, "k": 10},
            "cpu_baseline": cpu, "extra": out}
    print(json.dumps(line), flush=True)


def band_cells(p, n, m):
    """In-band DP cells of pairs of lengths n, m (arrays) for a DTW band or an
    LCSS window, as the engine counts them."""
    tot = 0
    for a, b in zip(np.asarray(n, np.int64), np.asarray(m, np.int64)):
        R, C = max(a, b), min(a, b)
        r = np.arange(R, dtype=np.int64)
        if p.measure == 0 and p.band != 0:
            w = p.band if p.band > 0 else (-p.band * R + 99) // 100
            bw = w * R
            x = r * C - bw
            lo = np.where(x <= 0, 0, (x + R - 1) // R)
            hi = np.minimum(C - 1, (r * C + bw) // R)
            tot += int(np.maximum(hi - lo + 1, 0).sum())
        elif p.measure == 1 and p.delta >= 0:
            r = r[r - p.delta <= C - 1]
            tot += int((np.minimum(C - 1, r + p.delta) - np.maximum(0, r - p.delta) + 1).sum())
        else:
            tot += int(R * C)
    return tot


def run_c5(ts, args):
    """Query serving: a 2,000,000-trajectory DB (log-uniform lengths [16, 512])
    resident in HBM; Poisson arrivals at each stated QPS; query lengths
    log-uniform [16, 1024]; measure mix DTW 40% / LCSS 20% / Frechet 20% /
    Hausdorff 20%; k uniform in {1, 10, 100}; EXACT top-k. Latency is the
    server's own clock from submit to result ready."""
    import threading
    t0 = time.time()
    db_len = ts.gen_lengths(DB_SEED, args.n_db, 16, 512, True)
    db_xy, db_off = ts.gen_walks(DB_SEED, db_len)
    db = ts.TrajSet(db_xy, db_off)
    log("c5 DB ready in %.1fs: %d trajectories, %.2f GB resident" % (time.time() - t0, db.count, db.nbytes() / 1e9))
    rng = np.random.default_rng(0xA11CE)
    nreq_max = int(max(args.qps) * args.duration) + 16
    q_len = ts.gen_lengths(Q_SEED, nreq_max, 16, 1024, True)
    q_xy, q_off = ts.gen_walks(Q_SEED, q_len)
    mix = rng.choice(4, size=nreq_max, p=[0.4, 0.2, 0.2, 0.2])
    ks = rng.choice([1, 10, 100], size=nreq_max)
    names = ["dtw", "lcss", "frechet", "hausdorff"]
    prm = {(m, k): ts.params(names[m], mode=ts.EXACT, k=int(k), epsilon=2e-3) for m in range(4) for k in (1, 10, 100)}
    results = []
    for qps in args.qps:
        srv = ts.Server(db, streams=4, max_batch=256, batch_wait_us=500)
        n = int(qps * args.duration)
        arrivals = np.cumsum(rng.exponential(1.0 / qps, size=n))
        tickets = []
        refused = 0
        lat = []
        lock = threading.Lock()
        done = threading.Event()

        def waiter():
            while True:
                with lock:
                    if not tickets:
                        if done.is_set():
                            return
                        item = None
                    else:
                        item = tickets.pop(0)
                if item is None:
                    time.sleep(0.0005)
                    continue
                t, k = item
                _, _, ms = srv.wait(t, k, timeout_ms=120000)
                with lock:
                    lat.append(ms)

        ws = [threading.Thread(target=waiter) for _ in range(4)]
        for w in ws:
            w.start()
        start = time.perf_counter()
        for i in range(n):
            while time.perf_counter() - start < arrivals[i]:
                pass
            p = prm[(int(mix[i]), int(ks[i]))]
            try:
                t = srv.submit(q_xy[q_off[i]:q_off[i + 1]], p)
                with lock:
                    tickets.append((t, p.k))
            except ts.TrajsimError as e:
                if e.status != 6:
                    raise
                refused += 1
        submit_s = time.perf_counter() - start
        done.set()
        for w in ws:
            w.join()
        wall = time.perf_counter() - start
        c = srv.counters()
        srv.close()
        lat = np.array(lat)
        r = {"offered_qps": qps, "requests": n, "completed": int(len(lat)), "refused": refused,
             "achieved_qps": round(len(lat) / wall, 1), "submit_s": round(submit_s, 3), "wall_s": round(wall, 3),
             "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
             "mean_ms": round(float(lat.mean()), 3), "batches": c["batches"],
             "mean_batch": round(len(lat) / max(1, c["batches"]), 2)}
        log("c5", r)
        results.append(r)
    cpu = None
    if not args.no_cpu:
        # CPU QPS: the oracle answers the first 200 requests of the stream,
        # grouped by (measure, k), each group parallel over its queries.
        import oracle
        ncpu = min(200, nreq_max)
        t1 = time.perf_counter()
        for (m, k), _ in prm.items():
            sel = [i for i in range(ncpu) if mix[i] == m and ks[i] == k]
            if not sel:
                continue
            sub = [q_xy[q_off[i]:q_off[i + 1]] for i in sel]
            so = ts.offsets_of([len(t) for t in sub])
            meas = {0: oracle.DTW, 1: oracle.LCSS, 2: oracle.FRECHET, 3: oracle.HAUSDORFF}[m]
            op = oracle.params(meas, oracle.EXACT, k=int(k), epsilon=2e-3)
            oracle.knn(op, db_xy, db_off, np.concatenate(sub), so)
        dt = time.perf_counter() - t1
        cpu = {"value": round(ncpu / dt, 2), "unit": "queries/s", "cores": oracle.num_threads(), "kind": "port",
               "sample": "the first %d requests of the stream (same mix and k), oracle EXACT k-NN grouped by "
                         "(measure, k), %.2f s" % (ncpu, dt)}
        log("c5 cpu", cpu)
    best = results[0]
    line = {"metric": METRIC, "value": best["p99_ms"], "unit": "ms (p99 latency at %d QPS)" % best["offered_qps"],
            "n_gpus": 1, "steps": len(results), "warmup": 0, "ms_per_step": None, "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "serving: %d-trajectory DB (log-uniform [16,512]) resident, Poisson arrivals, "
                                   "query lengths log-uniform [16,1024], mix DTW 40/LCSS 20/Frechet 20/Hausdorff 20, "
                                   "k in {1,10,100}, EXACT top-k, continuous batching over 4 streams" % db.count,
                       "db_trajectories": db.count, "duration_s": args.duration},
            "cpu_baseline": cpu, "extra": {"sweep": results}}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--qps", type=lambda v: [int(x) for x in v.split(",")], default=[2000, 5000, 10000, 20000])
    ap.add_argument("--duration", type=float, default=2.0)
    ap.add_argument("--n-db", type=int, default=None)
    ap.add_argument("--n-q", type=int, default=None)
    ap.add_argument("--plain-sample", type=int, default=256)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="c2: launch eagerly instead of replaying HIP graphs")
    ap.add_argument("--oracle-queries", type=int, default=0,
                    help="c4: also check the top-10 of this many queries against the CPU oracle (about a minute each)")
    ap.add_argument("--quick", action="store_true", help="skip plain/latency/frechet legs (profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (HBM byte-counter calibration for rocprof --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one GPU per box); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import cudaruntime1_amd as ts
    ts.init(0)
    if args.config == "c1":
        return run_c1(ts, args)
    if args.config == "c3":
        args.n_db = args.n_db or 1_000_000
        args.n_q = args.n_q or 4096
        return run_c3(ts, args)
    if args.config == "c4":
        args.n_db = args.n_db or 20_000
        args.n_q = args.n_q or 256
        args.plain_sample = min(args.plain_sample, 16)
        return run_c4(ts, args)
    if args.config == "c5":
        args.n_db = args.n_db or 2_000_000
        return run_c5(ts, args)
    args.n_db = args.n_db or 100_000
    args.n_q = args.n_q or 64
    if args.copy_calibration:
        copy_calibration()
    dev = ts.device_name()
    log("device", dev)
    t0 = time.time()
    db_xy, db_off, q_xy, q_off = build_workload(ts, args.n_db, args.n_q, 64, 256)
    db = ts.TrajSet(db_xy, db_off)
    qs = ts.TrajSet(q_xy, q_off)
    log("workload ready in %.1fs: %d DB trajectories (%d points), %d queries" %
        (time.time() - t0, db.count, len(db_xy), qs.count))
    Q, N = qs.count, db.count
    eng = ts.Engine(graphs=not args.no_graphs)
    p_dtw = ts.params("dtw", mode=ts.EXACT, k=10)
    p_fr = ts.params("frechet", mode=ts.EXACT, k=10)

    # Timed region: K batched DTW steps, replayed from a HIP graph (the
    # warm-up records it; --no-graphs launches eagerly).
    for _ in range(max(args.warmup, 2)):
        eng.knn(db, qs, p_dtw)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        idx, val, st = eng.knn(db, qs, p_dtw)
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    value = Q * N / (ms_step * 1e-3)
    cells = st["cells"]
    ginfo = eng.graph_info()
    log("DTW batched: %.3f ms/step, %.3e pairs/s, %.3e cells/step, graphs %s" % (ms_step, value, cells, ginfo))
    # Per-kernel durations (HIP events around every launch, so eager): the
    # same K steps again, for the roofline and the kernel table.
    eng.set_timing(True)
    eng.reset_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ti, tv, _ = eng.knn(db, qs, p_dtw)
    ms_eager = (time.perf_counter() - t0) * 1e3 / args.steps
    kstats = eng.kernel_stats()
    eng.set_timing(False)
    assert np.array_equal(ti, idx) and np.array_equal(tv, val)
    log("DTW batched eager with kernel events: %.3f ms/step" % ms_eager)

    pmc = None
    pmc_path = os.path.join(ROOT, "profiles", "pmc_traffic_c2.json")
    if os.path.exists(pmc_path):
        pmc = json.load(open(pmc_path))
    roof = roofline_of(kstats, pmc)
    extra = {"device": dev, "dtw_ms_per_step": round(ms_step, 4), "dtw_cells_per_step": cells,
             "graphs": ginfo, "dtw_ms_per_step_eager_timed": round(ms_eager, 4),
             "dtw_cell_updates_per_s": cells / (ms_step * 1e-3),
             "dtw_device_ms": round(st["ms_device"], 4), "launches_per_step": st["launches"],
             "kernels": {k: {"launches": v["launches"], "ms_mean": round(v["ms"] / max(v["launches"], 1), 5)}
                         for k, v in kstats.items()}}

    if not args.quick:
        ms_fr, (fi, fv, fst) = timed(lambda: eng.knn(db, qs, p_fr), args.steps, 2)
        extra.update({"frechet_pairs_per_s": Q * N / (ms_fr * 1e-3), "frechet_ms_per_step": round(ms_fr, 4),
                      "frechet_cell_updates_per_s": fst["cells"] / (ms_fr * 1e-3)})
        ms_plain, (pi, pv, _) = timed(lambda: eng.knn_plain(db, qs, p_dtw), max(1, args.steps // 3), 1)
        same = bool(np.array_equal(pi, idx) and np.array_equal(pv, val))
        extra.update({"dtw_plain_pairs_per_s": Q * N / (ms_plain * 1e-3), "dtw_plain_ms_per_step": round(ms_plain, 4),
                      "batched_vs_plain": round(ms_plain / ms_step, 3), "plain_equals_batched": same})
        # Per-query latency of a single-query request on the resident DB.
        lat = []
        singles = [ts.TrajSet(q_xy[q_off[i]:q_off[i + 1]], np.array([0, q_off[i + 1] - q_off[i]], np.int64))
                   for i in range(Q)]
        for s in singles[:4]:
            eng.knn(db, s, p_dtw)
        for s in singles:
            t1 = time.perf_counter()
            eng.knn(db, s, p_dtw)
            lat.append((time.perf_counter() - t1) * 1e3)
        extra.update({"query_latency_ms_p50": round(float(np.percentile(lat, 50)), 4),
                      "query_latency_ms_p99": round(float(np.percentile(lat, 99)), 4),
                      "latency_note": "single-query DTW top-10 calls, serial, DB resident"})
        log("frechet %.3f ms, plain %.3f ms, p50 %.3f ms" % (ms_fr, ms_plain, extra["query_latency_ms_p50"]))
        extra["measures"] = c2_other_measures(ts, eng, db, qs, args)

    cpu = None
    if not args.no_cpu:
        cpu, ci, cv = cpu_baseline(db_xy, db_off, q_xy, q_off, args.cpu_sample)
        extra["cpu_matches_gpu"] = bool(np.array_equal(ci, idx[:args.cpu_sample]) and
                                        np.array_equal(cv, val[:args.cpu_sample]))
        log("cpu baseline", cpu)
        extra["cpu_other"] = cpu_extra(db_xy, db_off, q_xy, q_off)
        log("cpu other", extra["cpu_other"])

    line = {"metric": METRIC, "value": round(value, 1), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "DTW top-10 k-NN, %d query random walks vs %d-trajectory DB (lengths U[64,256], "
                                   "sigma 1e-3), bounding-box/endpoint lower-bound filter, batched" % (Q, N),
                       "queries": Q, "db_trajectories": N, "k": 10, "measure": "dtw", "mode": "exact"},
            "roofline": roof, "cpu_baseline": cpu, "extra": extra}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
