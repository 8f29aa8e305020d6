"""Benchmark of the trajectory k-NN hot path on one MI355X.

Default workload (the configuration the headline metric, batched vs plain
pairs/s, is stated on): 4,096 query random walks with lengths log-uniform in
[16, 2048] against a resident database of 1,000,000 random walks with
lengths log-uniform in [16, 1024]; start U[0,1)^2, Gaussian steps
sigma=1e-3 (DB seed 0xDB5EED, query seed 0x9E3779B9); DTW and discrete
Frechet top-10 over the whole database (EXACT mode: bounding-box/endpoint
lower bounds, tight point-to-box bounds, persistent length-bucketed DP).

One step = one batched k-NN call (ts_knn) over all 4,096 queries, sets
already resident in HBM. The line carries:
  value          DTW pairs/s of the batched call (4096 * 1e6 / seconds);
  extra          Frechet, the plain method (one launch sequence per query) on
                 a query sample, batched/plain ratios, cell-updates/s, the
                 64 x 100k five-measure figures, the database-file leg;
  parity         the batched rows of the longest, shortest and middle-length
                 queries against the CPU oracle, bit for bit;
  roofline       the dominant kernel (by time in the timed region) from HIP
                 events on the stream it runs on;
  cpu_baseline   the reference's pipeline (envelope filter, 10 candidates,
                 exact refine), as restated by the oracle for variable
                 lengths, on a bounded query sample over the full database
                 on this host's cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1..c5]
"""
import argparse
import ctypes
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
# counting one operation per VALU lane-instruction: 256 CUs x 128 lanes x
# 2.4 GHz (the issue rate behind the published figure, whose FMA counts two
# flops; SQ counters show more than one wave64 VALU instruction per CU-cycle).
VALU_ATTAINABLE_TOPS = 78.6
HBM_PEAK_GBS = 8000.0
# Source-level f32 operations per unit of work (DP cell).
OPS = {"dtw_wavefront": 9.0, "frechet_wavefront": 9.0, "lcss_wavefront": 7.0, "dtw_wavefront_seed": 9.0,
       "frechet_wavefront_seed": 9.0, "lcss_wavefront_seed": 7.0}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


_hip = None


def device_sync():
    """hipDeviceSynchronize: the timed region starts and ends with the device idle."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
    assert _hip.hipDeviceSynchronize() == 0


def build_workload(ts, n_db, n_q, lo, hi):
    db_len = ts.gen_lengths(DB_SEED, n_db, lo, hi, False)
    db_xy, db_off = ts.gen_walks(DB_SEED, db_len)
    q_len = ts.gen_lengths(Q_SEED, n_q, lo, hi, False)
    q_xy, q_off = ts.gen_walks(Q_SEED, q_len)
    return db_xy, db_off, q_xy, q_off


def timed(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    device_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = fn()
    device_sync()
    return (time.perf_counter() - t0) * 1e3 / steps, out


def ops_of(name):
    """Source-level ops per work unit of a kernel label ("dtw_wavefront/b3/w32u4"
    is a DP launch of bucket 3 with 32 lanes x 4 cells)."""
    return OPS.get(name.split("/")[0])


def load_pmc(name):
    path = os.path.join(ROOT, "profiles", name)
    return json.load(open(path)) if os.path.exists(path) else None


def roofline_of(kstats, pmc, pmc_source):
    """Dominant DP launch (one label: measure, phase, bucket, shape) by total
    time; achieved = algorithmic ops per launch / mean launch duration, both
    from the engine's HIP events on the stream the kernel runs on."""
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
            "traffic_source": pmc_source if traffic is not None else None,
            "units_per_launch": per_launch_units, "unit_name": "DP cell", "ms_per_launch": round(mean_ms, 4),
            "launches": s["launches"], "ops_per_unit": ops_of(name),
            "attainable": VALU_ATTAINABLE_TOPS, "frac_of_attainable": round(achieved / VALU_ATTAINABLE_TOPS, 4),
            "cell_updates_per_s": per_launch_units / (mean_ms * 1e-3),
            "peak_note": "peak: MI355X's published FP32 vector figure (FMA = 2 flops, packed FP32). The DP "
                         "is bound by VALU issue, neither by HBM nor by MFMA (a min-plus recurrence); built "
                         "without packed FP32 and at one op per lane-instruction its ceiling is 'attainable' "
                         "(256 CUs x 128 lanes x 2.4 GHz)"}


def copy_calibration(nbytes=256 << 20, reps=3):
    """Plain device-to-device copies of known size through the HIP runtime's
    copy kernel: the reference point that scales rocprof's FETCH_SIZE and
    WRITE_SIZE counters to bytes."""
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


def sub_queries(q_xy, q_off, sel):
    parts = [q_xy[q_off[i]:q_off[i + 1]] for i in sel]
    return np.concatenate(parts), np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.int64)


# ------------------------------------------------------------------ c3 (default)
def run_c3(ts, args):
    """Uneven-length batching workload: 4,096 queries (log-uniform lengths
    [16, 2048]) against 1,000,000 DB random walks (log-uniform [16, 1024]),
    DTW and Frechet top-10. Batched = all queries in one ts_knn call; plain =
    one launch sequence per query (ts_knn_plain) on a sample of queries, whose
    top-10 must equal the batched rows."""
    t0 = time.time()
    db_len = ts.gen_lengths(DB_SEED, args.n_db, 16, 1024, True)
    db_xy, db_off = ts.gen_walks(DB_SEED, db_len)
    q_len = ts.gen_lengths(Q_SEED, args.n_q, 16, 2048, True)
    q_xy, q_off = ts.gen_walks(Q_SEED, q_len)
    db = ts.TrajSet(db_xy, db_off)
    qs = ts.TrajSet(q_xy, q_off)
    Q, N = qs.count, db.count
    ns = min(args.plain_sample, Q)
    sample = ts.TrajSet(q_xy[:q_off[ns]], q_off[:ns + 1])
    dev = ts.device_name()
    log("device", dev)
    log("c3 workload ready in %.1fs: %d DB trajectories (%d points), %d queries" %
        (time.time() - t0, N, len(db_xy), Q))
    eng = ts.Engine(graphs=not args.no_graphs)
    K, W = args.steps, args.warmup
    pmc = load_pmc("pmc_traffic_c3.json")
    out, rows, roof = {}, {}, None
    for meas in ("dtw", "frechet"):
        p = ts.params(meas, mode=ts.EXACT, k=10)
        for _ in range(max(W, 1)):
            eng.knn(db, qs, p)
        # Timed region: K batched calls. Batches this large launch eagerly
        # (no graph replay), and the engine brackets every launch with HIP
        # events on the stream it runs on, for the per-kernel table.
        eng.set_timing(True)
        eng.reset_stats()
        device_sync()
        t1 = time.perf_counter()
        for _ in range(K):
            idx, val, st = eng.knn(db, qs, p)
        device_sync()
        tb = (time.perf_counter() - t1) / K
        kst = eng.kernel_stats()
        eng.set_timing(False)
        rows[meas] = (idx, val)
        t1 = time.perf_counter()
        pi, pv, _ = eng.knn_plain(db, sample, p)
        tp = time.perf_counter() - t1
        same = bool(np.array_equal(pi, idx[:ns]) and np.array_equal(pv.view(np.int32), val[:ns].view(np.int32)))
        out[meas] = {"batched_pairs_per_s": Q * N / tb, "batched_s": round(tb, 4), "graphs": eng.graph_info(),
                     "batched_cell_updates_per_s": st["cells"] / tb, "cells_per_step": st["cells"],
                     "refined_pairs_per_step": st["refined_pairs"], "launches_per_step": st["launches"],
                     "plain_pairs_per_s": ns * N / tp, "plain_s_for_sample": round(tp, 4), "plain_sample": ns,
                     "batched_vs_plain": round((Q * N / tb) / (ns * N / tp), 2), "plain_topk_equals_batched": same,
                     "kernels": {k: {"launches": v["launches"], "ms_mean": round(v["ms"] / max(v["launches"], 1), 4)}
                                 for k, v in sorted(kst.items(), key=lambda kv: -kv[1]["ms"])[:12]}}
        if meas == "dtw":
            roof = roofline_of(kst, pmc, "profiles/pmc_traffic_c3.json")
        if not args.no_graphs:
            out[meas]["graph_replay"] = c3_graph_leg(ts, db, qs, p, idx, val, K, tb)
        log(meas, {k: v for k, v in out[meas].items() if k != "kernels"})
    out["db_file"] = db_file_leg(ts, eng, db, sample)
    parity = None
    if not args.no_oracle:
        parity = c3_parity(ts, rows, db_xy, db_off, q_xy, q_off, q_len)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_ref_pipeline(db_xy, db_off, q_xy, q_off, min(args.cpu_sample, Q),
                               "c3 (queries log-uniform [16,2048], DB 1M log-uniform [16,1024])")
        out["cpu_other"] = cpu_other_c3(db_xy, db_off, q_xy, q_off)
    sample.close()
    qs.close()
    db.close()
    if not args.no_c2:
        out["c2"] = c2_leg(ts, args)
    line = {"metric": METRIC, "value": round(out["dtw"]["batched_pairs_per_s"], 1), "unit": "pairs/s", "n_gpus": 1,
            "steps": K, "warmup": W, "ms_per_step": round(out["dtw"]["batched_s"] * 1e3, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "uneven-length batching: %d query random walks log-uniform [16,2048] vs %d-"
                                   "trajectory DB log-uniform [16,1024], DTW top-10 (Frechet beside it), EXACT "
                                   "bounding-box filter, batched one call (plain method on a %d-query sample)"
                                   % (Q, N, ns),
                       "queries": Q, "db_trajectories": N, "k": 10, "measure": "dtw", "mode": "exact",
                       "device": dev},
            "batched_vs_plain": {"dtw": out["dtw"]["batched_vs_plain"], "frechet": out["frechet"]["batched_vs_plain"]},
            "parity": parity, "roofline": roof, "cpu_baseline": cpu, "extra": out}
    print(json.dumps(line), flush=True)


def c3_graph_leg(ts, db, qs, p, idx, val, K, eager_s):
    """The same batched call recorded as one HIP graph (graph mode 2: every
    batch size): the first call runs eagerly and sizes the workspace, the
    second records, then K replays are timed and must equal the eager rows."""
    g = ts.Engine(graphs=2)
    try:
        g.knn(db, qs, p)
        g.knn(db, qs, p)
        device_sync()
        t1 = time.perf_counter()
        for _ in range(K):
            gi, gv, st = g.knn(db, qs, p)
        device_sync()
        tr = (time.perf_counter() - t1) / K
        same = bool(np.array_equal(gi, idx) and np.array_equal(gv.view(np.int32), val.view(np.int32)))
        return {"replay_s": round(tr, 4), "eager_s": round(eager_s, 4), "replay_over_eager": round(tr / eager_s, 4),
                "graphs": g.graph_info(), "replay_equals_eager": same, "overflow_retries": st["overflow_retries"]}
    finally:
        g.close()


def c3_parity(ts, rows, db_xy, db_off, q_xy, q_off, q_len):
    """The batched rows of the two longest, two shortest and two
    middle-length queries against the oracle's EXACT top-10, bit for bit."""
    import oracle
    order = np.argsort(q_len, kind="stable")
    sel = np.concatenate([order[-2:], order[:2], order[[len(order) // 3, 2 * len(order) // 3]]])
    s_xy, s_off = sub_queries(q_xy, q_off, sel)
    res = {"queries": sel.tolist(), "lengths": q_len[sel].tolist()}
    t0 = time.perf_counter()
    for meas, om in (("dtw", oracle.DTW), ("frechet", oracle.FRECHET)):
        wi, wv, _ = oracle.knn(oracle.params(om, oracle.EXACT, k=10), db_xy, db_off, s_xy, s_off)
        gi, gv = rows[meas]
        res[meas + "_topk_equal"] = bool(np.array_equal(gi[sel], wi) and
                                         np.array_equal(gv[sel].view(np.int32), wv.view(np.int32)))
    res["seconds"] = round(time.perf_counter() - t0, 2)
    res["rule"] = "indices and f32 values of the whole top-10 identical (ties by index)"
    log("c3 parity", res)
    return res


def cpu_ref_pipeline(db_xy, db_off, q_xy, q_off, n_sample, what):
    """The reference's CPU pipeline (dtw.cpp:14-113: envelope r=5, LB_DTW of
    every DB trajectory, the 10 best, exact DTW, best first) as the oracle
    restates it for variable lengths, on the first n_sample queries over the
    whole database, parallel over queries on every core."""
    import oracle
    prm = oracle.params(oracle.DTW, oracle.REF, k=10)
    s_off = q_off[:n_sample + 1]
    t0 = time.perf_counter()
    oracle.knn(prm, db_xy, db_off, q_xy[:s_off[-1]], s_off)
    dt = time.perf_counter() - t0
    n_db = len(db_off) - 1
    cpu = {"value": round(n_sample * n_db / dt, 1), "unit": "pairs/s", "cores": oracle.num_threads(), "kind": "port",
           "sample": "%s: the first %d queries against the whole %d-trajectory DB, the reference's DTW pipeline "
                     "(envelope LB filter, 10 candidates, exact refine), %.2f s" % (what, n_sample, n_db, dt)}
    log("cpu baseline", cpu)
    return cpu


def cpu_other_c3(db_xy, db_off, q_xy, q_off, n_exact=8, ref_num_t=2000, ref_len=256):
    """More CPU figures: the oracle's EXACT search (the same answer as the
    GPU's) on the first n_exact queries, and the reference's own compiled
    launch_dtw_batch_cpu (oracle/_ref, dtw.cpp:14-113 built from the
    reference sources) on its own fixed-length shape: ref_num_t queries
    against ref_num_t DB walks of ref_len points (the workload's mean DB
    length is 242), pairs/s = ref_num_t^2 / seconds."""
    import oracle
    n_db = len(db_off) - 1
    prm = oracle.params(oracle.DTW, oracle.EXACT, k=10)
    t0 = time.perf_counter()
    oracle.knn(prm, db_xy, db_off, q_xy[:q_off[n_exact]], q_off[:n_exact + 1])
    dt = time.perf_counter() - t0
    res = {"oracle_exact_pairs_per_s": round(n_exact * n_db / dt, 1), "oracle_exact_sample": n_exact,
           "oracle_exact_seconds": round(dt, 2), "cores": oracle.num_threads()}
    if oracle.ref_available():
        xy, _ = _walks_fixed(2 * ref_num_t, ref_len)
        t1, t2 = xy[:ref_num_t * ref_len], xy[ref_num_t * ref_len:]
        t0 = time.perf_counter()
        _, ms = oracle.ref_batch(oracle.DTW, t1, t2, ref_num_t, ref_len, ref_len)
        dt = time.perf_counter() - t0
        res.update({"reference_build_pairs_per_s": round(ref_num_t * ref_num_t / dt, 1),
                    "reference_build_ms": round(ms, 1),
                    "reference_build_shape": "launch_dtw_batch_cpu, %d x %d walks of %d points (fixed length, "
                                             "Q = N as the reference requires)" % (ref_num_t, ref_num_t, ref_len)})
    log("cpu other", res)
    return res


def _walks_fixed(count, length):
    import cudaruntime1_amd as ts
    return ts.gen_walks(DB_SEED + 7, np.full(count, length, np.int32))


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


# ------------------------------------------------------------------ c2
def c2_leg(ts, args, want_kernels=False):
    """The five-measure workload: 64 queries vs 100,000 DB random walks,
    lengths uniform [64, 256]; EXACT top-10 batched (graph replay), the plain
    method and single-query latency for DTW; LCSS with eps 2e-3 and window
    16; lock-step Euclidean on its own 100,000 x 128 DB; and "dtw_ref", the
    reference's own DTW pipeline (envelope filter over every DB point, 10
    candidates, exact refine)."""
    db_xy, db_off, q_xy, q_off = build_workload(ts, 100_000, 64, 64, 256)
    db, qs = ts.TrajSet(db_xy, db_off), ts.TrajSet(q_xy, q_off)
    Q, N = qs.count, db.count
    eng = ts.Engine(graphs=not args.no_graphs)
    steps = max(args.steps, 10)
    res = {}
    legs = [("dtw", ts.params("dtw", mode=ts.EXACT, k=10), db, qs),
            ("dtw_ref", ts.params("dtw", mode=ts.REF, k=10), db, qs),
            ("frechet", ts.params("frechet", mode=ts.EXACT, k=10), db, qs),
            ("lcss", ts.params("lcss", mode=ts.EXACT, k=10, epsilon=2e-3, delta=16), db, qs),
            ("hausdorff", ts.params("hausdorff", mode=ts.EXACT, k=10), db, qs)]
    e_db_xy, e_db_off = ts.gen_walks(DB_SEED + 1, np.full(N, 128, np.int32))
    e_q_xy, e_q_off = ts.gen_walks(Q_SEED + 1, np.full(Q, 128, np.int32))
    e_db, e_qs = ts.TrajSet(e_db_xy, e_db_off), ts.TrajSet(e_q_xy, e_q_off)
    legs.append(("euclidean", ts.params("euclidean", mode=ts.EXACT, k=10), e_db, e_qs))
    dtw_rows = None
    for name, p, d, q in legs:
        ms, (gi, gv, st) = timed(lambda: eng.knn(d, q, p), steps, 3)
        res[name] = {"pairs_per_s": q.count * d.count / (ms * 1e-3), "ms_per_step": round(ms, 4),
                     "cells_per_step": st["cells"], "cell_updates_per_s": st["cells"] / (ms * 1e-3),
                     "refined_pairs_per_step": st["refined_pairs"]}
        if name == "dtw":
            dtw_rows = (gi, gv)
            res["graphs"] = eng.graph_info()
    p = legs[0][1]
    ms_plain, (pi, pv, _) = timed(lambda: eng.knn_plain(db, qs, p), max(1, steps // 3), 1)
    res["dtw"].update({"plain_pairs_per_s": Q * N / (ms_plain * 1e-3), "plain_ms_per_step": round(ms_plain, 4),
                       "batched_vs_plain": round(ms_plain / res["dtw"]["ms_per_step"], 2),
                       "plain_equals_batched": bool(np.array_equal(pi, dtw_rows[0]) and
                                                    np.array_equal(pv.view(np.int32), dtw_rows[1].view(np.int32)))})
    lat = []
    singles = [ts.TrajSet(q_xy[q_off[i]:q_off[i + 1]], np.array([0, q_off[i + 1] - q_off[i]], np.int64))
               for i in range(Q)]
    for s in singles[:4]:
        eng.knn(db, s, p)
    for s in singles:
        t1 = time.perf_counter()
        eng.knn(db, s, p)
        lat.append((time.perf_counter() - t1) * 1e3)
    res["dtw"].update({"single_query_ms_p50": round(float(np.percentile(lat, 50)), 4),
                       "single_query_ms_p99": round(float(np.percentile(lat, 99)), 4)})
    for s in singles:
        s.close()
    kst = None
    if want_kernels:
        eng.set_timing(True)
        eng.reset_stats()
        for _ in range(steps):
            ti, tv, _ = eng.knn(db, qs, p)
        kst = eng.kernel_stats()
        eng.set_timing(False)
        assert np.array_equal(ti, dtw_rows[0]) and np.array_equal(tv.view(np.int32), dtw_rows[1].view(np.int32))
    for s in (db, qs, e_db, e_qs):
        s.close()
    eng.close()
    log("c2", {k: (v.get("ms_per_step") if isinstance(v, dict) else v) for k, v in res.items()})
    return (res, kst, (db_xy, db_off, q_xy, q_off, dtw_rows)) if want_kernels else res


def run_c2(ts, args):
    """The 64 x 100k five-measure workload as its own line (value: batched
    EXACT DTW pairs/s), with the DTW roofline of its dominant launch and the
    reference pipeline on the host beside it."""
    if args.copy_calibration:
        copy_calibration()
    res, kst, (db_xy, db_off, q_xy, q_off, dtw_rows) = c2_leg(ts, args, want_kernels=True)
    roof = roofline_of(kst, load_pmc("pmc_traffic_c2.json"), "profiles/pmc_traffic_c2.json")
    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle
        cpu = cpu_ref_pipeline(db_xy, db_off, q_xy, q_off, 64, "64 x 100k, lengths U[64,256]")
        wi, wv, _ = oracle.knn(oracle.params(oracle.DTW, oracle.EXACT, k=10), db_xy, db_off, q_xy, q_off)
        parity = {"queries": 64, "dtw_topk_equal": bool(np.array_equal(wi, dtw_rows[0]) and
                                                        np.array_equal(wv.view(np.int32), dtw_rows[1].view(np.int32)))}
    d = res["dtw"]
    line = {"metric": METRIC, "value": round(d["pairs_per_s"], 1), "unit": "pairs/s", "n_gpus": 1,
            "steps": max(args.steps, 10), "warmup": 3, "ms_per_step": d["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "DTW top-10 k-NN, 64 query random walks vs 100000-trajectory DB (lengths "
                                   "U[64,256], sigma 1e-3), bounding-box/endpoint lower-bound filter, batched",
                       "queries": 64, "db_trajectories": 100_000, "k": 10, "measure": "dtw", "mode": "exact"},
            "parity": parity, "roofline": roof, "cpu_baseline": cpu, "extra": res}
    print(json.dumps(line), flush=True)


# ------------------------------------------------------------------ c1
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
            "cpu_baseline": {"value": round(out["ref"]["cpu_pairs_per_s"], 1), "unit": "pairs/s",
                             "cores": 1, "kind": "port",
                             "sample": "the whole workload (one query: the oracle's loop over queries runs on one "
                                       "thread), the reference's DTW pipeline (REF), mean of 20 runs"},
            "extra": out}
    print(json.dumps(line), flush=True)


# ------------------------------------------------------------------ c4
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
    roof = None
    for meas, p in legs:
        eng.knn(db, sample, p)  # warm-up (workspace growth)
        eng.set_timing(True)
        eng.reset_stats()
        device_sync()
        t1 = time.perf_counter()
        idx, val, st = eng.knn(db, qs, p)
        device_sync()
        tb = time.perf_counter() - t1
        kst = eng.kernel_stats()
        eng.set_timing(False)
        if meas == "dtw":
            roof = roofline_of(kst, load_pmc("pmc_traffic_c4.json"), "profiles/pmc_traffic_c4.json")
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
                                 "topk_equal": bool(np.array_equal(oi, idx[:nq]) and
                                                    np.array_equal(ov.view(np.int32), val[:nq].view(np.int32)))}
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
            "roofline": roof, "cpu_baseline": cpu, "extra": out}
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


# ------------------------------------------------------------------ c5
def run_c5(ts, args):
    """Query serving: a 2,000,000-trajectory DB (log-uniform lengths [16, 512])
    resident in HBM; Poisson arrivals at each stated QPS for args.duration
    seconds; query lengths log-uniform [16, 1024]; measure mix DTW 40% / LCSS
    20% / Frechet 20% / Hausdorff 20%; k uniform in {1, 10, 100}; EXACT top-k.
    Latency is the server's own clock from submit to result ready. The first
    args.check_requests results of the first QPS point are checked against the
    oracle."""
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
    checked = {}
    for qi_point, qps in enumerate(args.qps):
        srv = ts.Server(db, streams=args.streams, max_batch=args.max_batch, batch_wait_us=500)
        n = int(qps * args.duration)
        arrivals = np.cumsum(rng.exponential(1.0 / qps, size=n))
        tickets = []
        refused = 0
        lat = []
        keep = {}
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
                i, t, k = item
                gi, gv, ms = srv.wait(t, k, timeout_ms=120000)
                with lock:
                    lat.append(ms)
                    if qi_point == 0 and i < args.check_requests:
                        keep[i] = (gi, gv)

        ws = [threading.Thread(target=waiter) for _ in range(4)]
        for w in ws:
            w.start()
        start = time.perf_counter()
        last_log = start
        for i in range(n):
            while time.perf_counter() - start < arrivals[i]:
                pass
            p = prm[(int(mix[i]), int(ks[i]))]
            try:
                t = srv.submit(q_xy[q_off[i]:q_off[i + 1]], p)
                with lock:
                    tickets.append((i, t, p.k))
            except ts.TrajsimError as e:
                if e.status != 6:
                    raise
                refused += 1
            if time.perf_counter() - last_log > 5.0:
                last_log = time.perf_counter()
                log("c5 %d QPS: %d of %d submitted" % (qps, i + 1, n))
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
        if qi_point == 0:
            checked = keep
    parity = None
    if not args.no_oracle and checked:
        import oracle
        om = {0: oracle.DTW, 1: oracle.LCSS, 2: oracle.FRECHET, 3: oracle.HAUSDORFF}
        bad = 0
        t1 = time.perf_counter()
        for (m, k), _ in prm.items():
            sel = sorted(i for i in checked if mix[i] == m and ks[i] == k)
            if not sel:
                continue
            s_xy, s_off = sub_queries(q_xy, q_off, sel)
            wi, wv, _ = oracle.knn(oracle.params(om[m], oracle.EXACT, k=int(k), epsilon=2e-3), db_xy, db_off,
                                   s_xy, s_off)
            for r_, i in enumerate(sel):
                gi, gv = checked[i]
                bad += not (np.array_equal(gi, wi[r_]) and np.array_equal(gv.view(np.int32), wv[r_].view(np.int32)))
        parity = {"requests": len(checked), "mismatches": bad, "seconds": round(time.perf_counter() - t1, 2),
                  "rule": "each served top-k (indices, f32 values) equal to the oracle's EXACT top-k"}
        log("c5 parity", parity)
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
            s_xy, s_off = sub_queries(q_xy, q_off, sel)
            meas = {0: oracle.DTW, 1: oracle.LCSS, 2: oracle.FRECHET, 3: oracle.HAUSDORFF}[m]
            op = oracle.params(meas, oracle.EXACT, k=int(k), epsilon=2e-3)
            oracle.knn(op, db_xy, db_off, s_xy, s_off)
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
                                   "k in {1,10,100}, EXACT top-k, continuous batching over %d streams, batches of "
                                   "up to %d queries" % (db.count, args.streams, args.max_batch),
                       "db_trajectories": db.count, "duration_s": args.duration},
            "parity": parity, "cpu_baseline": cpu, "extra": {"sweep": results}}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--qps", type=lambda v: [int(x) for x in v.split(",")], default=[2000, 5000, 10000, 20000])
    ap.add_argument("--duration", type=float, default=10.0)
    ap.add_argument("--streams", type=int, default=4, help="c5: serving lanes")
    ap.add_argument("--max-batch", type=int, default=256, help="c5: queries per serving batch")
    ap.add_argument("--check-requests", type=int, default=64, help="c5: served results checked against the oracle")
    ap.add_argument("--n-db", type=int, default=None)
    ap.add_argument("--n-q", type=int, default=None)
    ap.add_argument("--plain-sample", type=int, default=256)
    ap.add_argument("--cpu-sample", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-oracle", action="store_true", help="skip the parity checks against the CPU oracle")
    ap.add_argument("--no-c2", action="store_true", help="c3: skip the 64 x 100k five-measure figures")
    ap.add_argument("--no-graphs", action="store_true", help="launch eagerly instead of replaying HIP graphs")
    ap.add_argument("--oracle-queries", type=int, default=2,
                    help="c4: check the top-10 of this many queries against the CPU oracle")
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
    if args.config == "c2":
        return run_c2(ts, args)
    if args.config == "c4":
        args.n_db = args.n_db or 20_000
        args.n_q = args.n_q or 256
        args.plain_sample = min(args.plain_sample, 16)
        if args.copy_calibration:
            copy_calibration()
        return run_c4(ts, args)
    if args.config == "c5":
        args.n_db = args.n_db or 2_000_000
        return run_c5(ts, args)
    args.n_db = args.n_db or 1_000_000
    args.n_q = args.n_q or 4096
    if args.copy_calibration:
        copy_calibration()
    return run_c3(ts, args)


if __name__ == "__main__":
    main()
