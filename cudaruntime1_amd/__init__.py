"""Python binding of libtrajsim.so, the MI355X trajectory k-NN engine.

The binding is a thin ctypes layer over the C ABI in include/trajsim.h; all
compute runs in the HIP library. Importing this package does not touch the
GPU. If libtrajsim.so has not been built the import raises, so nothing can
silently run without the native library.

Measures and modes mirror the reference's entry points
(1li2yang3/CudaRuntime1: dtw.cuh, frechet.cuh, lcss.cuh, hausdorff.cuh,
euclidean.cuh); the 12 original C++ signatures themselves are exported by
libtrajsim.so (include/trajsim_compat.hpp).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtrajsim.so")

DTW, LCSS, FRECHET, HAUSDORFF, EUCLIDEAN = 0, 1, 2, 3, 4
MEASURES = {"dtw": DTW, "lcss": LCSS, "frechet": FRECHET, "hausdorff": HAUSDORFF, "euclidean": EUCLIDEAN}
REF, EXACT = 0, 1
NOALLOC_MAX_PAIRS = 16384  # include/trajsim.h TS_NOALLOC_MAX_PAIRS
SET_TINY_COORD = 1  # TrajSet.flags(): a coordinate c with 0 < |c| < 2^-25
SET_HUGE_COORD = 2  # TrajSet.flags(): a coordinate c with |c| >= 2^62, or not finite

if not os.path.exists(LIB_PATH):
    raise ImportError("libtrajsim.so is not built (run `python cudaruntime1_amd/build.py` or "
                      "__graft_entry__.build()); there is no fallback path")

_lib = ctypes.CDLL(LIB_PATH)


class Params(ctypes.Structure):
    _fields_ = [("measure", ctypes.c_int32), ("mode", ctypes.c_int32), ("k", ctypes.c_int32),
                ("kc", ctypes.c_int32), ("r", ctypes.c_int32), ("band", ctypes.c_int32),
                ("delta", ctypes.c_int32), ("epsilon", ctypes.c_float)]


class Stats(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("ms_device", ctypes.c_double), ("pairs", ctypes.c_int64),
                ("refined_pairs", ctypes.c_int64), ("cells", ctypes.c_int64), ("launches", ctypes.c_int64),
                ("overflow_retries", ctypes.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [("streams", ctypes.c_int32), ("max_batch", ctypes.c_int32), ("max_queue", ctypes.c_int32),
                ("mem_reserve", ctypes.c_int64), ("fake_free_mem", ctypes.c_int64),
                ("batch_wait_us", ctypes.c_int32)]


_P = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)
_sig = {
    "ts_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "ts_last_error": (ctypes.c_char_p, []),
    "ts_params_default": (None, [ctypes.POINTER(Params), ctypes.c_int32, ctypes.c_int32]),
    "ts_init": (ctypes.c_int, [ctypes.c_int32]),
    "ts_device_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32]),
    "ts_set_create": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.POINTER(_P)]),
    "ts_set_reload": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32]),
    "ts_set_destroy": (ctypes.c_int, [_P]),
    "ts_set_info": (ctypes.c_int, [_P, _i32p, ctypes.POINTER(ctypes.c_int64), _i32p, _i32p]),
    "ts_set_lengths": (ctypes.c_int, [_P, _P]),
    "ts_set_flags": (ctypes.c_int, [_P, _i32p]),
    "ts_set_append": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32]),
    "ts_set_save": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "ts_set_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_P)]),
    "ts_db_write": (ctypes.c_int, [ctypes.c_char_p, _P, _P, ctypes.c_int32]),
    "ts_set_bytes": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "ts_engine_create": (ctypes.c_int, [ctypes.POINTER(_P)]),
    "ts_engine_destroy": (ctypes.c_int, [_P]),
    "ts_engine_set_timing": (ctypes.c_int, [_P, ctypes.c_int32]),
    "ts_engine_set_graphs": (ctypes.c_int, [_P, ctypes.c_int32]),
    "ts_dp_shape_count": (ctypes.c_int32, []),
    "ts_dp_cells": (ctypes.c_int64, [_P, ctypes.c_int32, ctypes.c_int32]),
    "ts_dp_shape": (ctypes.c_int, [ctypes.c_int32, _P, _P]),
    "ts_engine_load_tuning": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "ts_engine_dp_fallbacks": (ctypes.c_int, [_P, _P]),
    "ts_engine_set_dp_shape": (ctypes.c_int, [_P] + [ctypes.c_int32] * 6),
    "ts_engine_get_dp_shape": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P, _P, _P]),
    "ts_engine_graph_info": (ctypes.c_int, [_P, _P, _P, _P]),
    "ts_engine_kernel_stats": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32,
                                              ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "ts_engine_reset_stats": (ctypes.c_int, [_P]),
    "ts_knn": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(Params), _P, _P, ctypes.POINTER(Stats)]),
    "ts_knn_plain": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(Params), _P, _P, ctypes.POINTER(Stats)]),
    "ts_knn_multi_k": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(Params), _P, _P, _P, ctypes.POINTER(Stats)]),
    "ts_filter_ref": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(Params), _P]),
    "ts_pair_values": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(Params), _P, _P, ctypes.c_int64, _P]),
    "ts_pair_values_noalloc": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(Params), _P, _P, ctypes.c_int64, _P]),
    "ts_topk_rows": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P, _P]),
    "ts_server_config_default": (None, [ctypes.POINTER(ServerConfig)]),
    "ts_server_create": (ctypes.c_int, [_P, ctypes.POINTER(ServerConfig), ctypes.POINTER(_P)]),
    "ts_server_submit": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.POINTER(Params),
                                        ctypes.POINTER(ctypes.c_uint64)]),
    "ts_server_wait": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_int32, _P, _P, ctypes.POINTER(ctypes.c_double)]),
    "ts_server_cancel": (ctypes.c_int, [_P, ctypes.c_uint64]),
    "ts_server_counters": (ctypes.c_int, [_P, _P]),
    "ts_server_destroy": (ctypes.c_int, [_P]),
    "ts_gen_lengths": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_int32, _P]),
    "ts_gen_walks": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int32, _P, ctypes.c_double, _P]),
}
for _name, (_res, _args) in _sig.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args


# DP variants (rows of the tuning table) and length buckets (largest row
# width the DP visits in each).
DP_VARIANTS = ["dtw", "dtw_band", "frechet", "lcss", "lcss_window"]
DP_BUCKET_CAPS = [31, 63, 127, 255, 1023, 1 << 24]


def dp_cells(p, n, m):
    """DP cells of an n x m pair under params p (band or window included), as
    the engine counts them."""
    return int(_lib.ts_dp_cells(ctypes.byref(p), int(n), int(m)))


def dp_shapes():
    """[(lanes per pair, cells per lane step)] the DP kernels are built for."""
    out = []
    for i in range(_lib.ts_dp_shape_count()):
        w, u = ctypes.c_int32(), ctypes.c_int32()
        _lib.ts_dp_shape(i, ctypes.byref(w), ctypes.byref(u))
        out.append((w.value, u.value))
    return out


class TrajsimError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        msg = _lib.ts_status_string(status).decode()
        detail = _lib.ts_last_error().decode()
        super().__init__("%s: %s (%s)" % (where, msg, detail))


def _check(status, where):
    if status != 0:
        raise TrajsimError(status, where)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def init(device=0):
    _check(_lib.ts_init(device), "ts_init")


def device_name():
    buf = ctypes.create_string_buffer(256)
    _check(_lib.ts_device_name(buf, 256), "ts_device_name")
    return buf.value.decode()


def params(measure, mode=REF, k=None, kc=None, r=None, band=None, delta=None, epsilon=None):
    """ts_params with the reference's defaults, overridden by the arguments."""
    if isinstance(measure, str):
        measure = MEASURES[measure]
    p = Params()
    _lib.ts_params_default(ctypes.byref(p), measure, mode)
    for name, v in (("k", k), ("kc", kc), ("r", r), ("band", band), ("delta", delta), ("epsilon", epsilon)):
        if v is not None:
            setattr(p, name, v)
    return p


# ------------------------------------------------------------------ data
def gen_lengths(seed, count, lo, hi, log_uniform=False):
    out = np.empty(count, np.int32)
    _check(_lib.ts_gen_lengths(seed, count, lo, hi, int(log_uniform), _ptr(out)), "ts_gen_lengths")
    return out


def offsets_of(lengths):
    off = np.zeros(len(lengths) + 1, np.int64)
    np.cumsum(np.asarray(lengths, np.int64), out=off[1:])
    return off


def db_write(path, xy, offsets):
    """Write a tsdb database file from host arrays (no GPU needed)."""
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    offsets = np.ascontiguousarray(offsets, np.int64)
    _check(_lib.ts_db_write(os.fsencode(path), _ptr(xy), _ptr(offsets), len(offsets) - 1), "ts_db_write")


def gen_walks(seed, lengths, sigma=1e-3):
    """Random walks: returns (xy float32 [P,2], offsets int64 [count+1])."""
    off = offsets_of(lengths)
    xy = np.empty((int(off[-1]), 2), np.float32)
    _check(_lib.ts_gen_walks(seed, len(lengths), _ptr(off), sigma, _ptr(xy)), "ts_gen_walks")
    return xy, off


def fixed_offsets(count, length):
    return np.arange(count + 1, dtype=np.int64) * length


# ------------------------------------------------------------------ sets
class TrajSet:
    """Device-resident trajectory set (database or query batch)."""

    def __init__(self, xy, offsets):
        xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
        offsets = np.ascontiguousarray(offsets, np.int64)
        self.count = len(offsets) - 1
        self.offsets = offsets
        self.lengths = np.diff(offsets).astype(np.int32)
        self._h = ctypes.c_void_p()
        _check(_lib.ts_set_create(_ptr(xy), _ptr(offsets), self.count, ctypes.byref(self._h)), "ts_set_create")

    def reload(self, xy, offsets):
        """Replace the trajectories in place (same handle, buffers reused)."""
        xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
        offsets = np.ascontiguousarray(offsets, np.int64)
        _check(_lib.ts_set_reload(self._h, _ptr(xy), _ptr(offsets), len(offsets) - 1), "ts_set_reload")
        self.count = len(offsets) - 1
        self.offsets = offsets
        self.lengths = np.diff(offsets).astype(np.int32)

    def append(self, xy, offsets):
        """Append trajectories (offsets[0] == 0) behind the current ones."""
        xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
        offsets = np.ascontiguousarray(offsets, np.int64)
        _check(_lib.ts_set_append(self._h, _ptr(xy), _ptr(offsets), len(offsets) - 1), "ts_set_append")
        self._sync_lengths()

    def save(self, path):
        """Write the set to a tsdb database file."""
        _check(_lib.ts_set_save(self._h, os.fsencode(path)), "ts_set_save")

    @classmethod
    def load(cls, path):
        """A device-resident set from a tsdb database file."""
        self = cls.__new__(cls)
        self._h = ctypes.c_void_p()
        _check(_lib.ts_set_load(os.fsencode(path), ctypes.byref(self._h)), "ts_set_load")
        self._sync_lengths()
        return self

    def _sync_lengths(self):
        self.count = self.info()["count"]
        self.lengths = np.empty(self.count, np.int32)
        _check(_lib.ts_set_lengths(self._h, _ptr(self.lengths)), "ts_set_lengths")
        self.offsets = np.zeros(self.count + 1, np.int64)
        np.cumsum(self.lengths, out=self.offsets[1:])

    @property
    def handle(self):
        return self._h

    def info(self):
        c, mn, mx = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        p = ctypes.c_int64()
        _check(_lib.ts_set_info(self._h, ctypes.byref(c), ctypes.byref(p), ctypes.byref(mn), ctypes.byref(mx)),
               "ts_set_info")
        return {"count": c.value, "points": p.value, "min_len": mn.value, "max_len": mx.value}

    def nbytes(self):
        b = ctypes.c_int64()
        _check(_lib.ts_set_bytes(self._h, ctypes.byref(b)), "ts_set_bytes")
        return b.value

    def flags(self):
        """Bits found in the points at load: SET_TINY_COORD (a coordinate c
        with 0 < |c| < 2^-25), SET_HUGE_COORD (|c| >= 2^62 or not finite)."""
        f = ctypes.c_int32()
        _check(_lib.ts_set_flags(self._h, ctypes.byref(f)), "ts_set_flags")
        return f.value

    def close(self):
        if self._h:
            _lib.ts_set_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """One stream, its workspace and optional per-kernel timing."""

    def __init__(self, timing=False, graphs=False):
        self._h = ctypes.c_void_p()
        _check(_lib.ts_engine_create(ctypes.byref(self._h)), "ts_engine_create")
        if timing:
            self.set_timing(True)
        if graphs:
            self.set_graphs(graphs)

    def set_timing(self, on):
        _check(_lib.ts_engine_set_timing(self._h, int(on)), "ts_engine_set_timing")

    def set_graphs(self, on):
        """Replay batched k-NN calls of a repeated shape from HIP graphs: 0 off,
        1 (True) batches up to 2^26 (query, DB) pairs, 2 batches of every size."""
        _check(_lib.ts_engine_set_graphs(self._h, int(on)), "ts_engine_set_graphs")

    def load_tuning(self, path):
        _check(_lib.ts_engine_load_tuning(self._h, str(path).encode()), "ts_engine_load_tuning")

    def set_dp_shape(self, variant, bucket, lanes, cells, waves_per_cu, table=2):
        """Force the DP launch shape of one (variant, bucket); see DP_VARIANTS.
        table 0: throughput, 1: small refines, 3: mid-size refines, 4: refines
        of at most 256 pairs, 2: all."""
        _check(_lib.ts_engine_set_dp_shape(self._h, table, variant, bucket, lanes, cells, waves_per_cu),
               "ts_engine_set_dp_shape")

    def get_dp_shape(self, variant, bucket, table=0):
        w, u, g = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(_lib.ts_engine_get_dp_shape(self._h, table, variant, bucket, ctypes.byref(w), ctypes.byref(u),
                                           ctypes.byref(g)), "ts_engine_get_dp_shape")
        return w.value, u.value, g.value

    def dp_fallbacks(self):
        n = ctypes.c_int64()
        _check(_lib.ts_engine_dp_fallbacks(self._h, ctypes.byref(n)), "ts_engine_dp_fallbacks")
        return n.value

    def graph_info(self):
        c, r, n = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(_lib.ts_engine_graph_info(self._h, ctypes.byref(c), ctypes.byref(r), ctypes.byref(n)),
               "ts_engine_graph_info")
        return {"captures": c.value, "replays": r.value, "cached": n.value}

    def reset_stats(self):
        _check(_lib.ts_engine_reset_stats(self._h), "ts_engine_reset_stats")

    def kernel_stats(self):
        out = {}
        i = 0
        name = ctypes.create_string_buffer(128)
        n, ms, b, u = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        while _lib.ts_engine_kernel_stats(self._h, i, name, 128, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b),
                                          ctypes.byref(u)) == 0:
            out[name.value.decode()] = {"launches": n.value, "ms": ms.value, "bytes": b.value, "units": u.value}
            i += 1
        return out

    def _knn(self, fn, db, queries, p):
        Q, k = queries.count, p.k
        idx = np.empty((Q, k), np.int32)
        val = np.empty((Q, k), np.float32)
        st = Stats()
        _check(fn(self._h, db.handle, queries.handle, ctypes.byref(p), _ptr(idx), _ptr(val), ctypes.byref(st)),
               fn.__name__)
        return idx, val, st.as_dict()

    def knn(self, db, queries, p):
        """Batched k-NN -> (idx [Q,k] int32, val [Q,k] float32, stats)."""
        return self._knn(_lib.ts_knn, db, queries, p)

    def knn_multi_k(self, db, queries, p, k_per_query):
        """Batched k-NN with a k per query (each <= p.k) -> (idx [Q,p.k], val, stats);
        row i holds its first k_per_query[i] results, then idx -1."""
        kq = np.ascontiguousarray(k_per_query, np.int32)
        assert len(kq) == queries.count
        Q, k = queries.count, p.k
        idx = np.empty((Q, k), np.int32)
        val = np.empty((Q, k), np.float32)
        st = Stats()
        _check(_lib.ts_knn_multi_k(self._h, db.handle, queries.handle, ctypes.byref(p), _ptr(kq), _ptr(idx),
                                   _ptr(val), ctypes.byref(st)), "ts_knn_multi_k")
        return idx, val, st.as_dict()

    def knn_plain(self, db, queries, p):
        """Plain k-NN (one launch sequence per query) -> (idx, val, stats)."""
        return self._knn(_lib.ts_knn_plain, db, queries, p)

    def filter_ref(self, db, queries, p):
        out = np.empty((queries.count, db.count), np.float32)
        _check(_lib.ts_filter_ref(self._h, db.handle, queries.handle, ctypes.byref(p), _ptr(out)), "ts_filter_ref")
        return out

    def pair_values(self, db, queries, p, qidx, didx):
        qidx = np.ascontiguousarray(qidx, np.int32)
        didx = np.ascontiguousarray(didx, np.int32)
        out = np.empty(len(qidx), np.float32)
        _check(_lib.ts_pair_values(self._h, db.handle, queries.handle, ctypes.byref(p), _ptr(qidx), _ptr(didx),
                                   len(qidx), _ptr(out)), "ts_pair_values")
        return out

    def pair_values_noalloc(self, db, queries, p, qidx, didx):
        """pair_values with no device allocation (at most NOALLOC_MAX_PAIRS pairs)."""
        qidx = np.ascontiguousarray(qidx, np.int32)
        didx = np.ascontiguousarray(didx, np.int32)
        out = np.empty(len(qidx), np.float32)
        _check(_lib.ts_pair_values_noalloc(self._h, db.handle, queries.handle, ctypes.byref(p), _ptr(qidx),
                                           _ptr(didx), len(qidx), _ptr(out)), "ts_pair_values_noalloc")
        return out

    def topk_rows(self, scores, k, descending=False):
        scores = np.ascontiguousarray(scores, np.float32)
        rows, cols = scores.shape
        idx = np.empty((rows, k), np.int32)
        val = np.empty((rows, k), np.float32)
        _check(_lib.ts_topk_rows(self._h, _ptr(scores), rows, cols, k, int(descending), _ptr(idx), _ptr(val)),
               "ts_topk_rows")
        return idx, val

    def close(self):
        if self._h:
            _lib.ts_engine_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching k-NN server over a resident database (ts_server_*)."""

    def __init__(self, db, streams=4, max_batch=256, max_queue=65536, mem_reserve=256 << 20, fake_free_mem=0,
                 batch_wait_us=200):
        cfg = ServerConfig()
        _lib.ts_server_config_default(ctypes.byref(cfg))
        cfg.streams, cfg.max_batch, cfg.max_queue = streams, max_batch, max_queue
        cfg.mem_reserve, cfg.fake_free_mem, cfg.batch_wait_us = mem_reserve, fake_free_mem, batch_wait_us
        self._db = db  # keep the set alive while the server runs
        self._h = ctypes.c_void_p()
        _check(_lib.ts_server_create(db.handle, ctypes.byref(cfg), ctypes.byref(self._h)), "ts_server_create")

    def submit(self, q, p):
        """Queue one query (float32 [len,2]); returns its ticket. Raises TrajsimError
        (status TS_ERR_BUSY = 6) when admission refuses it."""
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 2)
        t = ctypes.c_uint64()
        _check(_lib.ts_server_submit(self._h, _ptr(q), len(q), ctypes.byref(p), ctypes.byref(t)), "ts_server_submit")
        return t.value

    def wait(self, ticket, k, timeout_ms=-1):
        idx = np.empty(k, np.int32)
        val = np.empty(k, np.float32)
        lat = ctypes.c_double()
        _check(_lib.ts_server_wait(self._h, ticket, timeout_ms, _ptr(idx), _ptr(val), ctypes.byref(lat)),
               "ts_server_wait")
        return idx, val, lat.value

    def cancel(self, ticket):
        _check(_lib.ts_server_cancel(self._h, ticket), "ts_server_cancel")

    def counters(self):
        c = np.zeros(5, np.int64)
        _check(_lib.ts_server_counters(self._h, _ptr(c)), "ts_server_counters")
        return dict(zip(["admitted", "refused", "completed", "batches", "cancelled"], c.tolist()))

    def close(self):
        if self._h:
            _lib.ts_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def exported_symbols():
    """Names the library exports from include/trajsim.h (checked by tests)."""
    return sorted(_sig)
