"""ctypes binding of the CPU oracle (liboracle.so) and of the reference's own
CPU path (oracle/_ref/libref.so, built from the reference sources).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker. The product (cudaruntime1_amd)
never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(_HERE, "liboracle.so")
REF_PATH = os.path.join(_HERE, "_ref", "libref.so")

DTW, LCSS, FRECHET, HAUSDORFF, EUCLIDEAN = 0, 1, 2, 3, 4
REF, EXACT = 0, 1

if not os.path.exists(ORACLE_PATH):
    raise ImportError("oracle/liboracle.so is not built (make -C oracle)")
_o = ctypes.CDLL(ORACLE_PATH)
_ref = None  # the reference's own build, loaded on first use (_load_ref)

_P = ctypes.c_void_p
_i, _f = ctypes.c_int, ctypes.c_float
for name, res, args in [
    ("orc_envelope", None, [_P, _i, _i, _P]),
    ("orc_lb_dtw", _f, [_P, _i, _P, _i]),
    ("orc_lb_frechet", _f, [_P, _i, _P, _i]),
    ("orc_ub_lcss_count", _i, [_P, _i, _P, _i, _f]),
    ("orc_dtw", _f, [_P, _i, _P, _i, _i]),
    ("orc_dtw_f64", ctypes.c_double, [_P, _i, _P, _i, _i]),
    ("orc_frechet", _f, [_P, _i, _P, _i]),
    ("orc_frechet_f64", ctypes.c_double, [_P, _i, _P, _i]),
    ("orc_lcss", _i, [_P, _i, _P, _i, _f, _i]),
    ("orc_hausdorff", _f, [_P, _i, _P, _i]),
    ("orc_hausdorff_f64", ctypes.c_double, [_P, _i, _P, _i]),
    ("orc_euclidean", _f, [_P, _P, _i]),
    ("orc_mbr", None, [_P, _i, _P]),
    ("orc_box_box_d2", _f, [_P, _P]),
    ("orc_summary_bound", _f, [_P, _P, _i, _P, _i]),
    ("orc_filter_ref", None, [_P, _P, _P, _i, _P, _P, _i, _P]),
    ("orc_knn", _i, [_P, _i, _P, _P, _i, _P, _P, _i, _P, _P, _P]),
    ("orc_pair_values_f64", None, [_P, _P, _P, _P, _P, _P, _P, _i, _P]),
    ("orc_pair_values", None, [_P, _P, _P, _P, _P, _P, _P, _i, _P]),
    ("orc_num_threads", _i, []),
    ("orc_set_threads", None, [_i]),
]:
    fn = getattr(_o, name)
    fn.restype = res
    fn.argtypes = args


def _load_ref():
    """oracle/_ref/libref.so, loaded only by the calls that run it, so a
    process that never times or checks against the reference's own build
    never maps it."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        lib = ctypes.CDLL(REF_PATH)
        for name in ("ref_dtw_cpu", "ref_frechet_cpu"):
            getattr(lib, name).restype = _f
            getattr(lib, name).argtypes = [_P, _P, _P, _i, _i, _i]
        lib.ref_lcss_cpu.restype = _f
        lib.ref_lcss_cpu.argtypes = [_P, _P, _P, _i, _i, _i, _f]
        _ref = lib
    return _ref


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _pts(a):
    return np.ascontiguousarray(a, np.float32).reshape(-1, 2)


def params(measure, mode=REF, k=10, kc=None, r=5, band=0, delta=-1, epsilon=0.5):
    """Packed parameter block (int32 x7 + float32), same field order as ts_params."""
    if kc is None:
        kc = 50 if measure == EUCLIDEAN else 10
    a = np.zeros(8, np.int32)
    a[:7] = [measure, mode, k, kc, r, band, delta]
    a[7:8] = np.array([epsilon], np.float32).view(np.int32)
    return a


def envelope(q, r=5):
    q = _pts(q)
    out = np.empty((len(q), 4), np.float32)
    _o.orc_envelope(_p(q), len(q), r, _p(out))
    return out


def lb_dtw(env, c):
    env, c = np.ascontiguousarray(env, np.float32), _pts(c)
    return _o.orc_lb_dtw(_p(env), len(env), _p(c), len(c))


def lb_frechet(env, c):
    env, c = np.ascontiguousarray(env, np.float32), _pts(c)
    return _o.orc_lb_frechet(_p(env), len(env), _p(c), len(c))


def ub_lcss_count(env, c, eps):
    env, c = np.ascontiguousarray(env, np.float32), _pts(c)
    return _o.orc_ub_lcss_count(_p(env), len(env), _p(c), len(c), eps)


def dtw(q, c, band=0, f64=False):
    q, c = _pts(q), _pts(c)
    return (_o.orc_dtw_f64 if f64 else _o.orc_dtw)(_p(q), len(q), _p(c), len(c), band)


def frechet(q, c, f64=False):
    q, c = _pts(q), _pts(c)
    return (_o.orc_frechet_f64 if f64 else _o.orc_frechet)(_p(q), len(q), _p(c), len(c))


def lcss(q, c, eps, delta=-1):
    q, c = _pts(q), _pts(c)
    return _o.orc_lcss(_p(q), len(q), _p(c), len(c), eps, delta)


def hausdorff(q, c, f64=False):
    q, c = _pts(q), _pts(c)
    return (_o.orc_hausdorff_f64 if f64 else _o.orc_hausdorff)(_p(q), len(q), _p(c), len(c))


def euclidean(q, c):
    q, c = _pts(q), _pts(c)
    assert len(q) == len(c)
    return _o.orc_euclidean(_p(q), _p(c), len(q))


def mbr(t):
    t = _pts(t)
    out = np.empty(4, np.float32)
    _o.orc_mbr(_p(t), len(t), _p(out))
    return out


def box_box_d2(a, b):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    return _o.orc_box_box_d2(_p(a), _p(b))


def summary_bound(prm, q, c):
    q, c = _pts(q), _pts(c)
    return _o.orc_summary_bound(_p(prm), _p(q), len(q), _p(c), len(c))


def filter_ref(prm, db_xy, db_off, q_xy, q_off):
    db_xy, q_xy = _pts(db_xy), _pts(q_xy)
    db_off, q_off = np.ascontiguousarray(db_off, np.int64), np.ascontiguousarray(q_off, np.int64)
    N, Q = len(db_off) - 1, len(q_off) - 1
    out = np.empty((Q, N), np.float32)
    _o.orc_filter_ref(_p(prm), _p(db_xy), _p(db_off), N, _p(q_xy), _p(q_off), Q, _p(out))
    return out


def knn(prm, db_xy, db_off, q_xy, q_off, brute=False):
    """-> (idx [Q,k], val [Q,k], refined_pairs)."""
    db_xy, q_xy = _pts(db_xy), _pts(q_xy)
    db_off, q_off = np.ascontiguousarray(db_off, np.int64), np.ascontiguousarray(q_off, np.int64)
    N, Q = len(db_off) - 1, len(q_off) - 1
    k = int(prm[2])
    idx = np.empty((Q, k), np.int32)
    val = np.empty((Q, k), np.float32)
    st = np.zeros(4, np.int64)
    _o.orc_knn(_p(prm), int(brute), _p(db_xy), _p(db_off), N, _p(q_xy), _p(q_off), Q, _p(idx), _p(val), _p(st))
    return idx, val, int(st[0])


def pair_values(prm, db_xy, db_off, q_xy, q_off, qidx, didx, f64=False):
    db_xy, q_xy = _pts(db_xy), _pts(q_xy)
    db_off, q_off = np.ascontiguousarray(db_off, np.int64), np.ascontiguousarray(q_off, np.int64)
    qidx, didx = np.ascontiguousarray(qidx, np.int32), np.ascontiguousarray(didx, np.int32)
    out = np.empty(len(qidx), np.float64 if f64 else np.float32)
    fn = _o.orc_pair_values_f64 if f64 else _o.orc_pair_values
    fn(_p(prm), _p(db_xy), _p(db_off), _p(q_xy), _p(q_off), _p(qidx), _p(didx), len(qidx), _p(out))
    return out


def num_threads():
    return _o.orc_num_threads()


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops (over queries or pairs)."""
    _o.orc_set_threads(int(n))


# ------------------------------------------------- the reference's own build
def ref_available():
    return os.path.exists(REF_PATH)


def ref_batch(measure, t1, t2, num_t, n, m, eps=0.5):
    """Run the reference's launch_<measure>_batch_cpu (dtw.cpp / frechet.cpp /
    lcss.cpp compiled from the reference sources) -> (h_results, ms)."""
    if _load_ref() is None:
        raise RuntimeError("oracle/_ref/libref.so not built")
    t1, t2 = _pts(t1), _pts(t2)
    out = np.empty(num_t, np.float32)
    if measure == DTW:
        ms = _ref.ref_dtw_cpu(_p(t1), _p(t2), _p(out), num_t, n, m)
    elif measure == FRECHET:
        ms = _ref.ref_frechet_cpu(_p(t1), _p(t2), _p(out), num_t, n, m)
    elif measure == LCSS:
        ms = _ref.ref_lcss_cpu(_p(t1), _p(t2), _p(out), num_t, n, m, eps)
    else:
        raise ValueError("the reference CPU build covers DTW, Frechet and LCSS")
    return out, ms
