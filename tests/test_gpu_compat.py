"""Every one of the 12 re-exported reference entry points, called through its
C++ symbol with the reference's argument conventions (fixed lengths, host
Point arrays, h_results per query, times in ms):

  DTW / Frechet / LCSS, the _cpu and the GPU names (dtw.cuh:6-7,
  frechet.cuh:6-8, lcss.cuh:6-8): the h_results of the reference's own CPU
  build in tests/golden, bit for bit;
  launch_hausdorff_batch_cpu_rtree, run_hausdorff_rtree_gpu_pipeline
  (hausdorff.cuh:7-13): best Hausdorff over the top_k nearest MBRs, against
  the oracle's REF pipeline; launch_hausdorff_rtree_gpu (hausdorff.cuh:16-19):
  best over given candidates, against the oracle per pair;
  launch_euclidean_batch_cpu_rtree, launch_euclidean_batch_gpu_rtree_exact
  (euclidean.cuh:7,10): best lock-step distance over the 50 nearest MBRs,
  against the oracle's REF pipeline; execute_rtree_csr_kernel_on_gpu
  (euclidean.cuh:11-13): best over CSR candidate lists, against the oracle.

The _cpu names run on the GPU as well (the library has no CPU path); their
values are those of the reference's CPU functions. The GPU Euclidean names
return euclidean.cpp's double-precision sum rounded to float, not the f32
tree sum of euclidean.cu:38-66."""
import ctypes
import os

import numpy as np
import pytest

import cudaruntime1_amd as ts
import oracle
from tests.helpers import say, walks

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
P = ctypes.c_void_p


@pytest.fixture(scope="module")
def lib():
    ts.init(0)
    return ctypes.CDLL(ts.LIB_PATH)


def ptr(a):
    return P(a.ctypes.data)


def fixtures():
    return sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))


@pytest.mark.parametrize("fname", fixtures())
def test_cpu_and_gpu_names_reproduce_reference_fixtures(lib, fname):
    z = np.load(os.path.join(GOLDEN, fname))
    num_t, n, m, meas = int(z["num_t"]), int(z["n"]), int(z["m"]), int(z["measure"])
    t1 = np.ascontiguousarray(z["t1"], np.float32)
    t2 = np.ascontiguousarray(z["t2"], np.float32)
    want = z["h_results"].view(np.int32)
    ms = ctypes.c_float()
    names = {oracle.DTW: ("_Z20launch_dtw_batch_cpuPK5PointS1_Pfiii", "_Z24launch_dtw_batch_gpu_knnPK5PointS1_PfiiiRf"),
             oracle.FRECHET: ("_Z24launch_frechet_batch_cpuPK5PointS1_Pfiii",
                              "_Z34launch_frechet_batch_gpu_wavefrontPK5PointS1_PfiiiRf"),
             oracle.LCSS: ("_Z21launch_lcss_batch_cpuPK5PointS1_Pfiiif",
                           "_Z35launch_lcss_batch_gpu_wavefront_knnPK5PointS1_PfiiifRf")}
    cpu_name, gpu_name = names[meas]
    say("compat", fname, cpu_name, gpu_name)
    for name, is_gpu in ((cpu_name, False), (gpu_name, True)):
        fn = getattr(lib, name)
        out = np.zeros(num_t, np.float32)
        args = [ptr(t1), ptr(t2), ptr(out), num_t, n, m]
        if meas == oracle.LCSS:
            args.append(ctypes.c_float(float(z["eps"])))
        if is_gpu:
            fn.restype = None
            fn(*args, ctypes.byref(ms))
            assert ms.value > 0
        else:
            fn.restype = ctypes.c_float
            assert fn(*args) > 0
        assert np.array_equal(out.view(np.int32), want), (name, out, z["h_results"])


def fixed_sets(seed, num_t, n, sigma=2e-3):
    xy, _ = walks(seed, np.full(num_t, n, np.int32), sigma)
    return np.ascontiguousarray(xy, np.float32)


def ref_best(measure, t1, num_t1, n, t2, num_t2, m, kc):
    prm = oracle.params(measure, oracle.REF, k=1, kc=kc)
    i, v, _ = oracle.knn(prm, t2, ts.fixed_offsets(num_t2, m), t1, ts.fixed_offsets(num_t1, n))
    return v[:, 0]


@pytest.mark.parametrize("num_t1,n,num_t2,m,top_k", [(40, 33, 300, 57, 10), (25, 64, 7, 20, 10), (16, 100, 500, 100, 3)])
def test_hausdorff_names_match_oracle(lib, num_t1, n, num_t2, m, top_k):
    t1, t2 = fixed_sets(51, num_t1, n), fixed_sets(52, num_t2, m)
    want = ref_best(oracle.HAUSDORFF, t1, num_t1, n, t2, num_t2, m, top_k)
    say("hausdorff names", num_t1, n, num_t2, m, top_k)
    for name in ("_Z32launch_hausdorff_batch_cpu_rtreePK5PointiiS1_iiPfi",
                 "_Z32run_hausdorff_rtree_gpu_pipelinePK5PointiiS1_iiPfi"):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_float
        out = np.zeros(num_t1, np.float32)
        assert fn(ptr(t1), num_t1, n, ptr(t2), num_t2, m, ptr(out), top_k) > 0
        assert np.array_equal(out.view(np.int32), want.view(np.int32)), (name, out, want)
    # given candidates: top_k of them per query, repeats allowed (the
    # reference pads short lists with the last one, hausdorff_rtree.cpp:66-68)
    rng = np.random.default_rng(num_t1 * 1000 + num_t2)
    cand = rng.integers(0, num_t2, size=(num_t1, top_k)).astype(np.int32)
    cand[:, -1] = cand[:, -2]
    fn = lib._Z26launch_hausdorff_rtree_gpuPK5PointiiS1_iiPKiiPfRf
    fn.restype = None
    out = np.zeros(num_t1, np.float32)
    ms = ctypes.c_float()
    fn(ptr(t1), num_t1, n, ptr(t2), num_t2, m, ptr(cand), top_k, ptr(out), ctypes.byref(ms))
    qi = np.repeat(np.arange(num_t1, dtype=np.int32), top_k)
    v = oracle.pair_values(oracle.params(oracle.HAUSDORFF), t2, ts.fixed_offsets(num_t2, m), t1,
                           ts.fixed_offsets(num_t1, n), qi, cand.ravel()).reshape(num_t1, top_k)
    assert np.array_equal(out.view(np.int32), v.min(1).view(np.int32)), (out, v.min(1))


@pytest.mark.parametrize("num_t,n", [(200, 128), (60, 17), (30, 1)])
def test_euclidean_names_match_oracle(lib, num_t, n):
    t1, t2 = fixed_sets(61, num_t, n), fixed_sets(62, num_t, n)
    want = ref_best(oracle.EUCLIDEAN, t1, num_t, n, t2, num_t, n, 50)
    say("euclidean names", num_t, n)
    f_cpu = lib._Z32launch_euclidean_batch_cpu_rtreePK5PointS1_Pfii
    f_cpu.restype = ctypes.c_float
    out = np.zeros(num_t, np.float32)
    assert f_cpu(ptr(t1), ptr(t2), ptr(out), num_t, n) > 0
    assert np.array_equal(out.view(np.int32), want.view(np.int32)), ("cpu", out, want)
    f_gpu = lib._Z38launch_euclidean_batch_gpu_rtree_exactPK5PointS1_PfiiRf
    f_gpu.restype = None
    out = np.zeros(num_t, np.float32)
    ms = ctypes.c_float()
    f_gpu(ptr(t1), ptr(t2), ptr(out), num_t, n, ctypes.byref(ms))
    assert np.array_equal(out.view(np.int32), want.view(np.int32)), ("gpu", out, want)
    # CSR candidate lists of uneven length (at least one each)
    rng = np.random.default_rng(num_t + n)
    counts = rng.integers(1, 12, size=num_t)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    cands = rng.integers(0, num_t, size=int(offsets[-1])).astype(np.int32)
    f_csr = lib._Z31execute_rtree_csr_kernel_on_gpuPK5PointS1_PKiS3_PfiiiRf
    f_csr.restype = None
    out = np.zeros(num_t, np.float32)
    f_csr(ptr(t1), ptr(t2), ptr(cands), ptr(offsets), ptr(out), num_t, n, int(offsets[-1]), ctypes.byref(ms))
    qi = np.repeat(np.arange(num_t, dtype=np.int32), counts)
    v = oracle.pair_values(oracle.params(oracle.EUCLIDEAN), t2, ts.fixed_offsets(num_t, n), t1,
                           ts.fixed_offsets(num_t, n), qi, cands)
    best = np.array([v[offsets[i]:offsets[i + 1]].min() for i in range(num_t)], np.float32)
    assert np.array_equal(out.view(np.int32), best.view(np.int32)), ("csr", out, best)
