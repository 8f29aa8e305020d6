// compat.cpp -- the reference's 12 entry points over the C ABI.
//
// Each call uploads its two fixed-length batches into device sets, runs the
// REF pipeline (the reference's filter -> K candidates -> exact refine ->
// best value) and writes the best value per query, as the reference does
// (dtw.cu:99-234, frechet.cu:120-253, lcss.cu:114-250, hausdorff_rtree.cpp:
// 32-83, euclidean_rtree.cpp:19-93). Times are end-to-end milliseconds like
// the reference's, except launch_hausdorff_rtree_gpu, which reports device
// time only (hausdorff.cu:158-166).
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "../../include/trajsim.h"
#include "../../include/trajsim_compat.hpp"

namespace {

std::mutex g_mu;
ts_engine* g_engine = nullptr;

void die(ts_status s, const char* where) {
  std::printf("Error: %s, code: %d, reason: %s (%s)\n", where, (int)s, ts_status_string(s), ts_last_error());
  std::exit(1);
}

#define COMPAT_CHECK(call)                  \
  do {                                      \
    ts_status _s = (call);                  \
    if (_s != TS_OK) die(_s, #call);        \
  } while (0)

ts_engine* engine() {
  if (!g_engine) {
    int dev = 0;
    COMPAT_CHECK(ts_init(dev));
    COMPAT_CHECK(ts_engine_create(&g_engine));
  }
  return g_engine;
}

struct SetGuard {
  ts_set* s = nullptr;
  ~SetGuard() { ts_set_destroy(s); }
};

ts_set* upload(const Point* pts, int count, int len) {
  if (count <= 0 || len <= 0) die(TS_ERR_INVALID, "empty trajectory batch");
  std::vector<int64_t> off(count + 1);
  for (int i = 0; i <= count; ++i) off[i] = (int64_t)i * len;
  ts_set* s = nullptr;
  COMPAT_CHECK(ts_set_create((const ts_point*)pts, off.data(), count, &s));
  return s;
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Best value per query of a REF k-NN with kc candidates.
void ref_best(int measure, const Point* t1, int nq, int n, const Point* t2, int nd, int m, int kc, float eps,
              float* h_results) {
  std::lock_guard<std::mutex> lk(g_mu);
  ts_engine* e = engine();
  SetGuard q, d;
  q.s = upload(t1, nq, n);
  d.s = upload(t2, nd, m);
  ts_params p;
  ts_params_default(&p, measure, TS_MODE_REF);
  p.kc = std::max(1, std::min(kc, nd));
  p.k = 1;
  p.epsilon = eps;
  std::vector<int32_t> idx(nq);
  std::vector<float> val(nq);
  COMPAT_CHECK(ts_knn(e, d.s, q.s, &p, idx.data(), val.data(), nullptr));
  for (int i = 0; i < nq; ++i) h_results[i] = val[i];
}

}  // namespace

void launch_dtw_batch_gpu_knn(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n, int m,
                              float& gpu_time) {
  double t0 = now_ms();
  ref_best(TS_DTW, h_t1, num_t, n, h_t2, num_t, m, 10, 0.0f, h_results);  // K = min(10, num_t), dtw.cu:161
  gpu_time = (float)(now_ms() - t0);
}

float launch_dtw_batch_cpu(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n, int m) {
  double t0 = now_ms();
  ref_best(TS_DTW, h_t1, num_t, n, h_t2, num_t, m, 10, 0.0f, h_results);
  return (float)(now_ms() - t0);
}

void launch_frechet_batch_gpu_wavefront(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n,
                                        int m, float& gpu_time) {
  double t0 = now_ms();
  ref_best(TS_FRECHET, h_t1, num_t, n, h_t2, num_t, m, 10, 0.0f, h_results);  // frechet.cu:181
  gpu_time = (float)(now_ms() - t0);
}

float launch_frechet_batch_cpu(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n, int m) {
  double t0 = now_ms();
  ref_best(TS_FRECHET, h_t1, num_t, n, h_t2, num_t, m, 10, 0.0f, h_results);
  return (float)(now_ms() - t0);
}

void launch_lcss_batch_gpu_wavefront_knn(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n,
                                         int m, float epsilon, float& gpu_time) {
  double t0 = now_ms();
  ref_best(TS_LCSS, h_t1, num_t, n, h_t2, num_t, m, 10, epsilon, h_results);  // lcss.cu:176
  gpu_time = (float)(now_ms() - t0);
}

float launch_lcss_batch_cpu(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n, int m,
                            float epsilon) {
  double t0 = now_ms();
  ref_best(TS_LCSS, h_t1, num_t, n, h_t2, num_t, m, 10, epsilon, h_results);
  return (float)(now_ms() - t0);
}

float launch_hausdorff_batch_cpu_rtree(const Point* h_t1, int num_t1, int n, const Point* h_t2, int num_t2, int m,
                                       float* h_results, int top_k) {
  double t0 = now_ms();
  ref_best(TS_HAUSDORFF, h_t1, num_t1, n, h_t2, num_t2, m, top_k, 0.0f, h_results);
  return (float)(now_ms() - t0);
}

float run_hausdorff_rtree_gpu_pipeline(const Point* h_t1, int num_t1, int n, const Point* h_t2, int num_t2, int m,
                                       float* h_results, int top_k) {
  double t0 = now_ms();
  ref_best(TS_HAUSDORFF, h_t1, num_t1, n, h_t2, num_t2, m, top_k, 0.0f, h_results);
  return (float)(now_ms() - t0);
}

// Min over the given candidates of the exact value; `pairs_of(i)` lists the
// candidates of query i.
template <class F>
static void given_candidates(int measure, const Point* t1, int nq, int n, const Point* t2, int nd, int m,
                             F pairs_of, float empty_value, float* h_results, float& dev_ms) {
  std::lock_guard<std::mutex> lk(g_mu);
  ts_engine* e = engine();
  SetGuard q, d;
  q.s = upload(t1, nq, n);
  d.s = upload(t2, nd, m);
  std::vector<int32_t> qi, di;
  for (int i = 0; i < nq; ++i) {
    for (int c : pairs_of(i)) {
      if (c < 0 || c >= nd) die(TS_ERR_INVALID, "candidate index out of range");
      qi.push_back(i);
      di.push_back(c);
    }
  }
  ts_params p;
  ts_params_default(&p, measure, TS_MODE_REF);
  std::vector<float> v(qi.size());
  double t0 = now_ms();
  COMPAT_CHECK(ts_pair_values(e, d.s, q.s, &p, qi.data(), di.data(), (int64_t)qi.size(), v.data()));
  dev_ms = (float)(now_ms() - t0);
  for (int i = 0; i < nq; ++i) h_results[i] = empty_value;
  std::vector<float> best(nq, INFINITY);
  for (size_t t = 0; t < qi.size(); ++t) best[qi[t]] = std::min(best[qi[t]], v[t]);
  for (int i = 0; i < nq; ++i)
    if (std::isfinite(best[i])) h_results[i] = best[i];
}

void launch_hausdorff_rtree_gpu(const Point* h_t1, int num_t1, int n, const Point* h_t2, int num_t2, int m,
                                const int* h_candidate_indices, int top_k, float* h_results, float& gpu_time) {
  // hausdorff.cu:96-121: min over the query's top_k candidates, 1e38f start.
  given_candidates(
      TS_HAUSDORFF, h_t1, num_t1, n, h_t2, num_t2, m,
      [&](int i) { return std::vector<int>(h_candidate_indices + (size_t)i * top_k, h_candidate_indices + (size_t)(i + 1) * top_k); },
      std::sqrt(1e38f), h_results, gpu_time);
}

float launch_euclidean_batch_cpu_rtree(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n) {
  double t0 = now_ms();
  ref_best(TS_EUCLIDEAN, h_t1, num_t, n, h_t2, num_t, n, 50, 0.0f, h_results);  // euclidean.cpp:90
  return (float)(now_ms() - t0);
}

void launch_euclidean_batch_gpu_rtree_exact(const Point* h_t1, const Point* h_t2, float* h_results, int num_t, int n,
                                            float& gpu_time) {
  double t0 = now_ms();
  ref_best(TS_EUCLIDEAN, h_t1, num_t, n, h_t2, num_t, n, 50, 0.0f, h_results);  // euclidean_rtree.cpp:57
  gpu_time = (float)(now_ms() - t0);
}

void execute_rtree_csr_kernel_on_gpu(const Point* h_t1, const Point* h_t2, const int* h_candidates,
                                     const int* h_offsets, float* h_results, int num_t, int n, int total_candidates,
                                     float& gpu_time) {
  // euclidean.cu:24-72: min over the CSR candidates, FLT_MAX start.
  if (h_offsets[num_t] > total_candidates) die(TS_ERR_INVALID, "CSR offsets exceed total_candidates");
  given_candidates(
      TS_EUCLIDEAN, h_t1, num_t, n, h_t2, num_t, n,
      [&](int i) { return std::vector<int>(h_candidates + h_offsets[i], h_candidates + h_offsets[i + 1]); },
      std::sqrt(FLT_MAX), h_results, gpu_time);
}
