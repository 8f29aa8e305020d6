// server.hip -- continuous-batching request loop over a resident database.
//
// No counterpart in the reference (it has no serving mode). Requests of any
// measure, length and k are queued by ts_server_submit. `streams` worker
// lanes each own an engine (its own HIP stream and workspace) and a reusable
// device query set; a free lane takes every compatible request that is
// waiting (same parameters apart from k, up to max_batch), uploads them into
// its set and runs one batched k-NN at the batch's largest k, each query
// keeping the exact answer for its own k. Requests that arrive while other
// lanes' batches are on the GPU are taken by the next free lane, so their
// work overlaps on the device. Admission is by free device memory: a batch starts only when its
// workspace estimate fits into the free HBM (hipMemGetInfo, or a fixed value
// injected by tests) minus the reserve and minus what running batches hold;
// a request that could never fit is refused, and a full queue refuses with
// TS_ERR_BUSY.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "ts_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t ticket = 0;
  std::vector<ts_point> pts;
  ts_params p{};
  int64_t est_bytes = 0;
  Clock::time_point t_submit, t_done;
  enum State { QUEUED, RUNNING, DONE, CANCELLED, FAILED } state = QUEUED;
  ts_status status = TS_OK;
  std::vector<int32_t> idx;
  std::vector<float> val;
};

// Requests that differ only in k share a batch: it runs at the largest k
// and each query keeps its own first k results (ts_knn_multi_k gives every
// query the exact answer for its own k).
bool batchable(const ts_params& a, const ts_params& b) {
  ts_params x = a, y = b;
  x.k = y.k = 0;
  return std::memcmp(&x, &y, sizeof(ts_params)) == 0;
}

}  // namespace

struct ts_server {
  const ts_set* db = nullptr;
  ts_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<uint64_t, std::shared_ptr<Request>> reqs;
  uint64_t next_ticket = 1;
  int64_t held_bytes = 0;
  bool stop = false;
  int64_t counters[5] = {0, 0, 0, 0, 0};  // admitted, refused, completed, batches, cancelled
  std::vector<std::thread> lanes;
  std::vector<ts_engine*> engines;
  int device = 0;
};

namespace {

// Device workspace of a batch of n queries with sum_len points in all, run
// at k (query points, candidate regions / score rows and keys, seeds sized
// by k), a deliberate over-estimate.
int64_t estimate_bytes(const ts_server* s, int64_t sum_len, int64_t n, const ts_params& p, int32_t k) {
  const int64_t N = s->db->count;
  int64_t per = 256;
  if (p.mode == TS_MODE_REF) {
    per += N * 16 + (int64_t)p.kc * 64;
  } else {
    const int64_t sd = ts::exact_seeds(k, 1e300);  // the most seeds a batch may take
    const int64_t nch = (N + ts::bound_chunk_items((int32_t)sd) - 1) / ts::bound_chunk_items((int32_t)sd);
    per += nch * sd * 8 * 2 + std::min<int64_t>(N, 4096) * 48 + sd * 64;
  }
  return sum_len * 16 + n * per;
}

int64_t free_bytes(ts_server* s) {
  if (s->cfg.fake_free_mem > 0) return s->cfg.fake_free_mem;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
  return (int64_t)fr;
}

void lane_main(ts_server* s, int lane) {
  (void)hipSetDevice(s->device);
  ts_engine* eng = s->engines[lane];
  ts_set* qset = new ts_set();
  qset->device = s->device;
  hipStream_t st = nullptr;
  (void)hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  std::vector<ts_point> xy;
  std::vector<int64_t> off;
  for (;;) {
    std::vector<std::shared_ptr<Request>> batch;
    int64_t batch_bytes = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stop) break;
        if (!s->queue.empty()) {
          // Continuous batching: take the oldest request and every waiting
          // request with the same parameters but possibly another k, while
          // the memory budget allows.
          const int64_t budget = free_bytes(s) - s->cfg.mem_reserve - s->held_bytes;
          const ts_params p0 = s->queue.front()->p;
          int64_t sum_len = 0;
          int32_t kmax = 0;
          for (auto it = s->queue.begin(); it != s->queue.end() && (int)batch.size() < s->cfg.max_batch;) {
            if (batchable((*it)->p, p0)) {
              const int64_t sl = sum_len + (int64_t)(*it)->pts.size();
              const int32_t km = std::max(kmax, (*it)->p.k);
              const int64_t est = estimate_bytes(s, sl, (int64_t)batch.size() + 1, p0, km);
              if (est <= budget) {
                sum_len = sl;
                kmax = km;
                batch_bytes = est;
                batch.push_back(*it);
                it = s->queue.erase(it);
                continue;
              }
            }
            ++it;
          }
          if (!batch.empty()) break;
        }
        s->cv_work.wait_for(lk, std::chrono::microseconds(std::max(50, s->cfg.batch_wait_us)));
      }
      if (s->stop && batch.empty()) break;
      s->held_bytes += batch_bytes;
      for (auto& r : batch) r->state = Request::RUNNING;
      s->counters[3] += 1;
    }
    // Upload the batch into the lane's reusable query set and run it.
    xy.clear();
    off.assign(1, 0);
    for (auto& r : batch) {
      xy.insert(xy.end(), r->pts.begin(), r->pts.end());
      off.push_back((int64_t)xy.size());
    }
    ts_params p = batch.front()->p;
    const int32_t B = (int32_t)batch.size();
    std::vector<int32_t> kq(B);
    for (int32_t i = 0; i < B; ++i) {
      kq[i] = batch[i]->p.k;
      p.k = std::max(p.k, kq[i]);
    }
    std::vector<int32_t> idx((size_t)B * p.k);
    std::vector<float> val((size_t)B * p.k);
    ts_status stt = ts::set_load(qset, xy.data(), off.data(), B, st);
    if (stt == TS_OK) stt = ts_knn_multi_k(eng, s->db, qset, &p, kq.data(), idx.data(), val.data(), nullptr);
    const auto t_done = Clock::now();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->held_bytes -= batch_bytes;
      for (int32_t i = 0; i < B; ++i) {
        auto& r = batch[i];
        r->t_done = t_done;
        if (r->state == Request::CANCELLED) continue;
        r->status = stt;
        if (stt == TS_OK) {
          r->idx.assign(idx.begin() + (size_t)i * p.k, idx.begin() + (size_t)i * p.k + kq[i]);
          r->val.assign(val.begin() + (size_t)i * p.k, val.begin() + (size_t)i * p.k + kq[i]);
          r->state = Request::DONE;
          s->counters[2] += 1;
        } else {
          r->state = Request::FAILED;
        }
      }
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
  ts_set_destroy(qset);
  (void)hipStreamDestroy(st);
}

}  // namespace

using namespace ts;

extern "C" {

void ts_server_config_default(ts_server_config* c) {
  if (!c) return;
  c->streams = 4;
  c->max_batch = 256;
  c->max_queue = 65536;
  c->mem_reserve = 256ll << 20;
  c->fake_free_mem = 0;
  c->batch_wait_us = 200;
}

ts_status ts_server_create(const ts_set* db, const ts_server_config* c, ts_server** out) {
  if (!db || !out) return fail(TS_ERR_INVALID, "ts_server_create: null argument");
  *out = nullptr;
  auto s = new ts_server();
  s->db = db;
  if (c) s->cfg = *c;
  else ts_server_config_default(&s->cfg);
  if (s->cfg.streams < 1 || s->cfg.streams > 16 || s->cfg.max_batch < 1 || s->cfg.max_queue < 1) {
    delete s;
    return fail(TS_ERR_INVALID, "ts_server_create: streams in [1,16], max_batch >= 1, max_queue >= 1");
  }
  (void)hipGetDevice(&s->device);
  for (int i = 0; i < s->cfg.streams; ++i) {
    ts_engine* e = nullptr;
    ts_status st = ts_engine_create(&e);
    if (st != TS_OK) {
      for (auto x : s->engines) ts_engine_destroy(x);
      delete s;
      return st;
    }
    s->engines.push_back(e);
  }
  for (int i = 0; i < s->cfg.streams; ++i) s->lanes.emplace_back(lane_main, s, i);
  *out = s;
  return TS_OK;
}

ts_status ts_server_submit(ts_server* s, const ts_point* q, int32_t len, const ts_params* p, uint64_t* ticket) {
  if (!s || !q || !p || !ticket || len < 1) return fail(TS_ERR_INVALID, "ts_server_submit: bad argument");
  if (p->measure < TS_DTW || p->measure > TS_EUCLIDEAN || p->k < 1 || p->k > 1000 ||
      (p->mode == TS_MODE_REF && p->kc < p->k))
    return fail(TS_ERR_INVALID, "ts_server_submit: bad parameters");
  auto r = std::make_shared<Request>();
  r->pts.assign(q, q + len);
  r->p = *p;
  r->est_bytes = estimate_bytes(s, len, 1, *p, p->k);
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return fail(TS_ERR_CANCELLED, "server is shutting down");
  if ((int64_t)s->queue.size() >= s->cfg.max_queue) {
    s->counters[1] += 1;
    return fail(TS_ERR_BUSY, "request queue is full");
  }
  if (r->est_bytes > free_bytes(s) - s->cfg.mem_reserve) {
    s->counters[1] += 1;
    return fail(TS_ERR_BUSY, "request needs more device memory than the server can ever free");
  }
  r->ticket = s->next_ticket++;
  *ticket = r->ticket;
  s->reqs[r->ticket] = r;
  s->queue.push_back(r);
  s->counters[0] += 1;
  s->cv_work.notify_one();
  return TS_OK;
}

ts_status ts_server_wait(ts_server* s, uint64_t ticket, int32_t timeout_ms, int32_t* out_idx, float* out_val,
                         double* latency_ms) {
  if (!s) return fail(TS_ERR_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(TS_ERR_INVALID, "unknown ticket");
  std::shared_ptr<Request> r = it->second;
  auto finished = [&] {
    return r->state == Request::DONE || r->state == Request::CANCELLED || r->state == Request::FAILED;
  };
  if (timeout_ms < 0) {
    s->cv_done.wait(lk, finished);
  } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), finished)) {
    return fail(TS_ERR_TIMEOUT, "result not ready");
  }
  s->reqs.erase(ticket);
  if (r->state == Request::CANCELLED) return fail(TS_ERR_CANCELLED, "request cancelled");
  if (r->state == Request::FAILED) return r->status;
  if (out_idx) std::copy(r->idx.begin(), r->idx.end(), out_idx);
  if (out_val) std::copy(r->val.begin(), r->val.end(), out_val);
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  return TS_OK;
}

ts_status ts_server_cancel(ts_server* s, uint64_t ticket) {
  if (!s) return fail(TS_ERR_INVALID, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(TS_ERR_INVALID, "unknown ticket");
  auto r = it->second;
  if (r->state == Request::DONE || r->state == Request::FAILED) return TS_OK;  // too late: result stands
  if (r->state == Request::QUEUED) {
    for (auto q = s->queue.begin(); q != s->queue.end(); ++q)
      if (*q == r) {
        s->queue.erase(q);
        break;
      }
  }
  if (r->state != Request::CANCELLED) s->counters[4] += 1;
  r->state = Request::CANCELLED;
  s->cv_done.notify_all();
  return TS_OK;
}

ts_status ts_server_counters(ts_server* s, int64_t* out5) {
  if (!s || !out5) return fail(TS_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  std::copy(s->counters, s->counters + 5, out5);
  return TS_OK;
}

ts_status ts_server_destroy(ts_server* s) {
  if (!s) return TS_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (auto& r : s->queue) {
      r->state = Request::CANCELLED;
      s->counters[4] += 1;
    }
    s->queue.clear();
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  for (auto& t : s->lanes) t.join();
  for (auto e : s->engines) ts_engine_destroy(e);
  delete s;
  return TS_OK;
}

}  // extern "C"
