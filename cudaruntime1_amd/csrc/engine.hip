// engine.hip -- k-NN pipelines behind the C ABI.
//
// REF mode (the reference pipeline, e.g. launch_dtw_batch_gpu_knn dtw.cu:
// 99-234): envelope -> per-point filter score of every pair -> kc best
// candidates per query -> exact refine -> best k of the kc by exact value.
// The reference downloads the dense score matrix and selects on the host
// (dtw.cu:156-176); here selection stays on the device.
//
// EXACT mode (no reference counterpart): admissible summary bounds seed each
// query with its s best-bounded candidates, whose exact values give a
// threshold tau (the k-th best); one more filter pass compacts every pair
// whose bound can still reach tau, those are refined, and the top-k is taken
// over them. The result equals brute force over the whole database.
//
// Batched (ts_knn) runs each stage once for a whole batch of queries, with
// persistent refine kernels draining one work list. Plain (ts_knn_plain)
// runs the same stages once per query.
#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "ts_internal.h"

namespace ts {

// DP cells a pair needs: all of them, or those inside a Sakoe-Chiba band
// (|i*C - j*R| <= w*R, R the longer length) or an LCSS index window
// (|i - j| <= delta), counted row by row.
// Cells and pairs of an item list, per DP bucket (acc[b], acc_pairs[b]), so
// that each bucket's launch is credited with its own work. Measures without
// DP count into bucket 0.
__global__ void k_count_cells(const Item* items, const int32_t* nitems, int64_t max_items, SetView q, SetView db,
                              MeasureArgs ma, int32_t by_bucket, unsigned long long* acc,
                              unsigned long long* acc_pairs, int32_t* zero, int32_t nzero) {
  __shared__ unsigned long long part[kDpBuckets], partp[kDpBuckets];
  // the refine's DP work counters start at zero
  if (blockIdx.x == 0 && threadIdx.x < nzero) zero[threadIdx.x] = 0;
  if (threadIdx.x < kDpBuckets) {
    part[threadIdx.x] = 0;
    partp[threadIdx.x] = 0;
  }
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int64_t n = max_items;
  if (nitems) n = min<int64_t>(n, (int64_t)*nitems);
  if (i < n) {
    const Item it = items[i];
    if (it.d >= 0) {
      const int a = q.len[it.q], b = db.len[it.d];
      const int bk = by_bucket ? dp_bucket_of(dp_width_of(a, b, ma)) : 0;
      atomicAdd(&part[bk], pair_cells(a, b, ma.measure, ma.band, ma.delta));
      atomicAdd(&partp[bk], 1ull);
    }
  }
  __syncthreads();
  if (threadIdx.x < kDpBuckets && partp[threadIdx.x] != 0) {
    atomicAdd(acc + threadIdx.x, part[threadIdx.x]);
    atomicAdd(acc_pairs + threadIdx.x, partp[threadIdx.x]);
  }
}

// Everything a refine needs before its DP launches, in one pass over its
// items: with FROM_KEYS the items are first built from top-k keys (key i of
// query q0 + i / s -> Item{q0 + i / s, index or -1, slot i}; seed_idx[i] =
// the index, vals[i] = NaN); then cells and pairs per DP bucket as in
// k_count_cells and, with outb, each item's place in its bucket's two-ended
// list as in k_items_by_bucket. The list counters must be zero on entry.
// With FROM_KEYS and `via`, the keys hold candidate slots instead of DB
// indices: key i becomes a copy of the candidate via[slot] (its query, DB
// index and value slot), the DB index goes to seed_idx[i], no value is
// cleared, and mark[slot] is set to -inf (refined in this round). Without
// FROM_KEYS and with `tb`, items already refined (tb[slot] == -inf) are
// skipped, and items whose filter bound exceeds their query's threshold
// tau[q - tau_q0] are skipped with NaN in pvals[slot]; a skipped item is
// written back with d = -1.
struct PrepExtra {
  const Item* via = nullptr;
  float* mark = nullptr;
  const float* tb = nullptr;
  const float* tau = nullptr;
  int32_t tau_q0 = 0;
  float* pvals = nullptr;
};

template <bool FROM_KEYS>
__global__ void __launch_bounds__(256) k_refine_prep(const uint64_t* keys, int32_t s, int32_t q0, int32_t* seed_idx,
                                                     float* nan_vals, Item* items, const int32_t* nitems,
                                                     int64_t max_items, SetView q, SetView db, MeasureArgs ma,
                                                     int32_t by_bucket, unsigned long long* acc,
                                                     unsigned long long* acc_pairs, Item* outb, int64_t cap_each,
                                                     int32_t* heads, int32_t* tails, BucketCells long_cells,
                                                     PrepExtra x) {
  __shared__ unsigned long long part[kDpBuckets], partp[kDpBuckets];
  if (threadIdx.x < kDpBuckets) {
    part[threadIdx.x] = 0;
    partp[threadIdx.x] = 0;
  }
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int64_t n = max_items;
  if (nitems) n = min<int64_t>(n, (int64_t)*nitems);
  Item it{0, -1, 0, 0};
  if (i < n) {
    if (FROM_KEYS) {
      const uint64_t k = keys[i];
      if (x.via) {
        if (k != ~0ull) {
          const Item c = x.via[(uint32_t)k];
          it = Item{c.q, c.d, c.slot, 0};
          x.mark[c.slot] = __int_as_float(0xff800000);
        } else {
          it = Item{q0 + (int32_t)(i / s), -1, 0, 0};
        }
        items[i] = it;
        seed_idx[i] = it.d;
      } else {
        it = Item{q0 + (int32_t)(i / s), k == ~0ull ? -1 : (int32_t)(uint32_t)k, (int32_t)i, 0};
        items[i] = it;
        seed_idx[i] = it.d;
        nan_vals[i] = __int_as_float(0x7fc00000);
      }
    } else {
      it = items[i];
      if (x.tb && it.d >= 0) {
        const float b = x.tb[it.slot];
        bool skip = b == __int_as_float(0xff800000);
        if (!skip && b > x.tau[it.q - x.tau_q0]) {
          x.pvals[it.slot] = __int_as_float(0x7fc00000);
          skip = true;
        }
        if (skip) {
          it.d = -1;
          items[i] = it;
        }
      }
    }
  }
  int b = -1;
  bool lng = false;
  if (it.d >= 0) {
    const int na = q.len[it.q], nb = db.len[it.d];
    b = by_bucket ? dp_bucket_of(dp_width_of(na, nb, ma)) : 0;
    atomicAdd(&part[b], pair_cells(na, nb, ma.measure, ma.band, ma.delta));
    atomicAdd(&partp[b], 1ull);
    lng = (int64_t)na * nb >= long_cells.t[b];
  }
  if (outb) {  // grid-uniform
    const int lane = threadIdx.x & 63;
    for (int bb = 0; bb < kDpBuckets; ++bb) {
      for (int e = 0; e < 2; ++e) {
        const bool mine = b == bb && lng == (e == 0);
        const uint64_t bal = __ballot(mine);
        if (bal == 0) continue;
        const int leader = __ffsll((unsigned long long)bal) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(e == 0 ? &heads[bb] : &tails[bb], __popcll(bal));
        base = __shfl(base, leader);
        if (mine) {
          const int64_t pos = base + __popcll(bal & ((1ull << lane) - 1ull));
          if (pos < cap_each) outb[bb * cap_each + (e == 0 ? pos : cap_each - 1 - pos)] = it;
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < kDpBuckets && partp[threadIdx.x] != 0) {
    atomicAdd(acc + threadIdx.x, part[threadIdx.x]);
    atomicAdd(acc_pairs + threadIdx.x, partp[threadIdx.x]);
  }
}

// Upper bound of each seed pair's value from one warping path: along the
// longer trajectory (L points), point t is matched with point t*S/L of the
// shorter (S points), a monotone path from (0,0) to (L-1,S-1) with unit
// steps that lies in every Sakoe-Chiba band of width w >= 1 (|i*m - j*n| is
// then below max(n,m)). DTW's value is the least f32 running sum over its
// paths, so at most this path's running sum, which the lanes' sum (another
// order) times 1 + (L + 64) 2^-22 exceeds; Frechet's is at most the path's
// largest cost (max never rounds). One wave per seed; seed_idx[i] gets the
// seed's DB index (-1 and +inf for missing seeds).
// Upper bound of pair (query qi, DB trajectory d) by one wave (see k_seed_ub),
// in every lane.
__device__ float path_ub(const SetView& q, const SetView& db, int qi, int d, int32_t measure) {
  const int lane = threadIdx.x & 63;
  const int n = q.len[qi], m = db.len[d];
  const int64_t qo = q.off[qi], dof = db.off[d];
  const bool qlong = n >= m;
  const int L = qlong ? n : m, S = qlong ? m : n;
  const float *Ax = qlong ? q.x + qo : db.x + dof, *Ay = qlong ? q.y + qo : db.y + dof;
  const float *Bx = qlong ? db.x + dof : q.x + qo, *By = qlong ? db.y + dof : q.y + qo;
  float acc = 0.0f;
  for (int t = lane; t < L; t += 64) {
    const int j = (int)(((int64_t)t * S) / L);
    const float dx = Ax[t] - Bx[j], dy = Ay[t] - By[j];
    const float c2 = dx * dx + dy * dy;
    acc = measure == TS_DTW ? acc + sqrtf(c2) : fmaxf(acc, c2);
  }
  for (int sh = 32; sh > 0; sh >>= 1) {
    const float o = __shfl_xor(acc, sh);
    acc = measure == TS_DTW ? acc + o : fmaxf(acc, o);
  }
  return measure == TS_DTW ? acc * (1.0f + (float)(L + 64) * 0x1p-22f) : sqrtf(acc);
}

__global__ void __launch_bounds__(256) k_seed_ub(const uint64_t* keys, int32_t s, int32_t q0, int64_t n_seeds,
                                                 SetView q, SetView db, int32_t measure, int32_t* seed_idx,
                                                 float* ub) {
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n_seeds) return;
  const uint64_t key = keys[i];
  const int d = key == ~0ull ? -1 : (int32_t)(uint32_t)key;
  if (d < 0) {
    if (lane == 0) {
      seed_idx[i] = -1;
      ub[i] = __int_as_float(0x7f800000);
    }
    return;
  }
  const float u = path_ub(q, db, q0 + (int)(i / s), d, measure);
  if (lane == 0) {
    seed_idx[i] = d;
    ub[i] = u;
  }
}

// Second-round pool: keys of candidate slots (low 32 bits) to keys of
// (upper bound, slot), one wave per key; missing keys stay last.
__global__ void __launch_bounds__(256) k_slot_ub_keys(const uint64_t* keys, int64_t n, const Item* citems, SetView q,
                                                      SetView db, int32_t measure, uint64_t* out) {
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const uint64_t key = keys[i];
  if (key == ~0ull) {
    if ((threadIdx.x & 63) == 0) out[i] = ~0ull;
    return;
  }
  const uint32_t slot = (uint32_t)key;
  const Item it = citems[slot];
  const float u = path_ub(q, db, it.q, it.d, measure);
  if ((threadIdx.x & 63) == 0) out[i] = ((uint64_t)__float_as_uint(u) << 32) | slot;
}

// Keys (upper bound, DB index) of a seed pool, for picking the seeds of
// smallest upper bound: missing seeds (index < 0) sort last.
__global__ void k_ub_keys(const float* ub, const int32_t* idx, int64_t n, uint64_t* keys) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int32_t d = idx[i];
  keys[i] = d < 0 ? ~0ull : ((uint64_t)__float_as_uint(ub[i]) << 32) | (uint32_t)d;
}

__global__ void k_fill_f32(float* p, float v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

enum BufId {
  B_ENV, B_SCORES, B_KEYS, B_KEYS2, B_SCA, B_SCB, B_CAND, B_CANDIDX, B_ITEMS, B_ITEMS3, B_VALS, B_REGION,
  B_RVALS, B_COUNTS, B_CTR, B_TAU, B_TOP, B_OIDX, B_OVAL, B_CELLS, B_HOSTIN, B_ITEMS2, B_RITEMS, B_SEGC, B_NEED, B_BCNT,
  B_BITS, B_BOUNDS, B_KQ, B_TB, B_TOP2, B_CANDIDX2, B_ITEMS5, B_MKEYS, B_TOPA, B_UB, B_TAUUB, B_PRE, B_POOL,
  B_POOLK, B_POOLIDX, B_TOP2P, B_KEYS2P, B_NBUF
};

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
};

struct KStat {
  std::string name;
  int64_t launches = 0;
  double ms = 0, bytes = 0, units = 0;
};

// DP launch shape and persistent grid (waves per CU) of one bucket.
struct DpTune {
  int shape;
  int waves_per_cu;
};
constexpr int kDpVariants = 5;  // DTW, DTW banded, Frechet, LCSS, LCSS windowed
// DP shape tables by refine size (pairs the refine can hold): few pairs (at
// most one per CU: a CU's waves may share a pair, the multi-wave shapes),
// small refines (one round of the grid or less: a pair's latency decides),
// mid-size ones (a few rounds: the main refine of 64 queries against 100k
// trajectories), and throughput.
constexpr int64_t kFewPairs = 256;
constexpr int64_t kLatencyPairs = 4096;
constexpr int64_t kMidPairs = 65536;
constexpr int kTables = 4;  // 0: throughput, 1: small, 2: mid-size, 3: few
int dp_table_of(int64_t max_items) {
  if (max_items <= kFewPairs) return 3;
  return max_items <= kLatencyPairs ? 1 : (max_items <= kMidPairs ? 2 : 0);
}

struct Pending {
  int kid;
  hipEvent_t a, b;
  double bytes, units;
  int cells_slot;  // >= 0: units are read from the cell counter slot
};

// A captured batch: the key lists every host-side input its launches were
// built from (set handles and shapes, parameters, batch range, candidate
// capacity, workspace generation), so a replay is valid exactly when the key
// matches; the data behind the device pointers may change between replays.
struct GraphEntry {
  std::vector<int64_t> key;
  hipGraphExec_t exec = nullptr;
  int64_t launches = 0;
  int cells_to = 0;
  uint64_t last_use = 0;
};

}  // namespace ts

using namespace ts;

struct ts_engine {
  hipStream_t st = nullptr;
  int device = 0;
  bool timing = false;
  int64_t launches = 0;
  Buf bufs[B_NBUF];
  std::vector<KStat> ks;
  std::map<std::string, int> kidx;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> evpool;
  size_t ev_used = 0;
  hipEvent_t call_a = nullptr, call_b = nullptr;
  // Fork/join: per DP bucket a stream, so bucket launches overlap.
  hipStream_t aux[kDpBuckets] = {};
  hipEvent_t fork = nullptr, join[kDpBuckets] = {};
  std::vector<unsigned long long> cells_host;
  int cells_used = 0;
  int64_t overflow_retries = 0;
  // DP shape per (table, variant, bucket): defaults, then the tuning table
  // (tables as in dp_table_of).
  DpTune tune[kTables][kDpVariants][kDpBuckets];
  std::string tuning_source = "defaults";
  int64_t dp_fallbacks = 0;  // launches that fell back to the default shape
  // hipGraph replay of batches (off by default; never with timing or for
  // the plain method).
  bool graphs = false;
  bool graphs_all = false;  // also batches above kGraphMaxPairs
  bool capturing = false;       // launches are being recorded, not run
  bool capture_broken = false;  // the workspace had to grow while recording
  uint64_t ws_gen = 0;          // bumped whenever a workspace buffer moves
  uint64_t use_clock = 0;
  int64_t graph_captures = 0, graph_replays = 0;
  std::vector<GraphEntry> gcache;
  std::vector<std::vector<int64_t>> gseen;  // keys run once eagerly
  int64_t exact_cap = 0;                    // candidate capacity of device-sized batches
  unsigned long long* need_dev = nullptr;   // largest candidate total of the call
  // host staging of the eager candidate sizing; engine-owned so that async
  // copies from them stay valid
  std::vector<int32_t> h_counts;
  std::vector<int64_t> h_off;
  // Pinned host staging for a call's results and counters: one stream
  // synchronisation brings them all back.
  char* h_stage = nullptr;
  size_t h_stage_bytes = 0;
};

namespace ts {
namespace {


hipError_t zero_cells(hipStream_t st, unsigned long long* cells) {
  hipLaunchKernelGGL(k_zero_u64, dim3((kCellBuf + 255) / 256), dim3(256), 0, st, cells, kCellBuf);
  return hipGetLastError();
}

template <class T>
ts_status ensure(ts_engine* e, BufId id, int64_t n, T** out) {
  Buf& b = e->bufs[id];
  size_t bytes = (size_t)std::max<int64_t>(n, 1) * sizeof(T);
  if (bytes > b.cap) {
    if (e->capturing) {
      e->capture_broken = true;
      return fail(TS_ERR_BUSY, "workspace growth while recording a graph");
    }
    ++e->ws_gen;
    TS_HIP(hipStreamSynchronize(e->st));
    if (b.p) TS_HIP(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    size_t want = std::max(bytes, (size_t)4096);
    hipError_t err = hipMalloc(&b.p, want);
    if (err != hipSuccess) {
      (void)hipGetLastError();
      return fail(TS_ERR_OOM, "device workspace allocation of " + std::to_string(want) + " bytes failed");
    }
    b.cap = want;
  }
  *out = (T*)b.p;
  return TS_OK;
}

int kernel_id(ts_engine* e, const std::string& name) {
  auto it = e->kidx.find(name);
  if (it != e->kidx.end()) return it->second;
  int id = (int)e->ks.size();
  e->ks.push_back(KStat{name});
  e->kidx[name] = id;
  return id;
}

hipEvent_t next_event(ts_engine* e) {
  if (e->ev_used == e->evpool.size()) {
    hipEvent_t ev;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    e->evpool.push_back(ev);
  }
  return e->evpool[e->ev_used++];
}

// Enqueue one launcher, optionally bracketed by timing events.
template <class F>
ts_status run_on(ts_engine* e, hipStream_t st, const char* name, double bytes, double units, F f,
                 int cells_slot = -1, int nlaunch = 1) {
  hipEvent_t a = nullptr;
  if (e->timing) {
    a = next_event(e);
    if (a) (void)hipEventRecord(a, st);
  }
  hipError_t err = f();
  if (err != hipSuccess) return hip_fail(err, name, __FILE__, __LINE__);
  e->launches += nlaunch;
  if (e->timing && a) {
    hipEvent_t b = next_event(e);
    if (b) {
      (void)hipEventRecord(b, st);
      e->pending.push_back(Pending{kernel_id(e, name), a, b, bytes, units, cells_slot});
    }
  }
  return TS_OK;
}

template <class F>
ts_status run(ts_engine* e, const char* name, double bytes, double units, F f, int cells_slot = -1,
              int nlaunch = 1) {
  return run_on(e, e->st, name, bytes, units, f, cells_slot, nlaunch);
}

ts_status resolve_timing(ts_engine* e) {
  if (!e->pending.empty()) {
    TS_HIP(hipStreamSynchronize(e->st));
    if (e->cells_used > 0) {
      unsigned long long* dc = (unsigned long long*)e->bufs[B_CELLS].p;
      e->cells_host.resize(e->cells_used);
      TS_HIP(hipMemcpy(e->cells_host.data(), dc, e->cells_used * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    }
    for (const Pending& p : e->pending) {
      float ms = 0;
      TS_HIP(hipEventElapsedTime(&ms, p.a, p.b));
      KStat& k = e->ks[p.kid];
      k.launches += 1;
      k.ms += ms;
      k.bytes += p.bytes;
      k.units += p.cells_slot >= 0 && p.cells_slot < (int)e->cells_host.size() ? (double)e->cells_host[p.cells_slot]
                                                                               : p.units;
    }
  }
  e->pending.clear();
  e->ev_used = 0;
  return TS_OK;
}

// Halo of the tuned shape; a shape whose halo does not fit a block's LDS
// (a long band on a narrow group) falls back to the bucket's default shape.
DpHalo shape_halo(ts_engine* e, int& sh, int b, const MeasureArgs& ma, int32_t hi, int32_t whi, int32_t rmax) {
  DpHalo h = dp_halo(sh, b, ma, hi, whi);
  if (!dp_fits(sh, h, rmax) && sh != kDpDefaultShape[b]) {
    sh = kDpDefaultShape[b];
    h = dp_halo(sh, b, ma, hi, whi);
    ++e->dp_fallbacks;
  }
  return h;
}

// Kernel-statistics label of a DP launch: measure, bucket and shape, e.g.
// "dtw_wavefront/b3/w32u4" (the k_dp<measure, 32, 4, ...> instantiation).
std::string dp_label(const char* base, int b, int sh) {
  return std::string(base) + "/b" + std::to_string(b) + "/w" + std::to_string(kDpShapeList[sh].W) + "u" +
         std::to_string(kDpShapeList[sh].U);
}

// DP variant of a parameter set: its row in the tuning table.
int dp_variant(const ts_params& p) {
  switch (p.measure) {
    case TS_DTW: return p.band != 0 ? 1 : 0;
    case TS_FRECHET: return 2;
    default: return p.delta >= 0 ? 4 : 3;
  }
}

MeasureArgs margs(const ts_params& p) { return MeasureArgs{p.measure, p.band, p.delta, p.epsilon}; }

bool desc_measure(int m) { return m == TS_LCSS; }

ts_status check_params(const ts_params* p, const ts_set* db, const ts_set* qs) {
  if (!p || !db || !qs) return fail(TS_ERR_INVALID, "null argument");
  if (p->measure < TS_DTW || p->measure > TS_EUCLIDEAN) return fail(TS_ERR_INVALID, "unknown measure");
  if (p->mode != TS_MODE_REF && p->mode != TS_MODE_EXACT) return fail(TS_ERR_INVALID, "unknown mode");
  if (p->k < 1 || p->k > 1000) return fail(TS_ERR_INVALID, "k must be in [1, 1000]");
  if (p->mode == TS_MODE_REF && (p->kc < p->k || p->kc > 1000))
    return fail(TS_ERR_INVALID, "REF mode needs k <= kc <= 1000");
  if (p->band < -100) return fail(TS_ERR_INVALID, "band must be >= -100 (a percentage when negative)");
  if (p->band != 0 && p->measure != TS_DTW) return fail(TS_ERR_UNSUPPORTED, "band is a DTW option");
  if (p->delta >= 0 && p->measure != TS_LCSS) return fail(TS_ERR_UNSUPPORTED, "delta is an LCSS option");
  if (p->measure == TS_LCSS && !(p->epsilon > 0.0f)) return fail(TS_ERR_INVALID, "LCSS needs epsilon > 0");
  if (p->mode == TS_MODE_REF && p->r < 0) return fail(TS_ERR_INVALID, "r must be >= 0");
  if (db->device != qs->device) return fail(TS_ERR_INVALID, "sets live on different devices");
  return TS_OK;
}

// Items a seed refine builds on the device from the seeds' top-k keys
// (k_refine_prep<true>): into `items`, the DB indices into seed_idx.
struct KeyItems {
  const uint64_t* keys;
  int32_t s, q0;
  int32_t* seed_idx;
  const Item* via = nullptr;  // keys hold candidate slots of via (second seed round)
  float* mark = nullptr;      // tb of the candidates: -inf where refined
};

// Refine a device item list into vals[slot]. nitems_dev may be null (all
// max_items valid positions; items with d < 0 are skipped). ctr_zeroed: a
// set of 32 DP counters the caller has already zeroed, which lets one
// kernel prepare the refine (required with from_keys); otherwise the
// refine's own set is zeroed by its cell count.
// kind: 0 a main refine, 1 a seed refine, 2 the second seed round (launch
// labels). prune: the main refine's candidate bounds tb[slot] and the
// threshold they are tested against (items over it, or refined in the
// second round, are skipped).
struct Prune {
  const float* tb;
  const float* tau;
};
ts_status refine(ts_engine* e, const ts_params& p, const ts_set* db, const ts_set* qs, Item* items,
                 const int32_t* nitems_dev, int64_t max_items, float* vals, int kind = 0,
                 int32_t* ctr_zeroed = nullptr, const KeyItems* from_keys = nullptr,
                 const RefineBufs* fixed = nullptr, const float* tau = nullptr, int32_t tau_q0 = 0,
                 const Prune* prune = nullptr) {
  const bool seed = kind == 1;
  if (max_items == 0) return TS_OK;
  if (from_keys && !ctr_zeroed) return fail(TS_ERR_INVALID, "items from keys need zeroed counters");
  if (fixed && max_items > fixed->max_items) return fail(TS_ERR_INVALID, "refine larger than its fixed buffers");
  const SetView qv = qs->view(), dv = db->view();
  const MeasureArgs ma = margs(p);
  const int cus = device_cu_count();
  unsigned long long* cells = nullptr;
  if (fixed) cells = fixed->cells;
  else TS_TRY(ensure(e, B_CELLS, kCellBuf, &cells));
  // kDpBuckets counter slots for this refine (the last group is shared once
  // the slots run out; totals stay right, per-launch credit blurs)
  int slot = e->cells_used;
  if (slot + kDpBuckets <= kCellSlots) e->cells_used += kDpBuckets;
  else slot = kCellSlots - kDpBuckets;
  const bool dp = p.measure == TS_DTW || p.measure == TS_FRECHET || p.measure == TS_LCSS;
  // DP counters: [0, B) work counters of the bucket launches, [B, 2B) and
  // [2B, 3B) the heads and tails of the bucket lists
  int32_t* ctr = ctr_zeroed;
  if (!ctr) {
    if (fixed) ctr = fixed->ctr;
    else TS_TRY(ensure(e, B_CTR, 32, &ctr));
  }
  static_assert(3 * kDpBuckets <= 32, "DP counters");
  // Buckets of the row width the DP visits: one persistent launch per bucket
  // that can occur for these two sets (the width grows with both lengths).
  const int32_t hi = std::min(qs->max_len, db->max_len);
  const int32_t wlo = dp_width_of(qs->min_len, db->min_len, ma), whi = dp_width_of(qs->max_len, db->max_len, ma);
  const int b_lo = dp_bucket_of(wlo), b_hi = dp_bucket_of(whi);
  const int32_t rmax = std::max(qs->max_len, db->max_len);  // longest row of any pair
  if (dp && p.measure == TS_DTW && p.band != 0 && (int64_t)rmax * rmax >= (1ll << 31))
    return fail(TS_ERR_UNSUPPORTED, "band needs trajectory lengths below 46341");
  // With several buckets, each gets a work list, longest pairs first: pairs
  // of at least half the bucket's largest n*m lead it.
  const bool lists = dp && b_lo != b_hi;
  Item* itemsb = nullptr;
  BucketCells long_cells{};
  if (lists) {
    if (fixed) itemsb = fixed->itemsb;
    else TS_TRY(ensure(e, B_ITEMS3, (int64_t)kDpBuckets * max_items, &itemsb));
    for (int b = 0; b < kDpBuckets; ++b) {
      const int64_t cb = dp_ranged(ma) ? hi : std::min<int64_t>(kDpBucketCap[b], hi);
      long_cells.t[b] = cb * rmax / 2;
    }
  }
  if (prune && (!ctr_zeroed || from_keys)) return fail(TS_ERR_INVALID, "pruning needs a prepared main refine");
  if (ctr_zeroed) {
    TS_TRY(run(e, "refine_prep", 0, 0, [&] {
      const dim3 grid((unsigned)((max_items + 255) / 256));
      PrepExtra x;
      if (from_keys) {
        x.via = from_keys->via;
        x.mark = from_keys->mark;
        hipLaunchKernelGGL(k_refine_prep<true>, grid, dim3(256), 0, e->st, from_keys->keys, from_keys->s,
                           from_keys->q0, from_keys->seed_idx, vals, items, nitems_dev, max_items, qv, dv, ma,
                           dp ? 1 : 0, cells + slot, cells + kCellSlots + slot, lists ? itemsb : nullptr, max_items,
                           ctr + kDpBuckets, ctr + 2 * kDpBuckets, long_cells, x);
      } else {
        if (prune) {
          x.tb = prune->tb;
          x.tau = prune->tau;
          x.tau_q0 = tau_q0;
          x.pvals = vals;
        }
        hipLaunchKernelGGL(k_refine_prep<false>, grid, dim3(256), 0, e->st, nullptr, 1, 0, nullptr, nullptr, items,
                           nitems_dev, max_items, qv, dv, ma, dp ? 1 : 0, cells + slot, cells + kCellSlots + slot,
                           lists ? itemsb : nullptr, max_items, ctr + kDpBuckets, ctr + 2 * kDpBuckets, long_cells,
                           x);
      }
      return hipGetLastError();
    }));
  } else {
    TS_TRY(run(e, "count_cells", 0, 0, [&] {
      hipLaunchKernelGGL(k_count_cells, dim3((unsigned)((max_items + 255) / 256)), dim3(256), 0, e->st, items,
                         nitems_dev, max_items, qv, dv, ma, dp ? 1 : 0, cells + slot, cells + kCellSlots + slot,
                         ctr, dp ? 3 * kDpBuckets : 0);
      return hipGetLastError();
    }));
    if (lists) {
      TS_TRY(run(e, "items_by_bucket", 0, 0, [&] {
        return launch_items_by_bucket(e->st, items, nitems_dev, max_items, qv, dv, ma, itemsb, max_items,
                                      ctr + kDpBuckets, ctr + 2 * kDpBuckets, long_cells);
      }));
    }
  }
  if (p.measure == TS_HAUSDORFF) {
    return run(e, seed ? "hausdorff_refine_seed" : "hausdorff_refine", 0, 0, [&] {
      return launch_hausdorff(e->st, qv, dv, items, nitems_dev, max_items, vals);
    }, slot);
  }
  if (p.measure == TS_EUCLIDEAN) {
    return run(e, seed ? "euclidean_refine_seed" : "euclidean_refine", 0, 0, [&] {
      return launch_euclidean(e->st, qv, dv, items, nitems_dev, max_items, vals);
    }, slot);
  }
  // Wavefront DP, one launch per bucket.
  static const char* kNames[3][3] = {{"dtw_wavefront", "frechet_wavefront", "lcss_wavefront"},
                                     {"dtw_wavefront_seed", "frechet_wavefront_seed", "lcss_wavefront_seed"},
                                     {"dtw_wavefront_seed2", "frechet_wavefront_seed2", "lcss_wavefront_seed2"}};
  const char* nm = kNames[kind][p.measure == TS_DTW ? 0 : (p.measure == TS_FRECHET ? 1 : 2)];
  const DpTune* tune = e->tune[dp_table_of(max_items)][dp_variant(p)];
  if (b_lo == b_hi) {
    int sh = tune[b_lo].shape;
    const DpHalo halo = shape_halo(e, sh, b_lo, ma, hi, whi, rmax);
    // launch labels name the per-kernel statistics, kept only while timing
    const std::string label = e->timing ? dp_label(nm, b_lo, sh) : std::string();
    return run(e, e->timing ? label.c_str() : nm, 0, 0, [&] {
      return launch_dp(e->st, ma, sh, qv, dv, items, nitems_dev, nullptr, max_items, ctr, vals, halo, rmax, hi,
                       cus * tune[b_lo].waves_per_cu, tau, tau_q0, cells + slot + b_lo);
    }, slot + b_lo);
  }
  // Buckets run concurrently on their own streams (fork/join by events), so
  // the latency-bound tail of one overlaps the others.
  TS_HIP(hipEventRecord(e->fork, e->st));
  for (int b = b_lo; b <= b_hi; ++b) {
    int sh = tune[b].shape;
    const DpHalo halo = shape_halo(e, sh, b, ma, hi, whi, rmax);
    const std::string label = e->timing ? dp_label(nm, b, sh) : std::string();
    TS_HIP(hipStreamWaitEvent(e->aux[b], e->fork, 0));
    TS_TRY(run_on(e, e->aux[b], e->timing ? label.c_str() : nm, 0, 0, [&] {
      return launch_dp(e->aux[b], ma, sh, qv, dv, itemsb + (int64_t)b * max_items, ctr + kDpBuckets + b,
                       ctr + 2 * kDpBuckets + b, max_items, ctr + b, vals, halo, rmax, hi,
                       cus * tune[b].waves_per_cu, tau, tau_q0, cells + slot + b);
    }, slot + b));
    TS_HIP(hipEventRecord(e->join[b], e->aux[b]));
    TS_HIP(hipStreamWaitEvent(e->st, e->join[b], 0));
  }
  return TS_OK;
}

ts_status fill_nan(ts_engine* e, float* p, int64_t n) {
  if (n == 0) return TS_OK;
  return run(e, "fill", 0, 0, [&] {
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e->st, p,
                       __builtin_nanf(""), n);
    return hipGetLastError();
  });
}

ts_status topk(ts_engine* e, const uint64_t* in, int64_t stride, int32_t len, const int32_t* counts, int32_t nseg,
               int32_t k, uint64_t* out, const int64_t* seg_off = nullptr) {
  int64_t need = select_scratch_elems(len, nseg, k);
  uint64_t *a = nullptr, *b = nullptr;
  TS_TRY(ensure(e, B_SCA, need, &a));
  TS_TRY(ensure(e, B_SCB, need, &b));
  int nl = 0;
  return run(e, "topk_select", (double)nseg * len * 8.0, (double)nseg * len, [&] {
    return select_topk(e->st, in, stride, len, counts, nseg, k, out, a, b, need, &nl, seg_off);
  });
}

// REF pipeline for queries [q0, q0+nq); results into device top keys [nq][k].
ts_status knn_ref_batch(ts_engine* e, const ts_params& p, const ts_set* db, const ts_set* qs, int32_t q0, int32_t nq,
                        const float4* env, const TileView* tiles, uint64_t* top) {
  const int32_t N = db->count;
  const int32_t kc = p.kc;
  const bool desc = desc_measure(p.measure);
  float* scores = nullptr;
  uint64_t *keys = nullptr, *cand = nullptr, *keys2 = nullptr;
  int32_t* cand_idx = nullptr;
  Item* items = nullptr;
  float* vals = nullptr;
  TS_TRY(ensure(e, B_SCORES, (int64_t)nq * N, &scores));
  TS_TRY(ensure(e, B_KEYS, (int64_t)nq * N, &keys));
  TS_TRY(ensure(e, B_CAND, (int64_t)nq * kc, &cand));
  TS_TRY(ensure(e, B_CANDIDX, (int64_t)nq * kc, &cand_idx));
  TS_TRY(ensure(e, B_ITEMS, (int64_t)nq * kc, &items));
  TS_TRY(ensure(e, B_VALS, (int64_t)nq * kc, &vals));
  TS_TRY(ensure(e, B_KEYS2, (int64_t)nq * kc, &keys2));
  const SetView qv = qs->view(), dv = db->view();
  const double dbpts = (double)db->points;
  // algorithmic bytes: the DB's points once per pass of 8 queries (the
  // tiled filter's query group); units: point evaluations
  TS_TRY(run(e, "filter_ref", ((nq + 7) / 8) * (dbpts * 8.0), nq * dbpts, [&] {
    return launch_filter_ref(e->st, p.measure, qv, q0, nq, env, dv, tiles, p.epsilon, scores);
  }));
  TS_TRY(run(e, "keys", 0, 0, [&] { return launch_keys_from_rows(e->st, scores, nq, N, desc, keys); }));
  TS_TRY(topk(e, keys, N, N, nullptr, nq, kc, cand));
  TS_TRY(run(e, "items", 0, 0, [&] { return launch_items_from_keys(e->st, cand, nq, kc, q0, items, cand_idx, vals); }));
  TS_TRY(refine(e, p, db, qs, items, nullptr, (int64_t)nq * kc, vals));
  TS_TRY(run(e, "keys", 0, 0, [&] {
    return launch_keys_from_regions(e->st, vals, cand_idx, nullptr, nq, kc, desc, keys2);
  }));
  return topk(e, keys2, kc, kc, nullptr, nq, p.k, top);
}

// EXACT for small problems: when every pair of the batch fits in the DP
// kernels' concurrent pair slots at once, refining all of them takes one DP
// round, where seeds + filter + refine take two. The result is the exact
// top-k by definition (ties by index, as everywhere).
constexpr int64_t kAllPairsMax = 2048;

ts_status knn_all_batch(ts_engine* e, const ts_params& p, const ts_set* db, const ts_set* qs, int32_t q0,
                        int32_t nq, uint64_t* top) {
  const int32_t N = db->count;
  const int64_t n = (int64_t)nq * N;
  Item* items = nullptr;
  float* vals = nullptr;
  uint64_t* keys = nullptr;
  TS_TRY(ensure(e, B_ITEMS, n, &items));
  TS_TRY(ensure(e, B_VALS, n, &vals));
  TS_TRY(ensure(e, B_KEYS, n, &keys));
  TS_TRY(run(e, "items", 0, 0, [&] { return launch_items_all(e->st, q0, nq, N, items); }));
  TS_TRY(fill_nan(e, vals, n));
  TS_TRY(refine(e, p, db, qs, items, nullptr, n, vals));
  TS_TRY(run(e, "keys", 0, 0, [&] {
    return launch_keys_from_items(e->st, vals, items, n, desc_measure(p.measure), keys);
  }));
  return topk(e, keys, N, N, nullptr, nq, p.k, top);
}

// DP work scale of an EXACT batch of nq queries (for exact_seeds).
double exact_work(const ts_set* qs, const ts_set* db, int64_t nq) {
  const double ql = qs->count ? (double)qs->points / qs->count : 0.0;
  const double dl = db->count ? (double)db->points / db->count : 0.0;
  return (double)nq * db->count * ql * dl;
}

// DP cells of a pair of mean lengths (the row width the DP visits times the
// longer length).
double mean_pair_cells(const ts_set* qs, const ts_set* db, const MeasureArgs& ma) {
  const double ql = qs->count ? (double)qs->points / qs->count : 0.0;
  const double dl = db->count ? (double)db->points / db->count : 0.0;
  const int32_t a = (int32_t)std::min(ql, 2e9), b = (int32_t)std::min(dl, 2e9);
  return (double)dp_width_of(a, b, ma) * std::max(ql, dl);
}

// Second seed round: from this batch work (as exact_work), kSeeds2 or twice
// that candidates a query; the TS_TWO_ROUND environment variable overrides
// the work rule ("0": never, "1": always), and TS_SEEDS2 the count.
// The second round adds a DP launch round to the batch's critical path, so
// it starts where refines are long: 2^44 keeps the serving loop's batches
// of up to ~250 queries against 2M trajectories of 16-512 points on one
// round (from 2^40, p99 at 5k QPS went 21 -> 33 ms), and 4096 queries x 1M
// (2^48.6) or 256 x 20k of 2048-8192 points (2^46.9) on two.
constexpr double kTwoRoundWork = 17592186044416.0;  // 2^44
constexpr int32_t kSeeds2 = 32;

// Seed pool (DTW, Frechet): the bound pass keeps pool x s candidates a
// query, each gets its upper bound (one in-band warping path), and the s of
// smallest upper bound are the seeds; the k-th smallest upper bound of the
// whole pool is their threshold (its k pairs are among the seeds). The
// summary bound ranks long pairs poorly; the path's cost ranks them by
// their actual alignment. Factor 8 for DTW, 4 for Frechet (4096 x 1M: DTW
// 135.2 -> 120.2 ms, Frechet 43.4 -> 40.9 ms; 256 x 20k banded DTW of
// 2048-8192 points 189.7 -> 160.2 ms), at most 256 / s, and only with the
// second seed round (from kTwoRoundWork: smaller batches lose more to its
// launches than they gain, 64 x 100k DTW went 0.74 -> 0.81 ms a step).
// TS_SEED_POOL overrides the factor (1: no pool).
int32_t seed_pool(const ts_params& p, int32_t s, bool two) {
  if (!two || (p.measure != TS_DTW && p.measure != TS_FRECHET)) return 1;
  const char* env = std::getenv("TS_SEED_POOL");
  const int32_t f = env && *env ? std::max(1, std::atoi(env)) : (p.measure == TS_DTW ? 8 : 4);
  return std::max(1, std::min(f, 256 / std::max(1, s)));
}

// Largest batch (queries x DB trajectories) whose seed-pass bounds are kept
// for the compaction (4 bytes a pair).
constexpr int64_t kStoredBoundsMax = 1ll << 26;

// EXACT pipeline for queries [q0, q0+nq).
ts_status knn_exact_batch(ts_engine* e, const ts_params& p, const ts_set* db, const ts_set* qs, int32_t q0,
                          int32_t nq, uint64_t* top, const int32_t* kq_dev) {
  const int32_t N = db->count;
  const int32_t k = p.k;
  const bool desc = desc_measure(p.measure);
  const double work = exact_work(qs, db, nq);
  // Second seed round (DTW and Frechet, batches in the throughput regime,
  // from kTwoRoundWork):
  // after the tight filter, each query's s2 candidates of smallest bound are
  // refined against the first threshold, and the k-th best of both rounds'
  // values becomes the main refine's threshold. The first round's seeds
  // leave the candidates; their values join the final top-k.
  const char* two_env = std::getenv("TS_TWO_ROUND");  // "0" never, "1" always
  const bool two = (p.measure == TS_DTW || p.measure == TS_FRECHET) &&
                   (two_env && *two_env ? two_env[0] == '1' : work >= kTwoRoundWork);
  // Two rounds seed less at first (the second round tightens the
  // threshold): s = max(2k, 32), and s2 = 64 candidates a query for DTW
  // when a mean pair is under 2^20 DP cells, else 32 (long banded pairs:
  // each extra refine costs more than the tighter threshold saves), and 16
  // for Frechet (its max-based values leave little to tighten after the
  // seed pool: 4096 x 1M Frechet 40.5 ms at 32, 38.7 ms at 16).
  const int32_t s = exact_seeds(k, two ? 0.0 : work, two ? 32 : 0);
  const char* s2_env = std::getenv("TS_SEEDS2");
  const int32_t s2 = !two ? 0
                          : std::max(1, std::min(1024, s2_env && *s2_env ? std::atoi(s2_env)
                                                                         : (p.measure == TS_FRECHET ? kSeeds2 / 2
                                                                            : mean_pair_cells(qs, db, margs(p)) <
                                                                                      1048576.0
                                                                                  ? 2 * kSeeds2
                                                                                  : kSeeds2)));
  const int32_t pool = seed_pool(p, s, two), sp = pool * s;  // candidates a query from the bound pass
  // Second-round pool: DTW with short pairs (s2 = 64) picks its s2 from the
  // 2 s2 candidates of smallest bound by their path upper bounds (4096 x 1M
  // DTW 122.0 -> 118.1 ms; Frechet and the long banded pairs of 256 x 20k
  // measured no gain); TS_SEED2_POOL overrides the factor.
  const char* p2_env = std::getenv("TS_SEED2_POOL");
  const int32_t pool2 = !two ? 1
                             : (p2_env && *p2_env ? std::max(1, std::min(16, std::atoi(p2_env)))
                                                  : (p.measure == TS_DTW && s2 == 2 * kSeeds2 ? 2 : 1));
  const int32_t chunk = bound_chunk_items(sp, nq, N);
  const int32_t nch = (N + chunk - 1) / chunk;
  const SetView qv = qs->view(), dv = db->view();
  const MeasureArgs ma = margs(p);
  uint64_t *bkeys = nullptr, *seeds = nullptr;
  int32_t *seed_idx = nullptr, *counts = nullptr;
  int64_t* d_off = nullptr;
  Item* items = nullptr;
  float *vals = nullptr, *tau = nullptr;
  TS_TRY(ensure(e, B_KEYS, (int64_t)nq * nch * sp, &bkeys));
  TS_TRY(ensure(e, B_CAND, (int64_t)nq * s, &seeds));
  TS_TRY(ensure(e, B_CANDIDX, (int64_t)nq * s, &seed_idx));
  TS_TRY(ensure(e, B_TAU, nq, &tau));
  TS_TRY(ensure(e, B_VALS, (int64_t)nq * s, &vals));
  TS_TRY(ensure(e, B_ITEMS, (int64_t)nq * s, &items));
  TS_TRY(ensure(e, B_COUNTS, nq + 2, &counts));
  TS_TRY(ensure(e, B_REGION, 2 * (int64_t)(nq + 1), (int32_t**)&d_off));
  // DP counter sets of the seed refine [0, 32) and the main refine [32, 64),
  // zeroed by the seed pass
  int32_t* ctrs = nullptr;
  TS_TRY(ensure(e, B_CTR, 96, &ctrs));
  const double pairs = (double)nq * N;
  // The seed pass keeps every pair's bound for the compaction's count pass
  // when the batch is small enough; larger batches evaluate them again.
  float* bvals = nullptr;
  if ((int64_t)nq * N <= kStoredBoundsMax) TS_TRY(ensure(e, B_BOUNDS, (int64_t)nq * N, &bvals));
  // 1. seeds: good candidates by bound, refined exactly
  TS_TRY(run(e, "bound_seed", pairs * 36.0, pairs, [&] {
    return launch_bound_chunks(e->st, ma, k, qv, q0, nq, dv, sp, chunk, bkeys, bvals, ctrs, 96);
  }));
  float* tau_ub = nullptr;
  if (pool > 1) {
    // the pool: sp candidates a query by bound, their upper bounds, the
    // threshold from all of them, and the s of smallest upper bound
    uint64_t *pk = nullptr, *pkeys = nullptr;
    int32_t* pidx = nullptr;
    float* ub = nullptr;
    const int64_t np = (int64_t)nq * sp;
    TS_TRY(ensure(e, B_POOL, np, &pk));
    TS_TRY(ensure(e, B_POOLK, np, &pkeys));
    TS_TRY(ensure(e, B_POOLIDX, np, &pidx));
    TS_TRY(ensure(e, B_UB, np, &ub));
    TS_TRY(ensure(e, B_TAUUB, nq, &tau_ub));
    TS_TRY(topk(e, bkeys, (int64_t)nch * sp, nch * sp, nullptr, nq, sp, pk));
    TS_TRY(run(e, "seed_ub", 0, 0, [&] {
      hipLaunchKernelGGL(k_seed_ub, dim3((unsigned)((np + 3) / 4)), dim3(256), 0, e->st, pk, sp, q0, np, qv, dv,
                         p.measure, pidx, ub);
      return hipGetLastError();
    }));
    TS_TRY(run(e, "threshold", 0, 0, [&] {
      return launch_seed_tau(e->st, ub, pidx, nq, sp, k, 0, tau_ub, kq_dev ? kq_dev + q0 : nullptr);
    }));
    TS_TRY(run(e, "keys", 0, 0, [&] {
      hipLaunchKernelGGL(k_ub_keys, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, e->st, ub, pidx, np, pkeys);
      return hipGetLastError();
    }));
    TS_TRY(topk(e, pkeys, sp, sp, nullptr, nq, s, seeds));
  } else {
    TS_TRY(topk(e, bkeys, (int64_t)nch * s, nch * s, nullptr, nq, s, seeds));
  }
  const KeyItems seed_items{seeds, s, q0, seed_idx};
  // The seeds stop early against the k-th smallest upper bound of their
  // values (DTW, Frechet): a seed above it cannot be among the k best seeds,
  // whose own bounds keep k of them below it, so the threshold the seeds
  // give is unchanged.
  if (pool == 1 && (p.measure == TS_DTW || p.measure == TS_FRECHET)) {
    float* ub = nullptr;
    TS_TRY(ensure(e, B_UB, (int64_t)nq * s, &ub));
    TS_TRY(ensure(e, B_TAUUB, nq, &tau_ub));
    TS_TRY(run(e, "seed_ub", 0, 0, [&] {
      const int64_t ns = (int64_t)nq * s;
      hipLaunchKernelGGL(k_seed_ub, dim3((unsigned)((ns + 3) / 4)), dim3(256), 0, e->st, seeds, s, q0, ns, qv, dv,
                         p.measure, seed_idx, ub);
      return hipGetLastError();
    }));
    TS_TRY(run(e, "threshold", 0, 0, [&] {
      return launch_seed_tau(e->st, ub, seed_idx, nq, s, k, 0, tau_ub, kq_dev ? kq_dev + q0 : nullptr);
    }));
  }
  TS_TRY(refine(e, p, db, qs, items, nullptr, (int64_t)nq * s, vals, 1, ctrs, &seed_items, nullptr, tau_ub, q0));
  // 2. threshold: k-th best exact value among the seeds
  TS_TRY(run(e, "threshold", 0, 0, [&] {
    return launch_seed_tau(e->st, vals, seed_idx, nq, s, k, desc, tau, kq_dev ? kq_dev + q0 : nullptr);
  }));
  // 3. compaction of every pair that can still reach tau: count, size each
  // query's region, write. Eager calls size the regions exactly from the
  // counts read back to the host; graph-captured calls cannot read back, so
  // the device sizes them within the engine's candidate capacity and the
  // call is repeated with a larger capacity if the total exceeded it.
  const int32_t gx = compact_blocks(N);
  int32_t* bcounts = nullptr;
  uint64_t* bits = nullptr;
  TS_TRY(ensure(e, B_BCNT, (int64_t)nq * gx, &bcounts));
  TS_TRY(ensure(e, B_BITS, (int64_t)nq * compact_mask_words(N), &bits));
  TS_TRY(run(e, "bound_compact", pairs * 36.0, pairs, [&] {
    return launch_bound_compact(e->st, ma, k, qv, q0, nq, dv, tau, nullptr, bvals, bcounts, bits, nullptr, 0);
  }));
  TS_TRY(run(e, "block_prefix", 0, 0, [&] { return launch_block_prefix(e->st, bcounts, nq, gx, counts); }));
  const bool dev_sized = e->capturing;
  int64_t total = 0;  // candidates written (eager) or the capacity (device-sized)
  int64_t cap = 0;    // length of the candidate buffers
  int32_t seg_len = 1;
  int32_t *segc = nullptr, *ntot = nullptr;
  if (dev_sized) {
    total = cap = e->exact_cap;
    // a region holds at most N candidates, and at most the capacity (the
    // scan clamps region lengths to it)
    seg_len = (int32_t)std::min<int64_t>(N, cap);
    TS_TRY(ensure(e, B_SEGC, nq + 2, &segc));
    ntot = segc + nq;
    TS_TRY(run(e, "scan", 0, 0, [&] {
      return launch_scan_offsets(e->st, counts, nq, cap, d_off, ntot, segc, e->need_dev);
    }));
  } else {
    std::vector<int32_t>& hc = e->h_counts;
    std::vector<int64_t>& hoff = e->h_off;
    hc.resize(nq);
    hoff.resize(nq + 1);
    TS_HIP(hipMemcpyAsync(hc.data(), counts, sizeof(int32_t) * nq, hipMemcpyDeviceToHost, e->st));
    TS_HIP(hipStreamSynchronize(e->st));
    hoff[0] = 0;
    for (int32_t i = 0; i < nq; ++i) {
      hoff[i + 1] = hoff[i] + hc[i];
      seg_len = std::max(seg_len, hc[i]);
    }
    total = hoff[nq];
    if (total >= (1ll << 31)) return fail(TS_ERR_UNSUPPORTED, "more than 2^31 candidate pairs in one batch");
    cap = total;
    if (e->graphs) {
      // Size everything as a recorded replay of this batch will need it, so
      // that recording allocates nothing.
      e->exact_cap = std::max<int64_t>(e->exact_cap, std::min<int64_t>(total + total / 4 + 4096, (1ll << 31) - 1));
      cap = e->exact_cap;
      TS_TRY(ensure(e, B_SEGC, nq + 2, &segc));
      // the selections over each query's region: k, and s2 (times its pool)
      // with two rounds
      const int64_t need = select_scratch_elems(N, nq, std::max(k, s2 * pool2));
      uint64_t* sc = nullptr;
      TS_TRY(ensure(e, B_SCA, need, &sc));
      TS_TRY(ensure(e, B_SCB, need, &sc));
    }
    TS_HIP(hipMemcpyAsync(d_off, hoff.data(), sizeof(int64_t) * (nq + 1), hipMemcpyHostToDevice, e->st));
  }
  Item *citems = nullptr, *items2 = nullptr;
  float* rvals = nullptr;
  uint64_t* rkeys = nullptr;
  TS_TRY(ensure(e, B_ITEMS2, cap, &citems));
  TS_TRY(ensure(e, B_RITEMS, cap, &items2));
  TS_TRY(ensure(e, B_RVALS, cap, &rvals));
  TS_TRY(ensure(e, B_SCORES, 2 * cap, (float**)&rkeys));
  TS_TRY(run(e, "compact_write", 0, 0, [&] {
    return launch_bound_compact(e->st, ma, k, qv, q0, nq, dv, tau, d_off, nullptr, bcounts, bits, citems, cap);
  }));
  // 3b. tighter O(n+m) bound on the survivors (with two rounds: every
  // survivor's bound into tb, the first round's seeds rejected)
  int32_t* nitems2 = counts + nq;
  float* tb = nullptr;
  if (two) TS_TRY(ensure(e, B_TB, cap, &tb));
  // Unbanded DTW and Frechet: a lane-per-pair pre-check on superblock boxes
  // first, the tight filter on what it keeps. Units: the pairs each tests.
  const Item* tight_in = citems;
  const int32_t* tight_n = ntot;
  if ((p.measure == TS_DTW && p.band == 0) || p.measure == TS_FRECHET) {
    Item* pre = nullptr;
    TS_TRY(ensure(e, B_PRE, cap, &pre));
    int32_t* npre = counts + nq + 1;
    TS_TRY(run(e, "tight_pre", 0, (double)total, [&] {
      return launch_tight_pre(e->st, ma, qv, q0, dv, tau, citems, ntot, total, pre, npre, rvals, tb);
    }));
    tight_in = pre;
    tight_n = npre;
  }
  TS_TRY(run(e, "tight_filter", 0, (double)total, [&] {
    return launch_tight_filter(e->st, ma, k, qv, q0, dv, tau, tight_in, tight_n, total, items2, nitems2, rvals,
                               device_cu_count() * 32, tb, two ? seed_idx : nullptr, s, tight_in != citems);
  }));
  const int32_t* seg_counts = dev_sized ? segc : counts;
  if (two) {
    // 3c. second seed round: each query's s2 candidates of smallest bound
    uint64_t* top2 = nullptr;
    int32_t* seed_idx2 = nullptr;
    Item* items5 = nullptr;
    TS_TRY(ensure(e, B_TOP2, (int64_t)nq * s2, &top2));
    TS_TRY(ensure(e, B_CANDIDX2, (int64_t)nq * s2, &seed_idx2));
    TS_TRY(ensure(e, B_ITEMS5, (int64_t)nq * s2, &items5));
    TS_TRY(run(e, "keys", 0, 0, [&] { return launch_keys_from_slots(e->st, tb, citems, cap, rkeys); }));
    if (pool2 > 1) {
      // pool2 x s2 candidates by bound, then the s2 of smallest path upper bound
      const int32_t s2p = pool2 * s2;
      uint64_t *top2p = nullptr, *keys2p = nullptr;
      TS_TRY(ensure(e, B_TOP2P, (int64_t)nq * s2p, &top2p));
      TS_TRY(ensure(e, B_KEYS2P, (int64_t)nq * s2p, &keys2p));
      TS_TRY(topk(e, rkeys, 0, seg_len, seg_counts, nq, s2p, top2p, d_off));
      TS_TRY(run(e, "seed_ub", 0, 0, [&] {
        const int64_t n2 = (int64_t)nq * s2p;
        hipLaunchKernelGGL(k_slot_ub_keys, dim3((unsigned)((n2 + 3) / 4)), dim3(256), 0, e->st, top2p, n2, citems, qv,
                           dv, p.measure, keys2p);
        return hipGetLastError();
      }));
      TS_TRY(topk(e, keys2p, s2p, s2p, nullptr, nq, s2, top2));
    } else {
      TS_TRY(topk(e, rkeys, 0, seg_len, seg_counts, nq, s2, top2, d_off));
    }
    const KeyItems r2_items{top2, s2, q0, seed_idx2, citems, tb};
    TS_TRY(refine(e, p, db, qs, items5, nullptr, (int64_t)nq * s2, rvals, 2, ctrs + 64, &r2_items, nullptr, tau, q0));
    TS_TRY(run(e, "threshold", 0, 0, [&] {
      return launch_tau2(e->st, vals, seed_idx, nq, s, top2, s2, rvals, citems, k, tau, kq_dev ? kq_dev + q0 : nullptr);
    }));
  }
  // the main refine abandons pairs whose value must exceed the threshold
  // (with two rounds it skips the second round's pairs and those whose
  // bound exceeds the new threshold)
  const Prune prune{tb, tau};
  TS_TRY(refine(e, p, db, qs, items2, nitems2, cap, rvals, 0, ctrs + 32, nullptr, nullptr, tau, q0,
                two ? &prune : nullptr));
  // 4. top-k over each query's refined candidates (slot = position in citems)
  TS_TRY(run(e, "keys", 0, 0, [&] { return launch_keys_from_items(e->st, rvals, citems, cap, desc, rkeys); }));
  // Region lengths: clamped to the capacity (device-sized), or the counted
  // totals.
  if (!two) return topk(e, rkeys, 0, seg_len, seg_counts, nq, k, top, d_off);
  // with two rounds, then over the candidates' top-k and the seeds
  uint64_t *topa = nullptr, *mkeys = nullptr;
  TS_TRY(ensure(e, B_TOPA, (int64_t)nq * k, &topa));
  TS_TRY(ensure(e, B_MKEYS, (int64_t)nq * (k + s), &mkeys));
  TS_TRY(topk(e, rkeys, 0, seg_len, seg_counts, nq, k, topa, d_off));
  TS_TRY(run(e, "merge_seeds", 0, 0, [&] { return launch_merge_seed_keys(e->st, topa, k, vals, seed_idx, s, nq, mkeys); }));
  return topk(e, mkeys, k + s, k + s, nullptr, nq, k, top);
}

// Run one batch through the graph cache: replay a recorded graph whose key
// matches; run a key's first occurrence eagerly (it sizes the workspace);
// record on the second and replay from then on. A recording that fails
// (workspace growth, capture error) runs the batch eagerly instead.
constexpr size_t kGraphCacheMax = 16;
constexpr int64_t kGraphMaxPairs = 1ll << 26;

template <class F>
ts_status run_graph(ts_engine* e, const std::vector<int64_t>& base, F body) {
  const int c_entry = e->cells_used;
  auto full_key = [&] {
    std::vector<int64_t> k = base;
    k.push_back((int64_t)e->ws_gen);
    k.push_back(e->exact_cap);
    k.push_back(c_entry);
    return k;
  };
  const std::vector<int64_t> key = full_key();
  for (GraphEntry& g : e->gcache) {
    if (g.key != key) continue;
    TS_HIP(hipGraphLaunch(g.exec, e->st));
    ++e->graph_replays;
    e->launches += g.launches;
    e->cells_used = g.cells_to;
    g.last_use = ++e->use_clock;
    return TS_OK;
  }
  if (std::find(e->gseen.begin(), e->gseen.end(), key) == e->gseen.end()) {
    // First occurrence: run eagerly, and remember the key as it stands
    // afterwards (the run may have grown the workspace or the capacity).
    TS_TRY(body());
    if (e->gseen.size() >= 4 * kGraphCacheMax) e->gseen.erase(e->gseen.begin());
    e->gseen.push_back(full_key());
    return TS_OK;
  }
  const int c0 = e->cells_used;
  const int64_t l0 = e->launches;
  TS_HIP(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
  e->capturing = true;
  e->capture_broken = false;
  const ts_status st = body();
  e->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t err = hipStreamEndCapture(e->st, &g);
  hipGraphExec_t ex = nullptr;
  bool ok = st == TS_OK && err == hipSuccess && !e->capture_broken && g != nullptr;
  if (ok) ok = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess;
  if (g) (void)hipGraphDestroy(g);
  if (!ok) {
    (void)hipGetLastError();
    e->cells_used = c0;
    e->launches = l0;
    return body();
  }
  if (e->gcache.size() >= kGraphCacheMax) {
    auto lru = std::min_element(e->gcache.begin(), e->gcache.end(),
                                [](const GraphEntry& a, const GraphEntry& b) { return a.last_use < b.last_use; });
    (void)hipGraphExecDestroy(lru->exec);
    e->gcache.erase(lru);
  }
  e->gcache.push_back(GraphEntry{key, ex, e->launches - l0, e->cells_used, ++e->use_clock});
  ++e->graph_captures;
  TS_HIP(hipGraphLaunch(ex, e->st));
  return TS_OK;
}

std::vector<int64_t> set_key(const ts_set* s) {
  return {(int64_t)(intptr_t)s, s->count, s->min_len, s->max_len, s->dev_points, (int64_t)(intptr_t)s->x,
          (int64_t)(intptr_t)s->off};
}

// Waits for a k-NN call's stream by polling it: the call's host thread has
// nothing else to do, and a poll sees the end of the last copy sooner than a
// blocking wait's wake-up.
ts_status wait_stream(hipStream_t st) {
  for (;;) {
    const hipError_t r = hipStreamQuery(st);
    if (r == hipSuccess) return TS_OK;
    if (r != hipErrorNotReady) {
      TS_HIP(r);
    }
  }
}

// kq (host, optional): per-query k, each in [1, pp->k]. Every stage then
// runs at k = pp->k except the EXACT threshold, which takes query q's
// kq[q]-th best seed: the first kq[q] results of query q are its exact top
// kq[q] (the top-k_max of a candidate set that holds every pair better than
// the kq[q]-th begins with them); its later slots hold idx -1.
ts_status knn_impl(ts_engine* e, const ts_set* db, const ts_set* qs, const ts_params* pp, int32_t* out_idx,
                   float* out_val, ts_stats* stats, bool plain, int retry = 0, const int32_t* kq = nullptr) {
  auto t0 = std::chrono::steady_clock::now();
  if (!e) return fail(TS_ERR_INVALID, "null engine");
  TS_TRY(check_params(pp, db, qs));
  if (!out_idx || !out_val) return fail(TS_ERR_INVALID, "null output");
  const ts_params p = *pp;
  const int32_t Q = qs->count, N = db->count, k = p.k;
  const bool desc = desc_measure(p.measure);
  const int64_t l0 = e->launches;
  e->cells_used = 0;
  float4* env = nullptr;
  TileView tiles{};
  const bool point_filter =
      p.mode == TS_MODE_REF && (p.measure == TS_DTW || p.measure == TS_FRECHET || p.measure == TS_LCSS);
  if (point_filter) {
    TS_TRY(ensure(e, B_ENV, qs->dev_points, &env));
    // the DB's length-bucketed tiles, (re)built here, outside any recording
    TS_TRY(set_tiles(db, e->st, &tiles));
  }
  // Results block, laid out as the host staging buffer: indices, values
  // (padded to 16 bytes), then a copy of the cell / pair counters and the
  // candidate total, so one copy brings everything back.
  const int64_t res_words = ((int64_t)Q * k * 2 + 3) / 4 * 4;
  int32_t* d_out = nullptr;
  uint64_t* top = nullptr;
  TS_TRY(ensure(e, B_OIDX, res_words + 2 * kCellBuf, &d_out));
  int32_t* d_idx = d_out;
  float* d_val = (float*)(d_out + (int64_t)Q * k);
  unsigned long long* d_tail = (unsigned long long*)(d_out + res_words);
  TS_TRY(ensure(e, B_HOSTIN, (int64_t)Q * k, &top));
  unsigned long long* cells = nullptr;
  TS_TRY(ensure(e, B_CELLS, kCellBuf, &cells));
  e->need_dev = cells + 2 * kCellSlots;
  if (!e->call_a) {
    TS_HIP(hipEventCreate(&e->call_a));
    TS_HIP(hipEventCreate(&e->call_b));
  }
  // Queries per batch: bounded by the dense REF score matrix, or by the
  // EXACT seed pass's kept keys (8 bytes each, DB chunks x seed candidates a
  // query) within 2 GiB.
  int32_t bq;
  if (plain) {
    bq = 1;
  } else if (p.mode == TS_MODE_REF) {
    bq = (int32_t)std::max<int64_t>(1, std::min<int64_t>(Q, (int64_t)(256ll << 20) / (8ll * std::max(1, N))));
  } else {
    int64_t s = exact_seeds(k, exact_work(qs, db, Q));
    s *= seed_pool(p, (int32_t)s, true);  // at most this many
    int64_t nch = (N + bound_chunk_items((int32_t)s) - 1) / bound_chunk_items((int32_t)s);
    bq = (int32_t)std::max<int64_t>(1, std::min<int64_t>({(int64_t)Q, 4096, (int64_t)(2048ll << 20) / (8 * nch * s)}));
  }
  // Small EXACT problems refine every pair (Euclidean keeps its filter,
  // which also drops pairs of unequal length).
  const bool all_pairs = p.mode == TS_MODE_EXACT && p.measure != TS_EUCLIDEAN && (int64_t)bq * N <= kAllPairsMax;
  // Graph replay pays where launch overhead is a visible share of a batch:
  // at most kGraphMaxPairs pairs. Above that (measured on 4096 x 1M DTW:
  // replays 1.28-1.48 s against 1.25 s eager, with the DP buckets' parallel
  // branches) batches launch eagerly.
  const bool use_graphs =
      e->graphs && !e->timing && !plain && !kq && (e->graphs_all || (int64_t)bq * N <= kGraphMaxPairs);
  int32_t* kq_dev = nullptr;
  if (kq) {
    for (int32_t i = 0; i < Q; ++i)
      if (kq[i] < 1 || kq[i] > k) return fail(TS_ERR_INVALID, "per-query k must be in [1, k]");
    TS_TRY(ensure(e, B_KQ, Q, &kq_dev));
    TS_HIP(hipMemcpyAsync(kq_dev, kq, sizeof(int32_t) * Q, hipMemcpyHostToDevice, e->st));
  }
  // Results, the capacity check and the cell counters come back through the
  // pinned staging buffer in one copy behind the last kernel.
  const bool check_need = use_graphs && p.mode == TS_MODE_EXACT && !all_pairs;
  const bool tail = check_need || stats != nullptr;
  const size_t res_bytes = (size_t)res_words * 4;
  const size_t stage_bytes = res_bytes + sizeof(unsigned long long) * (2 * kCellSlots + 1);
  if (e->h_stage_bytes < stage_bytes) {
    if (e->h_stage) TS_HIP(hipHostFree(e->h_stage));
    e->h_stage = nullptr;
    e->h_stage_bytes = 0;
    TS_HIP(hipHostMalloc((void**)&e->h_stage, stage_bytes, hipHostMallocDefault));
    e->h_stage_bytes = stage_bytes;
  }
  auto batch = [&](int32_t q0, int32_t nq) -> ts_status {
    uint64_t* topq = top + (int64_t)q0 * k;
    if (p.mode == TS_MODE_REF) return knn_ref_batch(e, p, db, qs, q0, nq, env, &tiles, topq);
    if (all_pairs) return knn_all_batch(e, p, db, qs, q0, nq, topq);
    return knn_exact_batch(e, p, db, qs, q0, nq, topq, kq_dev);
  };
  auto prologue = [&]() -> ts_status {
    TS_TRY(run(e, "fill", 0, 0, [&] { return zero_cells(e->st, cells); }));
    if (env) {
      TS_TRY(run(e, "envelope", 0, 0, [&] { return launch_envelope(e->st, qs->view(), qs->max_len, p.r, env); }));
    }
    return TS_OK;
  };
  auto epilogue = [&]() -> ts_status {
    TS_TRY(run(e, "decode", 0, 0, [&] {
      return launch_decode_keys(e->st, top, (int64_t)Q * k, desc, d_idx, d_val, tail ? cells : nullptr, d_tail,
                                kCellBuf);
    }));
    TS_HIP(hipMemcpyAsync(e->h_stage, d_out, tail ? stage_bytes : res_bytes, hipMemcpyDeviceToHost, e->st));
    return TS_OK;
  };
  auto graph_key = [&](int32_t q0, int32_t nq) {
    std::vector<int64_t> key = set_key(db);
    const std::vector<int64_t> qk = set_key(qs);
    key.insert(key.end(), qk.begin(), qk.end());
    int32_t eps_bits;
    std::memcpy(&eps_bits, &p.epsilon, 4);
    key.insert(key.end(), {p.measure, p.mode, p.k, p.kc, p.r, p.band, p.delta, eps_bits, q0, nq,
                           (int64_t)(intptr_t)env, (int64_t)(intptr_t)(top + (int64_t)q0 * k),
                           (int64_t)(intptr_t)tiles.x, (int64_t)(intptr_t)tiles.off, tiles.ntiles});
    return key;
  };
  TS_HIP(hipEventRecord(e->call_a, e->st));
  if (use_graphs && Q <= bq) {
    // One batch: the whole call (counter clear, batch, decode, copy) is one
    // recorded graph.
    std::vector<int64_t> key = graph_key(0, Q);
    key.insert(key.end(), {-1, (int64_t)(intptr_t)d_out, (int64_t)(intptr_t)e->h_stage, tail ? 1 : 0});
    TS_TRY(run_graph(e, key, [&]() -> ts_status {
      TS_TRY(prologue());
      TS_TRY(batch(0, Q));
      return epilogue();
    }));
  } else {
    TS_TRY(prologue());
    for (int32_t q0 = 0; q0 < Q; q0 += bq) {
      const int32_t nq = std::min(bq, Q - q0);
      if (use_graphs) {
        TS_TRY(run_graph(e, graph_key(q0, nq), [&] { return batch(q0, nq); }));
      } else {
        TS_TRY(batch(q0, nq));
      }
    }
    TS_TRY(epilogue());
  }
  TS_HIP(hipEventRecord(e->call_b, e->st));
  const int used = stats ? std::min(e->cells_used, kCellSlots) : 0;
  int32_t* h_idx = (int32_t*)e->h_stage;
  float* h_val = (float*)(e->h_stage + (size_t)Q * k * 4);
  unsigned long long* h_cells = (unsigned long long*)(e->h_stage + res_bytes);
  unsigned long long* h_need = h_cells + 2 * kCellSlots;
  TS_TRY(wait_stream(e->st));
  std::memcpy(out_idx, h_idx, sizeof(int32_t) * Q * k);
  std::memcpy(out_val, h_val, sizeof(float) * Q * k);
  if (kq) {
    for (int32_t i = 0; i < Q; ++i)
      for (int32_t t = kq[i]; t < k; ++t) {
        out_idx[(int64_t)i * k + t] = -1;
        out_val[(int64_t)i * k + t] = 0.0f;
      }
  }
  if (check_need) {
    // A replayed batch sized its candidates on the device: if any total
    // exceeded the capacity its results are incomplete; grow and repeat.
    const unsigned long long need = *h_need;
    if ((int64_t)need > e->exact_cap) {
      if (retry > 0) return fail(TS_ERR_HIP, "candidate capacity still exceeded after growth");
      e->exact_cap = std::min<int64_t>((int64_t)need + (int64_t)need / 4 + 4096, (1ll << 31) - 1);
      ++e->overflow_retries;
      return knn_impl(e, db, qs, pp, out_idx, out_val, stats, plain, retry + 1, kq);
    }
  }
  if (stats) {
    std::vector<unsigned long long> hc(2 * kCellSlots, 0ull);
    for (int i = 0; i < used; ++i) {
      hc[i] = h_cells[i];
      hc[kCellSlots + i] = h_cells[kCellSlots + i];
    }
    float dms = 0;
    TS_HIP(hipEventElapsedTime(&dms, e->call_a, e->call_b));
    stats->ms_device = dms;
    stats->pairs = (int64_t)Q * N;
    stats->cells = 0;
    stats->refined_pairs = 0;
    for (int i = 0; i < e->cells_used && i < kCellSlots; ++i) {
      stats->cells += (int64_t)hc[i];
      stats->refined_pairs += (int64_t)hc[kCellSlots + i];
    }
    stats->launches = e->launches - l0;
    stats->overflow_retries = e->overflow_retries;
  }
  TS_TRY(resolve_timing(e));
  if (stats) stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return TS_OK;
}

}  // namespace

ts_status refine_fixed(ts_engine* e, const ts_params& p, const ts_set* db, const ts_set* qs, Item* items,
                       int64_t n, float* vals, const RefineBufs& bufs) {
  TS_TRY(check_params(&p, db, qs));
  const bool timing = e->timing;
  e->timing = false;
  e->cells_used = 0;
  const ts_status s = refine(e, p, db, qs, items, nullptr, n, vals, false, nullptr, nullptr, &bufs);
  e->timing = timing;
  return s;
}

ts_status validate_params(const ts_params* p, const ts_set* db, const ts_set* qs) { return check_params(p, db, qs); }

hipStream_t engine_stream(const ts_engine* e) { return e->st; }
int engine_device(const ts_engine* e) { return e->device; }

}  // namespace ts

extern "C" {

namespace ts {
namespace {

const char* const kVariantNames[kDpVariants] = {"dtw", "dtw_band", "frechet", "lcss", "lcss_window"};

int shape_of(int W, int U);

void tune_defaults(ts_engine* e) {
  const char* v = std::getenv("TS_DP_WAVES_PER_CU");
  int w = v ? std::atoi(v) : 0;
  if (w <= 0) w = 8;  // best of a 2..32 sweep on 64 x 100k DTW
  for (int m = 0; m < kTables; ++m)
    for (int var = 0; var < kDpVariants; ++var)
      for (int b = 0; b < kDpBuckets; ++b) e->tune[m][var][b] = DpTune{kDpDefaultShape[b], w};
  // few pairs of rows wider than 1023: four waves share each (one pair of
  // 8192 x 8192 DTW 60 -> 20 ms); LCSS windows stay on one wave
  for (int var = 0; var < kDpVariants; ++var)
    if (var != 4) e->tune[3][var][kDpBuckets - 1] = DpTune{shape_of(256, var == 1 ? 4 : 8), 4};
}

// The API's code of internal table m (2 means every table there).
int table_code(int m) { return m == 2 ? 3 : (m == 3 ? 4 : m); }

int shape_of(int W, int U) {
  for (int i = 0; i < kDpShapes; ++i)
    if (kDpShapeList[i].W == W && kDpShapeList[i].U == U) return i;
  return -1;
}

// Tuning table: text, one line per (variant, bucket[, table]),
//   <variant> <bucket> <W> <U> <waves_per_cu> [lat|mid|few]
// variant one of dtw, dtw_band, frechet, lcss, lcss_window; a trailing
// "lat" fills the small-refine table, "mid" the mid-size one, "few" the
// few-pair one, nothing the throughput table (and the others where no line
// of theirs came first); '#' comments. Lines not listed keep their current
// values.
ts_status load_tuning(ts_engine* e, const char* path) {
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(TS_ERR_INVALID, std::string("cannot open tuning table ") + path);
  char line[1024];
  int lineno = 0;
  ts_status st = TS_OK;
  bool set[kTables][kDpVariants][kDpBuckets] = {};
  while (std::fgets(line, sizeof line, f)) {
    ++lineno;
    char* hash = std::strchr(line, '#');
    if (hash) *hash = 0;
    char name[32], mode[8] = "";
    int b, W, U, waves;
    const int n = std::sscanf(line, "%31s %d %d %d %d %7s", name, &b, &W, &U, &waves, mode);
    if (n <= 0) continue;
    int var = -1;
    for (int i = 0; i < kDpVariants; ++i)
      if (n >= 5 && std::strcmp(name, kVariantNames[i]) == 0) var = i;
    const int sh = shape_of(W, U);
    const bool lat = n == 6 && std::strcmp(mode, "lat") == 0;
    const bool mid = n == 6 && std::strcmp(mode, "mid") == 0;
    const bool few = n == 6 && std::strcmp(mode, "few") == 0;
    if (n < 5 || (n == 6 && !lat && !mid && !few) || var < 0 || b < 0 || b >= kDpBuckets || sh < 0 ||
        waves < 1 || waves > 64) {
      st = fail(TS_ERR_INVALID, std::string(path) + ":" + std::to_string(lineno) + ": bad tuning line");
      break;
    }
    if (lat || mid || few) {
      const int t = lat ? 1 : (mid ? 2 : 3);
      e->tune[t][var][b] = DpTune{sh, waves};
      set[t][var][b] = true;
    } else {
      e->tune[0][var][b] = DpTune{sh, waves};
      for (int t = 1; t < kTables; ++t)
        if (!set[t][var][b]) e->tune[t][var][b] = DpTune{sh, waves};
    }
  }
  std::fclose(f);
  if (st == TS_OK) e->tuning_source = path;
  return st;
}

// The table beside the library (dp_tuning.txt), or TS_DP_TUNING; absent is
// fine (defaults), unreadable is an error.
ts_status load_default_tuning(ts_engine* e) {
  const char* env = std::getenv("TS_DP_TUNING");
  if (env && *env) return load_tuning(e, env);
  Dl_info info;
  if (dladdr((void*)&tune_defaults, &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    const size_t slash = p.rfind('/');
    p = (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/dp_tuning.txt";
    FILE* f = std::fopen(p.c_str(), "r");
    if (f) {
      std::fclose(f);
      return load_tuning(e, p.c_str());
    }
  }
  return TS_OK;
}

}  // namespace
}  // namespace ts

ts_status ts_engine_create(ts_engine** out) {
  if (!out) return fail(TS_ERR_INVALID, "null out");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(TS_ERR_NO_DEVICE, "no HIP device available");
  auto e = new ts_engine();
  (void)hipGetDevice(&e->device);
  tune_defaults(e);
  if (load_default_tuning(e) != TS_OK) {
    delete e;
    return TS_ERR_INVALID;
  }
  hipError_t err = hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking);
  for (int b = 0; b < kDpBuckets && err == hipSuccess; ++b) {
    err = hipStreamCreateWithFlags(&e->aux[b], hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&e->join[b], hipEventDisableTiming);
  }
  if (err == hipSuccess) err = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
  if (err != hipSuccess) {
    ts_engine_destroy(e);
    return hip_fail(err, "hipStreamCreate", __FILE__, __LINE__);
  }
  *out = e;
  return TS_OK;
}

ts_status ts_engine_destroy(ts_engine* e) {
  if (!e) return TS_OK;
  if (e->st) (void)hipStreamSynchronize(e->st);
  for (GraphEntry& g : e->gcache) (void)hipGraphExecDestroy(g.exec);
  for (auto& b : e->bufs)
    if (b.p) (void)hipFree(b.p);
  if (e->h_stage) (void)hipHostFree(e->h_stage);
  for (auto ev : e->evpool) (void)hipEventDestroy(ev);
  if (e->call_a) (void)hipEventDestroy(e->call_a);
  if (e->call_b) (void)hipEventDestroy(e->call_b);
  for (int b = 0; b < kDpBuckets; ++b) {
    if (e->aux[b]) (void)hipStreamDestroy(e->aux[b]);
    if (e->join[b]) (void)hipEventDestroy(e->join[b]);
  }
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->st) (void)hipStreamDestroy(e->st);
  delete e;
  return TS_OK;
}

ts_status ts_engine_set_timing(ts_engine* e, int32_t on) {
  if (!e) return fail(TS_ERR_INVALID, "null engine");
  e->timing = on != 0;
  return TS_OK;
}

ts_status ts_engine_load_tuning(ts_engine* e, const char* path) {
  if (!e || !path) return fail(TS_ERR_INVALID, "null argument");
  return load_tuning(e, path);
}

ts_status ts_engine_set_dp_shape(ts_engine* e, int32_t table, int32_t variant, int32_t bucket, int32_t lanes,
                                 int32_t cells, int32_t waves_per_cu) {
  if (!e) return fail(TS_ERR_INVALID, "null engine");
  const int sh = shape_of(lanes, cells);
  if (table < 0 || table > 4 || variant < 0 || variant >= kDpVariants || bucket < 0 || bucket >= kDpBuckets ||
      sh < 0 || waves_per_cu < 1 || waves_per_cu > 64)
    return fail(TS_ERR_INVALID, "ts_engine_set_dp_shape arguments");
  // table 0: throughput, 1: small, 2: every table, 3: mid-size, 4: few
  for (int m = 0; m < kTables; ++m)
    if (table == 2 || table == table_code(m)) e->tune[m][variant][bucket] = DpTune{sh, waves_per_cu};
  return TS_OK;
}

ts_status ts_engine_get_dp_shape(ts_engine* e, int32_t table, int32_t variant, int32_t bucket, int32_t* lanes,
                                 int32_t* cells, int32_t* waves_per_cu) {
  if (!e || table < 0 || table > 4 || table == 2 || variant < 0 || variant >= kDpVariants || bucket < 0 ||
      bucket >= kDpBuckets)
    return fail(TS_ERR_INVALID, "ts_engine_get_dp_shape arguments");
  const DpTune& t = e->tune[table == 3 ? 2 : (table == 4 ? 3 : table)][variant][bucket];
  if (lanes) *lanes = kDpShapeList[t.shape].W;
  if (cells) *cells = kDpShapeList[t.shape].U;
  if (waves_per_cu) *waves_per_cu = t.waves_per_cu;
  return TS_OK;
}

int32_t ts_dp_shape_count(void) { return kDpShapes; }

int64_t ts_dp_cells(const ts_params* p, int32_t n, int32_t m) {
  if (!p || n < 1 || m < 1) return -1;
  return (int64_t)pair_cells(n, m, p->measure, p->band, p->delta);
}

ts_status ts_engine_dp_fallbacks(ts_engine* e, int64_t* n) {
  if (!e || !n) return fail(TS_ERR_INVALID, "null argument");
  *n = e->dp_fallbacks;
  return TS_OK;
}

ts_status ts_dp_shape(int32_t i, int32_t* lanes, int32_t* cells) {
  if (i < 0 || i >= kDpShapes) return TS_ERR_INVALID;
  if (lanes) *lanes = kDpShapeList[i].W;
  if (cells) *cells = kDpShapeList[i].U;
  return TS_OK;
}

ts_status ts_engine_set_graphs(ts_engine* e, int32_t on) {
  if (!e) return fail(TS_ERR_INVALID, "null engine");
  if (on < 0 || on > 2) return fail(TS_ERR_INVALID, "ts_engine_set_graphs: on must be 0, 1 or 2");
  e->graphs = on != 0;
  e->graphs_all = on == 2;
  return TS_OK;
}

ts_status ts_engine_graph_info(ts_engine* e, int64_t* captures, int64_t* replays, int64_t* cached) {
  if (!e) return fail(TS_ERR_INVALID, "null engine");
  if (captures) *captures = e->graph_captures;
  if (replays) *replays = e->graph_replays;
  if (cached) *cached = (int64_t)e->gcache.size();
  return TS_OK;
}

ts_status ts_engine_kernel_stats(ts_engine* e, int32_t idx, char* name, int32_t namelen, int64_t* launches,
                                 double* ms, double* bytes, double* units) {
  if (!e || idx < 0 || idx >= (int32_t)e->ks.size()) return TS_ERR_INVALID;
  const KStat& k = e->ks[idx];
  if (name && namelen > 0) std::snprintf(name, namelen, "%s", k.name.c_str());
  if (launches) *launches = k.launches;
  if (ms) *ms = k.ms;
  if (bytes) *bytes = k.bytes;
  if (units) *units = k.units;
  return TS_OK;
}

ts_status ts_engine_reset_stats(ts_engine* e) {
  if (!e) return fail(TS_ERR_INVALID, "null engine");
  for (auto& k : e->ks) k = KStat{k.name};
  return TS_OK;
}

ts_status ts_knn(ts_engine* e, const ts_set* db, const ts_set* queries, const ts_params* p, int32_t* out_idx,
                 float* out_val, ts_stats* stats) {
  return knn_impl(e, db, queries, p, out_idx, out_val, stats, false);
}

ts_status ts_knn_multi_k(ts_engine* e, const ts_set* db, const ts_set* queries, const ts_params* p,
                         const int32_t* k_per_query, int32_t* out_idx, float* out_val, ts_stats* stats) {
  if (!k_per_query) return fail(TS_ERR_INVALID, "null k_per_query");
  return knn_impl(e, db, queries, p, out_idx, out_val, stats, false, 0, k_per_query);
}

ts_status ts_knn_plain(ts_engine* e, const ts_set* db, const ts_set* queries, const ts_params* p, int32_t* out_idx,
                       float* out_val, ts_stats* stats) {
  return knn_impl(e, db, queries, p, out_idx, out_val, stats, true);
}

ts_status ts_filter_ref(ts_engine* e, const ts_set* db, const ts_set* qs, const ts_params* p, float* scores) {
  if (!e || !scores) return fail(TS_ERR_INVALID, "null argument");
  TS_TRY(check_params(p, db, qs));
  float4* env = nullptr;
  float* d = nullptr;
  TileView tiles{};
  TS_TRY(ensure(e, B_ENV, qs->dev_points, &env));
  TS_TRY(ensure(e, B_SCORES, (int64_t)qs->count * db->count, &d));
  if (p->measure == TS_DTW || p->measure == TS_FRECHET || p->measure == TS_LCSS) {
    TS_TRY(set_tiles(db, e->st, &tiles));
    TS_TRY(run(e, "envelope", 0, 0, [&] { return launch_envelope(e->st, qs->view(), qs->max_len, p->r, env); }));
  }
  TS_TRY(run(e, "filter_ref", 0, 0, [&] {
    return launch_filter_ref(e->st, p->measure, qs->view(), 0, qs->count, env, db->view(), &tiles, p->epsilon, d);
  }));
  TS_HIP(hipMemcpyAsync(scores, d, sizeof(float) * qs->count * db->count, hipMemcpyDeviceToHost, e->st));
  TS_HIP(hipStreamSynchronize(e->st));
  return resolve_timing(e);
}

ts_status ts_pair_values(ts_engine* e, const ts_set* db, const ts_set* qs, const ts_params* p, const int32_t* qidx,
                         const int32_t* didx, int64_t npairs, float* out) {
  if (!e || !qidx || !didx || !out || npairs < 0) return fail(TS_ERR_INVALID, "null argument");
  TS_TRY(check_params(p, db, qs));
  if (npairs == 0) return TS_OK;
  // a pair's value slot is a 32-bit index
  if (npairs > INT32_MAX) return fail(TS_ERR_UNSUPPORTED, "ts_pair_values: more than 2^31 - 1 pairs in one call");
  std::vector<Item> h(npairs);
  for (int64_t i = 0; i < npairs; ++i) {
    if (qidx[i] < 0 || qidx[i] >= qs->count || didx[i] < 0 || didx[i] >= db->count)
      return fail(TS_ERR_INVALID, "pair index out of range");
    h[i] = Item{qidx[i], didx[i], (int32_t)i, 0};
  }
  Item* items = nullptr;
  float* vals = nullptr;
  TS_TRY(ensure(e, B_ITEMS, npairs, &items));
  TS_TRY(ensure(e, B_VALS, npairs, &vals));
  TS_HIP(hipMemcpyAsync(items, h.data(), sizeof(Item) * npairs, hipMemcpyHostToDevice, e->st));
  e->cells_used = 0;
  unsigned long long* cells = nullptr;
  TS_TRY(ensure(e, B_CELLS, kCellBuf, &cells));
  TS_TRY(run(e, "fill", 0, 0, [&] { return zero_cells(e->st, cells); }));
  TS_TRY(refine(e, *p, db, qs, items, nullptr, npairs, vals));
  TS_HIP(hipMemcpyAsync(out, vals, sizeof(float) * npairs, hipMemcpyDeviceToHost, e->st));
  TS_HIP(hipStreamSynchronize(e->st));
  return resolve_timing(e);
}

ts_status ts_topk_rows(ts_engine* e, const float* scores, int32_t rows, int32_t cols, int32_t k, int32_t descending,
                       int32_t* out_idx, float* out_val) {
  if (!e || !scores || !out_idx || !out_val || rows < 0 || cols < 1 || k < 1 || k > 1000)
    return fail(TS_ERR_INVALID, "ts_topk_rows arguments");
  if (rows == 0) return TS_OK;
  float* d = nullptr;
  uint64_t *keys = nullptr, *top = nullptr;
  int32_t* di = nullptr;
  float* dv = nullptr;
  TS_TRY(ensure(e, B_SCORES, (int64_t)rows * cols, &d));
  TS_TRY(ensure(e, B_KEYS, (int64_t)rows * cols, &keys));
  TS_TRY(ensure(e, B_TOP, (int64_t)rows * k, &top));
  TS_TRY(ensure(e, B_OIDX, (int64_t)rows * k, &di));
  TS_TRY(ensure(e, B_OVAL, (int64_t)rows * k, &dv));
  TS_HIP(hipMemcpyAsync(d, scores, sizeof(float) * rows * cols, hipMemcpyHostToDevice, e->st));
  TS_TRY(run(e, "keys", 0, 0, [&] { return launch_keys_from_rows(e->st, d, rows, cols, descending, keys); }));
  TS_TRY(topk(e, keys, cols, cols, nullptr, rows, k, top));
  TS_TRY(run(e, "decode", 0, 0, [&] { return launch_decode_keys(e->st, top, (int64_t)rows * k, descending, di, dv); }));
  TS_HIP(hipMemcpyAsync(out_idx, di, sizeof(int32_t) * rows * k, hipMemcpyDeviceToHost, e->st));
  TS_HIP(hipMemcpyAsync(out_val, dv, sizeof(float) * rows * k, hipMemcpyDeviceToHost, e->st));
  TS_HIP(hipStreamSynchronize(e->st));
  return resolve_timing(e);
}

}  // extern "C"
