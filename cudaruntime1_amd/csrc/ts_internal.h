// ts_internal.h -- shared declarations of the trajsim HIP engine.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdlib>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/trajsim.h"

namespace ts {

// DP border value: the reference's "infinity" (dtw.cpp:68-69, dtw.cu:59).
constexpr float kInf = 1e20f;

// A lower bound of sqrt(x) for admissible filter bounds (never for a value
// that is returned): v_sqrt_f32 is within 1 ulp, so scaling it by 1 - 2^-21
// stays below the true root (by at least 2^-21 - 2^-23 - 2^-24 relative);
// below 2^-96, where v_sqrt_f32 may flush, 0. Four instructions instead of
// the correctly rounded sqrtf's sixteen.
__device__ __forceinline__ float sqrt_lo(float x) {
  return x < 0x1p-96f ? 0.0f : __builtin_amdgcn_sqrtf(x) * (1.0f - 0x1p-21f);
}

// Squared distance from a point to a box (minx, maxx, miny, maxy). For any
// point b of the box, the f32 squared distance from p to b is at least this
// value: f32 subtraction, squaring of magnitudes and addition are monotone.
// Per axis, p minus p clamped into the box (v_med3_f32): 0 inside, else
// exactly the negated or plain f32 gap max(lo - p, p - hi), with one
// instruction fewer.
__device__ __forceinline__ float pt_box(float px, float py, float4 b) {
  const float gx = px - __builtin_amdgcn_fmed3f(px, b.x, b.y);
  const float gy = py - __builtin_amdgcn_fmed3f(py, b.z, b.w);
  return gx * gx + gy * gy;
}

// Correctly rounded sqrtf of U values. v_sqrt_f32 returns s within 1 ulp
// of the root; the root rounds to s - 1 ulp, s or s + 1 ulp, and fma
// residuals against the two midpoint products decide which: with
// sm = s - 1 ulp (0 for s = 0) and sp = s + 1 ulp, the result is
// sm + [x - sm*s > 0] + [x - sp*s > 0] in ulps. Each bracket is the residual's
// bit pattern clamped to [0, 1] as an integer (v_med3_i32: +0 and negative
// values give 0), and the sum one v_add3_u32, so no compare masks are
// needed. The residuals are exact for x = 0 and x in [2^-96, 2^127), checked
// on the CPU for every such float and every start within 1 ulp. Inputs in
// (0, 2^-96) need a tiny coordinate (0 < |c| < 2^-25), and +inf or NaN a huge
// or non-finite one (|c| >= 2^62); sets flag such trajectories, `tiny` says
// whether a pair of the wave's lanes has one (uniform), and then a wave
// holding any such x takes sqrtf itself.
// v clamped to [0, 1] as a signed integer, in one v_med3_i32 (written out:
// the compiler turns min/max into a compare and a carry, which costs a
// lane-mask hazard per use).
__device__ __forceinline__ uint32_t med3_01(int32_t v) {
  uint32_t r;
  asm("v_med3_i32 %0, %1, 0, 1" : "=v"(r) : "v"(v));
  return r;
}

template <int U>
__device__ __forceinline__ void sqrt_cr(float (&x)[U], bool tiny) {
  if (__builtin_expect(tiny, 0)) {
    uint32_t mb = __float_as_uint(x[0]) - 1u;  // 0 maps to the top
    uint32_t mx = __float_as_uint(x[0]);
#pragma unroll
    for (int j = 1; j < U; ++j) {
      mb = min(mb, __float_as_uint(x[j]) - 1u);
      mx = max(mx, __float_as_uint(x[j]));
    }
    if (__builtin_expect(__any(mb < 0x0f7fffffu || mx >= 0x7f800000u), 0)) {
#pragma unroll
      for (int j = 0; j < U; ++j) x[j] = sqrtf(x[j]);
      return;
    }
  }
#pragma unroll
  for (int j = 0; j < U; ++j) {
    const float s = __builtin_amdgcn_sqrtf(x[j]);
    const uint32_t sb = __float_as_uint(s);
    const uint32_t smb = __builtin_elementwise_sub_sat(sb, 1u);
    const float fm = __builtin_fmaf(-__uint_as_float(smb), s, x[j]);
    const float fp = __builtin_fmaf(-__uint_as_float(sb + 1u), s, x[j]);
    x[j] = __uint_as_float(smb + med3_01(__float_as_int(fm)) + med3_01(__float_as_int(fp)));
  }
}

// Thread-local error message behind ts_last_error().
void set_error(const std::string& msg);
ts_status fail(ts_status s, const std::string& msg);
ts_status hip_fail(hipError_t e, const char* what, const char* file, int line);

#define TS_HIP(call)                                                          \
  do {                                                                        \
    hipError_t _e = (call);                                                   \
    if (_e != hipSuccess) return ::ts::hip_fail(_e, #call, __FILE__, __LINE__); \
  } while (0)

#define TS_TRY(call)                    \
  do {                                  \
    ts_status _s = (call);              \
    if (_s != TS_OK) return _s;         \
  } while (0)

// Trajectories start on multiples of kPad points (32 B of x and of y), the
// tail of each padded with copies of its last point, so DP kernels can load
// up to kPad consecutive points (a chunk of U <= 8 columns) with aligned
// vector loads that never leave the padded trajectory.
constexpr int kPad = 8;

// Read-only device view of a trajectory set, passed to kernels by value.
struct SetView {
  const float* x;        // [points]
  const float* y;        // [points]
  const int64_t* off;    // [count + 1]
  const int32_t* len;    // [count]
  const float4* mbr;     // [count] (minx, maxx, miny, maxy)
  const float4* ends;    // [count] (x0, y0, x_last, y_last)
  const float4* blk;     // [points / 8] box of each 8-point block (blocks never straddle trajectories)
  const float4* sblk;    // [points / 8] box of blocks g..g+7 of the same trajectory (64 points)
  int32_t count;
  const uint8_t* tiny;   // [count] 1: the trajectory has a tiny coordinate
};

// Length-bucketed tiles of a set: its trajectories sorted by (length, index)
// and cut into tiles of 64, each tile's points interleaved, so that lane l of
// a wave reading point v of the tile's trajectory l reads
// x[off[t] + v * 64 + l]: one coalesced 256-byte load per coordinate for 64
// trajectories. len[t] is the tile's longest length (its last trajectory's),
// id[t * 64 + l] the trajectory in slot l (-1 past the last trajectory);
// slots shorter than len[t] repeat their last point.
struct TileView {
  const float* x;
  const float* y;
  const int64_t* off;  // [ntiles]
  const int32_t* len;  // [ntiles]
  const int32_t* id;   // [ntiles * 64]
  int32_t ntiles;
};
constexpr int kTile = 64;

// A tiny coordinate: 0 < |c| < 2^-25. Between two trajectories without one
// every squared point distance is 0 or at least 2^-96 (coordinates of
// magnitude 2^-25 or more are multiples of 2^-48, so are their nonzero
// differences), the range where the DP's square root needs no rescaling.

// Work item of the refine kernels: one (query, database) pair and the slot
// its value is written to.
struct Item {
  int32_t q, d, slot, pad;
};

// Launch-time parameters of the measures.
struct MeasureArgs {
  int32_t measure, band, delta;
  float eps;
};

// True when the DP visits only a band of each row: a DTW Sakoe-Chiba band
// or an LCSS index window.
__host__ __device__ inline bool dp_ranged(const MeasureArgs& m) {
  return (m.measure == TS_DTW && m.band != 0) || (m.measure == TS_LCSS && m.delta >= 0);
}

// Seeds per query of the EXACT pipeline (the k-th best of their exact
// values is the pruning threshold), for a batch of `work` = queries x DB
// trajectories x mean query length x mean DB length: max(2k, 64) from 2^40
// on, where a tighter threshold saves more refine work than the extra seeds
// cost (4096 x 1M DTW: 464 -> 356 ms; 256 x 20k banded DTW of 2048-8192
// points: 314 -> 271 ms), max(2k, 16) below (64 x 100k of 64-256 points
// is latency-bound: more seeds lengthen it); lo_override replaces the 64
// or 16 (a pipeline with a second seed round starts from 32).
// TS_EXACT_SEEDS overrides; at least k, at most 1024.
inline int32_t exact_seeds(int32_t k, double work, int32_t lo_override = 0) {
  const char* v = std::getenv("TS_EXACT_SEEDS");
  const int32_t env = v ? std::atoi(v) : 0;
  const int32_t lo = lo_override > 0 ? lo_override : (work >= 1099511627776.0 ? 64 : 16);
  const int32_t s = env > 0 ? env : (2 * k > lo ? 2 * k : lo);
  return s < k ? k : (s > 1024 ? 1024 : s);
}

// Sum over i in [0, n) of floor((a i + b) / m), for a, b >= 0 and m > 0, by
// the usual reduction: take a and b modulo m (summing what they carry),
// then solve the problem with m and a exchanged; O(log m) rounds. Every
// intermediate stays below max(n, m)^2 + m.
__host__ __device__ inline uint64_t floor_sum(uint64_t n, uint64_t m, uint64_t a, uint64_t b) {
  uint64_t s = 0;
  while (n > 0) {
    if (a >= m) {
      s += n * (n - 1) / 2 * (a / m);
      a %= m;
    }
    if (b >= m) {
      s += n * (b / m);
      b %= m;
    }
    const uint64_t y = a * n + b;
    if (y < m) break;
    n = y / m;
    b = y % m;
    const uint64_t t = m;
    m = a;
    a = t;
  }
  return s;
}

// DP cells of an a x b pair as the DP visits them (R = max, C = min): every
// cell; a band of w = band (or ceil(-band R / 100)) and bw = w R, the cells
// with |r C - j R| <= bw, i.e. rows r from max(0, ceil((r C - bw) / R)) to
// min(C - 1, floor((r C + bw) / R)) (never empty: bw >= R); a window delta,
// the cells with |r - j| <= delta. Closed forms over the rows: the sums of
// the clamped ends split where the clamps start to bind, floor sums between.
__host__ __device__ inline uint64_t pair_cells(int64_t a, int64_t b, int32_t measure, int32_t band, int32_t delta) {
  if (measure == TS_EUCLIDEAN) return (uint64_t)a;
  const int64_t R = a > b ? a : b, C = a < b ? a : b;
  if (measure == TS_DTW && band != 0) {
    const int64_t bw = (band > 0 ? (int64_t)band : (-(int64_t)band * R + 99) / 100) * R;
    // sum of the upper ends: floor((r C + bw) / R) until it reaches C - 1
    const int64_t t1 = (C - 1) * R - bw;
    int64_t r1 = t1 <= 0 ? 0 : (t1 + C - 1) / C;
    if (r1 > R) r1 = R;
    const uint64_t hi = floor_sum((uint64_t)r1, (uint64_t)R, (uint64_t)C, (uint64_t)bw) + (uint64_t)(R - r1) * (uint64_t)(C - 1);
    // sum of the lower ends: ceil((r C - bw) / R) from the first r with r C > bw
    const int64_t r0 = bw / C + 1;
    const uint64_t lo = r0 < R ? floor_sum((uint64_t)(R - r0), (uint64_t)R, (uint64_t)C, (uint64_t)(r0 * C + R - 1 - bw)) : 0;
    return hi - lo + (uint64_t)R;
  }
  if (measure == TS_LCSS && delta >= 0) {
    // rows r < min(R, C + delta); upper ends r + delta up to C - 1, lower
    // ends r - delta from r = delta + 1
    const int64_t d = delta, nr = R < C + d ? R : C + d;
    int64_t t = C - d;
    t = t < 0 ? 0 : (t > nr ? nr : t);
    const uint64_t hi = (uint64_t)(t * (t - 1) / 2 + t * d) + (uint64_t)(nr - t) * (uint64_t)(C - 1);
    const int64_t u = nr - 1 - d > 0 ? nr - 1 - d : 0;
    return hi - (uint64_t)(u * (u + 1) / 2) + (uint64_t)nr;
  }
  return (uint64_t)a * (uint64_t)b;
}

// Columns of a row the DP visits for an n x m pair: the shorter length C,
// or for a band of half-width w (w = band, or ceil(-band * max(n,m) / 100)
// for a negative band) or a window delta, at most 2w+1 or 2*delta+1.
__host__ __device__ inline int32_t dp_width_of(int32_t n, int32_t m, const MeasureArgs& a) {
  const int32_t c = n < m ? n : m, r = n < m ? m : n;
  int64_t e = c;
  if (a.measure == TS_DTW && a.band != 0) {
    const int64_t w = a.band > 0 ? a.band : ((int64_t)-a.band * r + 99) / 100;
    e = 2 * w + 1;
  } else if (a.measure == TS_LCSS && a.delta >= 0) {
    e = 2 * (int64_t)a.delta + 1;
  }
  return (int32_t)(e < c ? e : c);
}

}  // namespace ts

// Device-resident set (the opaque ts_set of the C ABI).
struct ts_set {
  int32_t count = 0;
  int64_t points = 0;      // caller's points
  int64_t dev_points = 0;  // device points including per-trajectory padding
  int32_t min_len = 0, max_len = 0;
  int device = 0;
  float* x = nullptr;
  float* y = nullptr;
  int64_t* off = nullptr;
  int32_t* len = nullptr;
  float4* mbr = nullptr;
  float4* ends = nullptr;
  float4* blk = nullptr;
  float4* sblk = nullptr;
  uint8_t* tiny = nullptr;  // [count]
  std::vector<int64_t> h_off;
  std::vector<int32_t> h_len;
  int64_t bytes = 0;
  // Capacity of the device arrays (sets reloaded in place by the server).
  int64_t cap_points = 0;
  int32_t cap_count = 0;
  // Contents generation: bumped by every load or append, so derived views
  // (the tiles) know when they are stale.
  uint64_t gen = 0;
  // Length-bucketed tile view (TileView), built on first use by set_tiles;
  // device arrays grow only.
  struct Tiles {
    float* x = nullptr;
    float* y = nullptr;
    int64_t* off = nullptr;
    int32_t* len = nullptr;
    int32_t* id = nullptr;
    int32_t ntiles = 0;
    int64_t cap_elems = 0;
    int32_t cap_tiles = 0;
    uint64_t gen = ~0ull;  // contents generation the tiles were built from
    int64_t bytes = 0;
  };
  mutable Tiles tiles;
  mutable std::mutex tiles_mu;
  ts::SetView view() const { return {x, y, off, len, mbr, ends, blk, sblk, count, tiny}; }
};

namespace ts {

// ------------------------------------------------------------ launchers
// Each launcher enqueues its kernel(s) on `st` and returns the launch status.

// set.hip
// MBR, endpoints and tiny-coordinate flag of trajectories [t0, count);
// 8-point block boxes from point point0 on and superblock boxes of
// trajectories [t0, count).
hipError_t launch_summaries(hipStream_t st, SetView s, float4* mbr, float4* ends, uint8_t* tiny, int32_t t0 = 0);
hipError_t launch_blocks(hipStream_t st, SetView s, int64_t points, float4* blk, float4* sblk, int32_t t0 = 0,
                         int64_t point0 = 0);
// The set's tile view, built on `st` (and waited for) when the set's
// contents changed since the last build; thread-safe.
ts_status set_tiles(const ts_set* s, hipStream_t st, TileView* out);
void free_tiles(const ts_set* s);
// (Re)load a set in place on `st`: grows the device arrays only when the new
// contents do not fit, so a warmed-up set reloads without hipMalloc/hipFree.
ts_status set_load(ts_set* s, const ts_point* xy, const int64_t* offsets, int32_t count, hipStream_t st);
// Append trajectories to a set (device arrays grow when needed, keeping the
// contents); summaries are computed for the new trajectories only.
ts_status set_append(ts_set* s, const ts_point* xy, const int64_t* offsets, int32_t count);
// Device arrays for at least points_cap padded points and count_cap
// trajectories; with keep, the set's current contents are copied over.
ts_status set_reserve(ts_set* s, int64_t points_cap, int32_t count_cap, bool keep);
// Points padded to kPad per trajectory (the device layout): padded offsets
// doff[count+1] (from base), x and y with each trajectory's last point
// repeated into its padding. Fails on a length outside 1..2^24.
ts_status pad_layout(const ts_point* xy, const int64_t* offsets, int32_t count, int64_t base,
                     std::vector<int32_t>& lens, std::vector<int64_t>& doff, std::vector<float>& hx,
                     std::vector<float>& hy);

// filter.hip
hipError_t launch_envelope(hipStream_t st, SetView q, int32_t max_len, int32_t r, float4* env);
// REF filter scores[q * N + j] for queries [q0, q0+nq); the point measures
// (DTW, Frechet, LCSS) read the DB through its tile view.
hipError_t launch_filter_ref(hipStream_t st, int32_t measure, SetView q, int32_t q0, int32_t nq,
                             const float4* env, SetView db, const TileView* tiles, float eps, float* scores);
// Admissible summary bounds (EXACT mode) of queries [q0,q0+nq) against the
// whole DB, reduced per DB chunk of `chunk` = bound_chunk_items(s, nq, N)
// trajectories to the best `s` keys per query: keys[(q * nchunks + c) * s + t].
// bound_chunk_items(s) is the smallest chunk any batch takes (for sizing).
int32_t bound_chunk_items(int32_t s);
int32_t bound_chunk_items(int32_t s, int32_t nq, int32_t n_db);
// With bstore, also stores every pair's bound at bstore[q * N + j]; zeroes
// zero[0, nzero) (nzero <= 256) for the batch's later kernels.
hipError_t launch_bound_chunks(hipStream_t st, MeasureArgs m, int32_t k, SetView q, int32_t q0,
                               int32_t nq, SetView db, int32_t s, int32_t chunk, uint64_t* keys, float* bstore,
                               int32_t* zero = nullptr, int32_t nzero = 0);
// Threshold compaction of every DB trajectory whose bound can still reach
// the query's threshold tau[q] (bounds read from bvals[q * N + j] when given,
// else evaluated), in compact_blocks(N) blocks per query:
// with items == null stores each block's survivors in
// bcounts[q * compact_blocks(N) + block] and the pass masks (64 DB
// trajectories a word) in bits[q * compact_mask_words(N) + word];
// launch_block_prefix then turns the counts into exclusive offsets within
// each query (counts[q] = total); with items, reads the masks and writes the
// survivors as Items at off[q] + bcounts[...] + rank; positions >= cap are
// dropped.
int32_t compact_blocks(int32_t n_db);
int64_t compact_mask_words(int32_t n_db);
hipError_t launch_bound_compact(hipStream_t st, MeasureArgs m, int32_t k, SetView q, int32_t q0, int32_t nq,
                                SetView db, const float* tau, const int64_t* off, const float* bvals,
                                int32_t* bcounts, uint64_t* bits, Item* items, int64_t cap);
hipError_t launch_block_prefix(hipStream_t st, int32_t* bcounts, int32_t nq, int32_t gx, int32_t* counts);
// Pre-check of the tight filter's input for unbanded DTW and Frechet (one
// lane per pair, superblock boxes against the other MBR, no point read):
// rejected pairs get NaN in rej[slot] (and tb[slot]); the rest are appended
// to `out`, *nout counting them (zeroed by the caller: launch_block_prefix
// clears counts[nq + 1]).
hipError_t launch_tight_pre(hipStream_t st, MeasureArgs m, SetView q, int32_t q0, SetView db, const float* tau,
                            const Item* items, const int32_t* nitems, int64_t max_items, Item* out, int32_t* nout,
                            float* rej, float* tb);
// Second, O(n+m) admissible bound over compacted pairs (one wave per pair);
// survivors are appended to `out` with count *nout (zeroed by the caller:
// launch_block_prefix clears counts[nq]); a rejected pair's rej[slot] is set
// to NaN. DTW and Frechet only: with tb, tb[slot] = a survivor's bound (the
// value compared with its threshold), NaN for every other pair; with
// seed_idx (s DB indices per query, row q - q0), those pairs are rejected
// too. pre: the items come from launch_tight_pre (unbanded DTW then skips
// the staged-tile test the pre-check already made).
hipError_t launch_tight_filter(hipStream_t st, MeasureArgs m, int32_t k, SetView q, int32_t q0, SetView db,
                               const float* tau, const Item* items, const int32_t* nitems, int64_t max_items,
                               Item* out, int32_t* nout, float* rej, int grid_waves, float* tb = nullptr,
                               const int32_t* seed_idx = nullptr, int32_t s = 0, bool pre = false);

// select.hip
// Key order: ascending u64. enc_asc(v,i) = ordered(v)<<32|i, desc uses ~ordered(v).
hipError_t launch_keys_from_rows(hipStream_t st, const float* vals, int32_t rows, int32_t cols,
                                 int32_t descending, uint64_t* keys);
// Segment top-k: segments of seg_len keys (or counts[s] if counts != null,
// clamped to seg_len) starting at in + s*seg_stride, or at in + seg_off[s]
// when seg_off != null; writes the k smallest keys of each segment (sorted,
// padded with ~0ull) to out + s*k.
hipError_t select_topk(hipStream_t st, const uint64_t* in, int64_t seg_stride, int32_t seg_len,
                       const int32_t* counts, int32_t nseg, int32_t k, uint64_t* out,
                       uint64_t* scratch_a, uint64_t* scratch_b, int64_t scratch_elems, int* launches,
                       const int64_t* seg_off = nullptr);
// keys[i] = key(vals[i], items[i].d) for the first n items (~0 for d < 0 or
// a missing value).
hipError_t launch_keys_from_items(hipStream_t st, const float* vals, const Item* items, int64_t n, int32_t descending,
                                  uint64_t* keys);
int64_t select_scratch_elems(int32_t seg_len, int32_t nseg, int32_t k);
// Decode sorted keys into (idx, value); with tail_src, also copies ntail
// counters to tail_dst.
hipError_t launch_decode_keys(hipStream_t st, const uint64_t* keys, int64_t n, int32_t descending,
                              int32_t* out_idx, float* out_val, const unsigned long long* tail_src = nullptr,
                              unsigned long long* tail_dst = nullptr, int32_t ntail = 0);
// Threshold per segment: value of key t (the k-th smallest), or +inf / -inf
// if the segment has fewer keys.
// tau[q] = k-th best of query q's s refined seed values (vals[q*s + i],
// seed_idx the DB index or -1); one sort per query.
// With kq, query q's threshold is its kq[q]-th best (each kq[q] in [1, k]).
hipError_t launch_seed_tau(hipStream_t st, const float* vals, const int32_t* seed_idx, int32_t nq, int32_t s,
                           int32_t k, int32_t descending, float* tau, const int32_t* kq = nullptr);
hipError_t launch_kth_threshold(hipStream_t st, const uint64_t* keys, int32_t nseg, int32_t s,
                                int32_t k, int32_t descending, float* tau);
// Second seed round (ascending measures): tau[q] = k-th best (kq[q]-th) of
// query q's s first-round values and the values rvals[slot] of the candidate
// slots in its s2 keys top2[q][...] (low 32 bits: slot; ~0: none).
hipError_t launch_tau2(hipStream_t st, const float* vals, const int32_t* seed_idx, int32_t nq, int32_t s,
                       const uint64_t* top2, int32_t s2, const float* rvals, const Item* citems, int32_t k,
                       float* tau, const int32_t* kq = nullptr);
// keys[i] = (ordered tb[i], i) for the first n candidate slots, ~0 where tb
// is NaN or the slot holds no pair.
hipError_t launch_keys_from_slots(hipStream_t st, const float* tb, const Item* citems, int64_t n, uint64_t* keys);
// out[q] = query q's k keys of top, then the keys of its s seed values.
hipError_t launch_merge_seed_keys(hipStream_t st, const uint64_t* top, int32_t k, const float* vals,
                                  const int32_t* seed_idx, int32_t s, int32_t nq, uint64_t* out);
// Build Items (q0+row, d=decoded idx, slot=row*s+t) from per-row sorted keys;
// with idx, also the decoded index (-1 for an empty key) per slot, and with
// nan_vals, NaN into every value slot (the refine overwrites the live ones).
hipError_t launch_items_from_keys(hipStream_t st, const uint64_t* keys, int32_t rows, int32_t s,
                                  int32_t q0, Item* items, int32_t* idx, float* nan_vals);
// Keys of refined values: key(value[slot], d) for every slot of the query
// regions, ~0 for empty slots.
hipError_t launch_keys_from_regions(hipStream_t st, const float* vals, const int32_t* region_d,
                                    const int32_t* counts, int32_t nq, int32_t cap, int32_t descending,
                                    uint64_t* keys);
hipError_t launch_fill_i32(hipStream_t st, int32_t* p, int32_t v, int64_t n);
// Split items by DP bucket (dp_bucket_of the visited row width,
// dp_width_of) into kDpBuckets lists of cap_each entries.
// Per-bucket work lists for the DP, longest pairs first: a pair of bucket b
// with n*m >= long_cells.t[b] is appended at the front of list b (count
// heads[b]), any other from the back (count tails[b], stored backwards from
// the list's last slot), so a persistent DP launch starts the long pairs
// first and the short ones fill in behind them.
hipError_t launch_items_by_bucket(hipStream_t st, const Item* items, const int32_t* nitems, int64_t max_items,
                                  SetView q, SetView db, MeasureArgs m, Item* outb, int64_t cap_each,
                                  int32_t* heads, int32_t* tails, struct BucketCells long_cells);
// Exclusive prefix of counts into off[nq+1] on the device; *ntot = total
// clamped to cap, segc[q] = region length clamped to cap, *need = max(*need,
// total).
hipError_t launch_scan_offsets(hipStream_t st, const int32_t* counts, int32_t nq, int64_t cap, int64_t* off,
                               int32_t* ntot, int32_t* segc, unsigned long long* need);
// Every pair of queries [q0, q0+nq) with all N DB trajectories, slot = index.
hipError_t launch_items_all(hipStream_t st, int32_t q0, int32_t nq, int32_t N, Item* items);
hipError_t launch_items_from_regions(hipStream_t st, const int32_t* region_d, const int32_t* counts,
                                     int32_t nq, int32_t cap, int32_t q0, Item* items, int32_t* nitems);

// dp.hip -- persistent wavefront DP over a device item list, one launch per
// bucket of the row width the DP visits (the shorter length C, or the band
// width): < 32, < 64, < 128, < 256, < 1024, >= 1024. Each launch runs one
// lane-group shape (lanes per pair W, cells per lane step U) out of
// kDpShapes; a tuning table picks the shape and grid per bucket.
constexpr int kDpBuckets = 6;
// Per-bucket n*m threshold of the DP work lists' long pairs (launch_items_by_bucket).
struct BucketCells {
  int64_t t[kDpBuckets];
};
// W > 64: a pair spans W / 64 waves of a block (dp_mw.hip).
constexpr int kDpShapes = 13;
struct DpShape {
  int W, U;
};
constexpr DpShape kDpShapeList[kDpShapes] = {{16, 1}, {16, 2}, {16, 4}, {32, 2}, {32, 4}, {64, 2}, {64, 4},
                                            {32, 8}, {64, 8}, {128, 4}, {128, 8}, {256, 4}, {256, 8}};
// Largest row width of each bucket.
constexpr int32_t kDpBucketCap[kDpBuckets] = {31, 63, 127, 255, 1023, 1 << 24};
// Shape used when no tuning table says otherwise.
constexpr int kDpDefaultShape[kDpBuckets] = {0, 1, 2, 4, 6, 6};
__host__ __device__ inline int dp_bucket_of(int32_t c) {
  return c <= 31 ? 0 : (c <= 63 ? 1 : (c <= 127 ? 2 : (c <= 255 ? 3 : (c <= 1023 ? 4 : 5))));
}
// LDS halo of bucket b: its length in floats, and the mask applied to
// column indexes (-1: the halo holds a whole row; otherwise a ring of
// len = mask + 1 floats that holds the band of one strip). cmax_c bounds
// the pairs' shorter length, wmax_c their dp_width_of.
struct DpHalo {
  int32_t len, mask;
};
DpHalo dp_halo(int shape, int b, const MeasureArgs& m, int32_t cmax_c, int32_t wmax_c);
// Whether a launch of `shape` with this halo fits a block's LDS (multi-wave
// shapes: with the strip table of pairs whose longer length is at most rmax).
bool dp_fits(int shape, DpHalo halo, int32_t rmax);
// tau (with tau_q0), cells: early abandoning of unbanded DTW and Frechet
// pairs whose value must exceed tau[q - tau_q0]; their slot gets NaN and
// `cells` loses their unfinished rows' cells. tau null: none.
hipError_t launch_dp(hipStream_t st, MeasureArgs m, int shape, SetView q, SetView db, const Item* items,
                     const int32_t* nitems, const int32_t* ntail, int64_t max_items, int32_t* next, float* out,
                     DpHalo halo, int32_t rmax_c, int32_t ccap_c, int grid_waves, const float* tau = nullptr,
                     int32_t tau_q0 = 0, unsigned long long* cells = nullptr);
// dp_mw.hip -- the multi-wave DP (nw waves per pair, u cells per lane step):
// halo floats per buffer and LDS bytes of a launch, and the launch itself
// (hcap from dp_mw_hcap, a grid of grid_waves waves).
int32_t dp_mw_hcap(int u, const MeasureArgs& m, int32_t cmax_c, int32_t wmax_c);
size_t dp_mw_lds_bytes(int32_t rmax_c, int32_t hcap);
// Lets the multi-wave kernels take the device's LDS per block (up to 160 KiB
// on gfx950); called by ts_init for each device. dp_mw_lds_max: the limit.
ts_status dp_mw_setup(int device);
int32_t dp_mw_lds_max();
hipError_t launch_dp_mw(hipStream_t st, MeasureArgs m, int nw, int u, SetView q, SetView db, const Item* items,
                        const int32_t* nitems, const int32_t* ntail, int64_t max_items, int32_t* next, float* out,
                        int32_t hcap, int32_t rmax_c, int grid_waves, const float* tau, int32_t tau_q0,
                        unsigned long long* cells);

// reduce.hip -- Hausdorff (wave per pair) and lock-step Euclidean (lane per pair).
hipError_t launch_hausdorff(hipStream_t st, SetView q, SetView db, const Item* items, const int32_t* nitems,
                            int64_t max_items, float* out);
hipError_t launch_euclidean(hipStream_t st, SetView q, SetView db, const Item* items, const int32_t* nitems,
                            int64_t max_items, float* out);

int device_cu_count();

// engine.hip -- cell / pair counters of a k-NN call or refine: kCellSlots
// of each, then the candidate total of device-sized EXACT batches.
constexpr int kCellSlots = 1024;
constexpr int kCellBuf = 2 * kCellSlots + 1;
// Device buffers a refine takes instead of growing the engine's workspace:
// cells[kCellBuf] (zero on entry), ctr[32] DP counters, and
// itemsb[kDpBuckets * max_items] bucket work lists.
struct RefineBufs {
  unsigned long long* cells;
  int32_t* ctr;
  Item* itemsb;
  int64_t max_items;
};
// Exact values of items[0, n) into vals[slot] on the engine's stream,
// every device buffer from `bufs`: no device allocation (per-kernel timing
// is off for the call). Does not wait for the stream.
ts_status refine_fixed(ts_engine* e, const ts_params& p, const ts_set* db, const ts_set* qs, Item* items,
                       int64_t n, float* vals, const RefineBufs& bufs);
hipStream_t engine_stream(const ts_engine* e);
int engine_device(const ts_engine* e);
// The argument checks of the k-NN and pair-value entry points.
ts_status validate_params(const ts_params* p, const ts_set* db, const ts_set* qs);

// noalloc.hip -- looks up the device address of the allocation-free entry
// point's scratch array on `device`, which must be current (once per device).
ts_status noalloc_setup(int device);

}  // namespace ts
