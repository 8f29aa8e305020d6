// select.hip -- per-query top-k on the device.
//
// Candidates are 64-bit keys: the value's order-preserving bits in the high
// word (inverted for descending measures) and the DB index in the low word,
// so ascending key order is the reference's candidate order: (value, index)
// for DTW/Frechet (std::partial_sort on pair<float,int>, dtw.cpp:62) and
// (value descending, index ascending) for LCSS (std::stable_sort, lcss.cpp:
// 67-70). Keys are unique, so the top-k set is well defined under ties.
//
// Selection is a tree of chunk selections: each block keeps the best k of up
// to 2048 keys of one segment (k <= 32: k rounds of a wave-wide minimum per
// quarter of the chunk, then over the quarters' 4k; larger k: a bitonic sort
// in LDS); stages repeat until one chunk per segment remains.
#include "ts_internal.h"

namespace ts {

constexpr int kSelChunk = 2048;

__device__ __forceinline__ uint32_t ord_bits_s(float v) {
  uint32_t b = __float_as_uint(v + 0.0f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ uint64_t key_of(float v, int32_t idx, bool desc) {
  if (!isfinite(v) || idx < 0) return ~0ull;
  uint32_t o = ord_bits_s(v);
  if (desc) o = ~o;
  return ((uint64_t)o << 32) | (uint32_t)idx;
}
__device__ __forceinline__ float value_of(uint64_t key, bool desc) {
  uint32_t o = (uint32_t)(key >> 32);
  if (desc) o = ~o;
  uint32_t b = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  return __uint_as_float(b);
}

__global__ void __launch_bounds__(256) k_keys_from_rows(const float* vals, int32_t rows, int32_t cols,
                                                        int32_t desc, uint64_t* keys) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t n = (int64_t)rows * cols;
  if (i >= n) return;
  keys[i] = key_of(vals[i], (int32_t)(i % cols), desc != 0);
}

hipError_t launch_keys_from_rows(hipStream_t st, const float* vals, int32_t rows, int32_t cols, int32_t descending,
                                 uint64_t* keys) {
  int64_t n = (int64_t)rows * cols;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_keys_from_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vals, rows, cols,
                     descending, keys);
  return hipGetLastError();
}

// Selections of at most kWaveK keys per chunk take the minimum key k times
// instead of sorting the chunk.
constexpr int kWaveK = 32;

// One DPP step of a wave-wide minimum of 64-bit keys: lanes whose source is
// outside the row or in a masked row see the empty key.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t min_dpp(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_update_dpp(0xffffffffu, (uint32_t)v, CTRL, ROWS, 0xf, false);
  const uint32_t hi = __builtin_amdgcn_update_dpp(0xffffffffu, (uint32_t)(v >> 32), CTRL, ROWS, 0xf, false);
  const uint64_t s = ((uint64_t)hi << 32) | lo;
  return s < v ? s : v;
}

// Minimum over the wave's 64 lanes (all lanes active): prefix minima within
// rows of 16 (row_shr 1, 2, 4, 8), rows combined by row_bcast 15 and 31, so
// lane 63 holds the minimum; returned to every lane.
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
  v = min_dpp<0x111, 0xf>(v);
  v = min_dpp<0x112, 0xf>(v);
  v = min_dpp<0x114, 0xf>(v);
  v = min_dpp<0x118, 0xf>(v);
  v = min_dpp<0x142, 0xa>(v);
  v = min_dpp<0x143, 0xc>(v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, 63), hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), 63);
  return ((uint64_t)hi << 32) | lo;
}

// Takes the k smallest of the wave's T keys per lane, in ascending order, into
// dst[0..k) (lane 0 writes; missing keys are empty). One occurrence is taken
// per round: the lowest lane holding the minimum gives up its first copy.
template <int T>
__device__ __forceinline__ void wave_extract(uint64_t (&v)[T], int k, uint64_t* dst) {
  const int lane = threadIdx.x & 63;
  for (int r = 0; r < k; ++r) {
    uint64_t lm = v[0];
    int at = 0;
#pragma unroll
    for (int t = 1; t < T; ++t)
      if (v[t] < lm) { lm = v[t]; at = t; }
    const uint64_t m = wave_min_u64(lm);
    if (m == ~0ull) {  // wave-uniform: no keys left
      for (int t = r + lane; t < k; t += 64) dst[t] = ~0ull;
      return;
    }
    if (lane == 0) dst[r] = m;
    const uint64_t holders = __ballot(lm == m);
    const bool mine = lane == __ffsll((unsigned long long)holders) - 1;
#pragma unroll
    for (int t = 0; t < T; ++t)
      if (mine && t == at) v[t] = ~0ull;
  }
}

// One selection stage: segment seg's keys [0, len) in chunks of ch; chunk c
// writes its best k to out[(seg * nch + c) * k]. Blocks stride over the
// chunks (the grid's x is capped), so an upper-bound seg_len costs little:
// chunks past a segment's end only write empty keys.
__global__ void __launch_bounds__(256) k_topk_stage(const uint64_t* in, int64_t seg_stride, int32_t seg_len,
                                                    const int32_t* counts, int32_t k, int32_t ch, int32_t nch,
                                                    uint64_t* out, const int64_t* seg_off) {
  __shared__ uint64_t s[kSelChunk];
  const int seg = blockIdx.y;
  if (k <= kWaveK) {
    // Each wave takes the best k of its quarter of the chunk (keys i with
    // i % 256 in the wave's 64), then wave 0 takes the best k of those 4k.
    uint64_t* part = s;  // [4][kWaveK]
    int len = seg_len;
    if (counts) len = min(counts[seg], seg_len);
    const uint64_t* base = in + (seg_off ? seg_off[seg] : (int64_t)seg * seg_stride);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int c = blockIdx.x; c < nch; c += gridDim.x) {
      const int begin = c * ch;
      const int nh = max(0, min(ch, len - begin));
      uint64_t* dst = out + ((int64_t)seg * nch + c) * k;
      if (nh == 0) {  // block-uniform
        for (int t = threadIdx.x; t < k; t += 256) dst[t] = ~0ull;
        continue;
      }
      const uint64_t* src = base + begin;
      uint64_t v[kSelChunk / 256];
#pragma unroll
      for (int t = 0; t < kSelChunk / 256; ++t) {
        const int i = t * 256 + threadIdx.x;
        v[t] = i < nh ? src[i] : ~0ull;
      }
      __syncthreads();  // the previous chunk's merge is done with part
      wave_extract(v, k, part + wv * kWaveK);
      __syncthreads();
      if (wv == 0) {
        uint64_t w[2];
        w[0] = lane < 2 * k ? part[(lane / k) * kWaveK + lane % k] : ~0ull;
        w[1] = lane < 2 * k ? part[(2 + lane / k) * kWaveK + lane % k] : ~0ull;
        wave_extract(w, k, dst);
      }
    }
    return;
  }
  int len = seg_len;
  if (counts) len = min(counts[seg], seg_len);
  const uint64_t* base = in + (seg_off ? seg_off[seg] : (int64_t)seg * seg_stride);
  for (int c = blockIdx.x; c < nch; c += gridDim.x) {
    const int begin = c * ch;
    const int nh = max(0, min(ch, len - begin));
    uint64_t* dst = out + ((int64_t)seg * nch + c) * k;
    if (nh == 0) {  // block-uniform
      for (int t = threadIdx.x; t < k; t += 256) dst[t] = ~0ull;
      continue;
    }
    // Sort only the smallest power of two (>= 64) holding the chunk's keys:
    // the rest would be empty keys, which sort last.
    int cs = ch;
    while (cs > 64 && (cs >> 1) >= nh) cs >>= 1;
    __syncthreads();  // the previous chunk's results are out of s
    const uint64_t* src = base + begin;
    for (int i = threadIdx.x; i < cs; i += 256) s[i] = i < nh ? src[i] : ~0ull;
    __syncthreads();
    for (int kk = 2; kk <= cs; kk <<= 1) {
      for (int j = kk >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < cs; i += 256) {
          int ixj = i ^ j;
          if (ixj > i) {
            uint64_t a = s[i], b = s[ixj];
            bool asc = (i & kk) == 0;
            if ((a > b) == asc) { s[i] = b; s[ixj] = a; }
          }
        }
        __syncthreads();
      }
    }
    for (int t = threadIdx.x; t < k; t += 256) dst[t] = t < cs ? s[t] : ~0ull;
  }
}

// Seed threshold, one block per query: keys of the query's s refined seed
// values (index -1 or a non-finite value: empty), sorted in LDS; tau is the
// k-th best value, or +-inf when fewer than k seeds have one.
__global__ void __launch_bounds__(256) k_seed_tau(const float* vals, const int32_t* seed_idx, int32_t s, int32_t k,
                                                  const int32_t* kq, int32_t cs, int32_t desc, float* tau) {
  __shared__ uint64_t sk[kSelChunk];
  const int q = blockIdx.x;
  for (int i = threadIdx.x; i < cs; i += 256)
    sk[i] = i < s ? key_of(vals[(int64_t)q * s + i], seed_idx[(int64_t)q * s + i], desc != 0) : ~0ull;
  __syncthreads();
  for (int kk = 2; kk <= cs; kk <<= 1) {
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < cs; i += 256) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const uint64_t a = sk[i], b = sk[ixj];
          const bool asc = (i & kk) == 0;
          if ((a > b) == asc) { sk[i] = b; sk[ixj] = a; }
        }
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    const uint64_t key = sk[(kq ? kq[q] : k) - 1];
    const float inf = __int_as_float(0x7f800000);
    tau[q] = key == ~0ull ? (desc ? -inf : inf) : value_of(key, desc != 0);
  }
}

static int chunk_for(int32_t len) {
  int ch = 64;
  while (ch < len && ch < kSelChunk) ch <<= 1;
  return ch;
}

int64_t select_scratch_elems(int32_t seg_len, int32_t nseg, int32_t k) {
  int64_t nch = (seg_len + kSelChunk - 1) / kSelChunk;
  if (nch <= 1) return 1;
  return (int64_t)nseg * nch * k;
}

hipError_t select_topk(hipStream_t st, const uint64_t* in, int64_t seg_stride, int32_t seg_len,
                       const int32_t* counts, int32_t nseg, int32_t k, uint64_t* out, uint64_t* scratch_a,
                       uint64_t* scratch_b, int64_t scratch_elems, int* launches, const int64_t* seg_off) {
  if (nseg == 0) return hipSuccess;
  if (k > kSelChunk / 2) return hipErrorInvalidValue;
  const uint64_t* cur = in;
  int64_t stride = seg_stride;
  int32_t len = seg_len;
  const int32_t* cnt = counts;
  uint64_t* bufs[2] = {scratch_a, scratch_b};
  int which = 0;
  for (;;) {
    int32_t nch = (len + kSelChunk - 1) / kSelChunk;
    if (nch < 1) nch = 1;
    int ch = nch == 1 ? chunk_for(len) : kSelChunk;
    uint64_t* dst = nch == 1 ? out : bufs[which];
    if (nch > 1 && (int64_t)nseg * nch * k > scratch_elems) return hipErrorInvalidValue;
    const int gx = nch < 64 ? nch : 64;
    hipLaunchKernelGGL(k_topk_stage, dim3(gx, nseg), dim3(256), 0, st, cur, stride, len, cnt, k, ch, nch, dst,
                       seg_off);
    if (launches) ++*launches;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nch == 1) return e;
    cur = dst;
    stride = (int64_t)nch * k;
    len = nch * k;
    cnt = nullptr;
    seg_off = nullptr;
    which ^= 1;
  }
}

// Threads past n copy tail_src[0, ntail) to tail_dst (counters that leave
// the device in the same copy as the results).
__global__ void __launch_bounds__(256) k_decode(const uint64_t* keys, int64_t n, int32_t desc, int32_t* idx,
                                                float* val, const unsigned long long* tail_src,
                                                unsigned long long* tail_dst, int32_t ntail) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) {
    if (i - n < ntail) tail_dst[i - n] = tail_src[i - n];
    return;
  }
  uint64_t k = keys[i];
  if (k == ~0ull) {
    if (idx) idx[i] = -1;
    if (val) val[i] = 0.0f;
  } else {
    if (idx) idx[i] = (int32_t)(uint32_t)k;
    if (val) val[i] = value_of(k, desc != 0);
  }
}

hipError_t launch_decode_keys(hipStream_t st, const uint64_t* keys, int64_t n, int32_t descending, int32_t* out_idx,
                              float* out_val, const unsigned long long* tail_src, unsigned long long* tail_dst,
                              int32_t ntail) {
  if (!tail_src) ntail = 0;
  if (n + ntail == 0) return hipSuccess;
  hipLaunchKernelGGL(k_decode, dim3((unsigned)((n + ntail + 255) / 256)), dim3(256), 0, st, keys, n, descending,
                     out_idx, out_val, tail_src, tail_dst, ntail);
  return hipGetLastError();
}

__global__ void k_kth(const uint64_t* keys, int32_t nseg, int32_t s, int32_t k, int32_t desc, float* tau) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nseg) return;
  uint64_t key = keys[(int64_t)i * s + (k - 1)];
  float inf = __int_as_float(0x7f800000);
  tau[i] = key == ~0ull ? (desc ? -inf : inf) : value_of(key, desc != 0);
}

hipError_t launch_seed_tau(hipStream_t st, const float* vals, const int32_t* seed_idx, int32_t nq, int32_t s,
                           int32_t k, int32_t descending, float* tau, const int32_t* kq) {
  if (nq == 0) return hipSuccess;
  if (s > kSelChunk || k < 1 || k > s) return hipErrorInvalidValue;
  const int cs = chunk_for(s);
  hipLaunchKernelGGL(k_seed_tau, dim3(nq), dim3(256), 0, st, vals, seed_idx, s, k, kq, cs, descending, tau);
  return hipGetLastError();
}

// Threshold of the second seed round: tau[q] = k-th best (kq[q]-th with kq)
// of query q's s first-round seed values and its second-round values, the
// value of the candidate slot in each of its s2 keys (rvals[slot], DB index
// citems[slot].d; ~0 keys and dropped pairs count as missing). The two sets
// are disjoint pairs.
__global__ void __launch_bounds__(256) k_tau2(const float* vals, const int32_t* seed_idx, int32_t s,
                                              const uint64_t* top2, int32_t s2, const float* rvals,
                                              const Item* citems, int32_t k, const int32_t* kq, int32_t cs,
                                              float* tau) {
  __shared__ uint64_t sk[kSelChunk];
  const int q = blockIdx.x;
  for (int i = threadIdx.x; i < cs; i += 256) {
    uint64_t key = ~0ull;
    if (i < s) {
      key = key_of(vals[(int64_t)q * s + i], seed_idx[(int64_t)q * s + i], false);
    } else if (i < s + s2) {
      const uint64_t t = top2[(int64_t)q * s2 + (i - s)];
      if (t != ~0ull) {
        const uint32_t slot = (uint32_t)t;
        key = key_of(rvals[slot], citems[slot].d, false);
      }
    }
    sk[i] = key;
  }
  __syncthreads();
  for (int kk = 2; kk <= cs; kk <<= 1) {
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < cs; i += 256) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const uint64_t a = sk[i], b = sk[ixj];
          const bool asc = (i & kk) == 0;
          if ((a > b) == asc) { sk[i] = b; sk[ixj] = a; }
        }
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    const uint64_t key = sk[(kq ? kq[q] : k) - 1];
    tau[q] = key == ~0ull ? __int_as_float(0x7f800000) : value_of(key, false);
  }
}

hipError_t launch_tau2(hipStream_t st, const float* vals, const int32_t* seed_idx, int32_t nq, int32_t s,
                       const uint64_t* top2, int32_t s2, const float* rvals, const Item* citems, int32_t k,
                       float* tau, const int32_t* kq) {
  if (nq == 0) return hipSuccess;
  if (s + s2 > kSelChunk || k < 1 || k > s) return hipErrorInvalidValue;
  const int cs = chunk_for(s + s2);
  hipLaunchKernelGGL(k_tau2, dim3(nq), dim3(256), 0, st, vals, seed_idx, s, top2, s2, rvals, citems, k, kq, cs, tau);
  return hipGetLastError();
}

// Keys of candidate slots by their filter bound: (ordered bound, slot), ~0
// for slots with no bound (NaN) or no pair. Slots are below 2^31.
__global__ void __launch_bounds__(256) k_keys_from_slots(const float* tb, const Item* citems, int64_t n,
                                                         uint64_t* keys) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float b = tb[i];
  keys[i] = (b == b && citems[i].d >= 0) ? ((uint64_t)ord_bits_s(b) << 32) | (uint32_t)i : ~0ull;
}

hipError_t launch_keys_from_slots(hipStream_t st, const float* tb, const Item* citems, int64_t n, uint64_t* keys) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_keys_from_slots, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, tb, citems, n, keys);
  return hipGetLastError();
}

// out[q][0, k) = top[q][0, k), out[q][k, k + s) = the keys of query q's s
// first-round seed values, for a top-k over both.
__global__ void __launch_bounds__(256) k_merge_seed_keys(const uint64_t* top, int32_t k, const float* vals,
                                                         const int32_t* seed_idx, int32_t s, int32_t nq,
                                                         uint64_t* out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t w = k + s;
  if (i >= (int64_t)nq * w) return;
  const int64_t q = i / w, j = i % w;
  out[i] = j < k ? top[q * k + j] : key_of(vals[q * s + (j - k)], seed_idx[q * s + (j - k)], false);
}

hipError_t launch_merge_seed_keys(hipStream_t st, const uint64_t* top, int32_t k, const float* vals,
                                  const int32_t* seed_idx, int32_t s, int32_t nq, uint64_t* out) {
  const int64_t n = (int64_t)nq * (k + s);
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_merge_seed_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, top, k, vals, seed_idx,
                     s, nq, out);
  return hipGetLastError();
}

hipError_t launch_kth_threshold(hipStream_t st, const uint64_t* keys, int32_t nseg, int32_t s, int32_t k,
                                int32_t descending, float* tau) {
  if (nseg == 0) return hipSuccess;
  hipLaunchKernelGGL(k_kth, dim3((nseg + 255) / 256), dim3(256), 0, st, keys, nseg, s, k, descending, tau);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_items_from_keys(const uint64_t* keys, int32_t rows, int32_t s, int32_t q0,
                                                         Item* items, int32_t* idx, float* nan_vals) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)rows * s) return;
  uint64_t k = keys[i];
  Item it;
  it.q = q0 + (int32_t)(i / s);
  it.d = k == ~0ull ? -1 : (int32_t)(uint32_t)k;
  it.slot = (int32_t)i;
  it.pad = 0;
  items[i] = it;
  if (idx) idx[i] = it.d;
  if (nan_vals) nan_vals[i] = __int_as_float(0x7fc00000);
}

hipError_t launch_items_from_keys(hipStream_t st, const uint64_t* keys, int32_t rows, int32_t s, int32_t q0,
                                  Item* items, int32_t* idx, float* nan_vals) {
  int64_t n = (int64_t)rows * s;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_items_from_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, keys, rows, s, q0,
                     items, idx, nan_vals);
  return hipGetLastError();
}

// Candidate regions sized on the device: exclusive prefix of counts[nq]
// into off[nq+1] (int64), the regions' total clamped to cap into *ntot, each
// region's length clamped to the capacity into segc[nq], and the unclamped
// total folded into *need (atomic max over the batches of one call). One
// block of 1024 threads.
__global__ void __launch_bounds__(1024) k_scan_offsets(const int32_t* counts, int32_t nq, int64_t cap, int64_t* off,
                                                       int32_t* ntot, int32_t* segc, unsigned long long* need) {
  __shared__ int64_t part[1024];
  const int per = (nq + 1023) / 1024;
  const int b = threadIdx.x * per;
  int64_t s = 0;
  for (int i = b; i < min(nq, b + per); ++i) s += counts[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const int64_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int64_t run = part[threadIdx.x] - s;
  for (int i = b; i < min(nq, b + per); ++i) {
    off[i] = run;
    segc[i] = (int32_t)max<int64_t>(0, min<int64_t>(counts[i], cap - run));
    run += counts[i];
  }
  if (threadIdx.x == 1023) {
    const int64_t total = part[1023];
    off[nq] = total;
    *ntot = (int32_t)min<int64_t>(total, cap);
    atomicMax(need, (unsigned long long)total);
  }
}

hipError_t launch_scan_offsets(hipStream_t st, const int32_t* counts, int32_t nq, int64_t cap, int64_t* off,
                               int32_t* ntot, int32_t* segc, unsigned long long* need) {
  hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, st, counts, nq, cap, off, ntot, segc, need);
  return hipGetLastError();
}

// Every (query, DB trajectory) pair of queries [q0, q0+nq): item i is
// (q0 + i / N, i % N) and writes slot i.
__global__ void __launch_bounds__(256) k_items_all(int32_t q0, int32_t nq, int32_t N, Item* items) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)nq * N) return;
  items[i] = Item{q0 + (int32_t)(i / N), (int32_t)(i % N), (int32_t)i, 0};
}

hipError_t launch_items_all(hipStream_t st, int32_t q0, int32_t nq, int32_t N, Item* items) {
  const int64_t n = (int64_t)nq * N;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_items_all, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, q0, nq, N, items);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_keys_from_regions(const float* vals, const int32_t* region_d,
                                                           const int32_t* counts, int32_t nq, int32_t cap,
                                                           int32_t desc, uint64_t* keys) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)nq * cap) return;
  const int q = (int)(i / cap), t = (int)(i % cap);
  const int c = counts ? counts[q] : cap;
  keys[i] = t < c ? key_of(vals[i], region_d[i], desc != 0) : ~0ull;
}

hipError_t launch_keys_from_regions(hipStream_t st, const float* vals, const int32_t* region_d,
                                    const int32_t* counts, int32_t nq, int32_t cap, int32_t descending,
                                    uint64_t* keys) {
  int64_t n = (int64_t)nq * cap;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_keys_from_regions, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vals, region_d,
                     counts, nq, cap, descending, keys);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_keys_from_items(const float* vals, const Item* items, int64_t n,
                                                         int32_t desc, uint64_t* keys) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  keys[i] = key_of(vals[i], items[i].d, desc != 0);
}

hipError_t launch_keys_from_items(hipStream_t st, const float* vals, const Item* items, int64_t n, int32_t descending,
                                  uint64_t* keys) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_keys_from_items, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vals, items, n,
                     descending, keys);
  return hipGetLastError();
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_i32(hipStream_t st, int32_t* p, int32_t v, int64_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_items_by_bucket(const Item* items, const int32_t* nitems,
                                                         int64_t max_items, SetView q, SetView db,
                                                         MeasureArgs m, Item* outb, int64_t cap_each,
                                                         int32_t* heads, int32_t* tails, BucketCells long_cells) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int64_t n = nitems ? (int64_t)*nitems : max_items;
  if (n > max_items) n = max_items;
  Item it{0, -1, 0, 0};
  if (i < n) it = items[i];
  int b = -1;
  bool lng = false;
  if (it.d >= 0) {
    const int na = q.len[it.q], nb = db.len[it.d];
    b = dp_bucket_of(dp_width_of(na, nb, m));
    lng = (int64_t)na * nb >= long_cells.t[b];
  }
  const int lane = threadIdx.x & 63;
  // One atomic per wave, bucket and end (wave ballot + prefix count).
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

hipError_t launch_items_by_bucket(hipStream_t st, const Item* items, const int32_t* nitems, int64_t max_items,
                                  SetView q, SetView db, MeasureArgs m, Item* outb, int64_t cap_each,
                                  int32_t* heads, int32_t* tails, BucketCells long_cells) {
  if (max_items == 0) return hipSuccess;
  hipLaunchKernelGGL(k_items_by_bucket, dim3((unsigned)((max_items + 255) / 256)), dim3(256), 0, st, items, nitems,
                     max_items, q, db, m, outb, cap_each, heads, tails, long_cells);
  return hipGetLastError();
}

// Dense item list from per-query regions: one block scans the counts
// (exclusive prefix), then every region entry is scattered.
__global__ void __launch_bounds__(1024) k_region_scan(const int32_t* counts, int32_t nq, int32_t cap, int32_t* offs,
                                                      int32_t* total) {
  __shared__ int32_t part[1024];
  const int per = (nq + 1023) / 1024;
  const int b = threadIdx.x * per;
  int32_t s = 0;
  for (int i = b; i < min(nq, b + per); ++i) s += min(counts[i], cap);
  part[threadIdx.x] = s;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    int32_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int32_t run = part[threadIdx.x] - s;
  for (int i = b; i < min(nq, b + per); ++i) {
    offs[i] = run;
    run += min(counts[i], cap);
  }
  if (threadIdx.x == 1023) *total = part[1023];
}

__global__ void __launch_bounds__(256) k_region_scatter(const int32_t* region_d, const int32_t* counts,
                                                        const int32_t* offs, int32_t nq, int32_t cap, int32_t q0,
                                                        Item* items) {
  const int q = blockIdx.y;
  const int c = min(counts[q], cap);
  for (int t = blockIdx.x * 256 + threadIdx.x; t < c; t += gridDim.x * 256) {
    Item it;
    it.q = q0 + q;
    it.d = region_d[(int64_t)q * cap + t];
    it.slot = q * cap + t;
    it.pad = 0;
    items[offs[q] + t] = it;
  }
}

hipError_t launch_items_from_regions(hipStream_t st, const int32_t* region_d, const int32_t* counts, int32_t nq,
                                     int32_t cap, int32_t q0, Item* items, int32_t* nitems) {
  if (nq == 0) return hipSuccess;
  // nitems[0] receives the total; nitems[1 .. nq] hold the offsets.
  hipLaunchKernelGGL(k_region_scan, dim3(1), dim3(1024), 0, st, counts, nq, cap, nitems + 1, nitems);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_region_scatter, dim3(std::min(64, (cap + 255) / 256), nq), dim3(256), 0, st, region_d, counts,
                     nitems + 1, nq, cap, q0, items);
  return hipGetLastError();
}

}  // namespace ts
