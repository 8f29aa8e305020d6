// filter.hip -- candidate filters.
//
// REF mode reproduces the reference's per-point envelope filters bit for bit:
// one lane per (query, DB trajectory) pair folding over the DB points in
// order (dtw.cpp:45-58 sum, frechet.cpp:46-62 max, lcss.cpp:46-63 count), so
// the f32 rounding sequence is the reference's. The index map v*n/m is
// advanced incrementally (quotient + remainder) instead of dividing per point.
//
// EXACT mode uses admissible summary bounds from 36 B per DB trajectory (MBR,
// endpoints, length; the query's summary is uniform per block) and either
// keeps the best s keys per DB chunk (seeding) or compacts every trajectory
// that can still reach the query's threshold with a wave ballot and one
// atomic per wave.
#include "ts_internal.h"

namespace ts {

__device__ __forceinline__ uint32_t ord_bits(float v) {
  uint32_t b = __float_as_uint(v + 0.0f);  // -0 -> +0, so equal values compare equal
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ uint64_t make_key(float v, int32_t idx, bool desc) {
  if (!isfinite(v)) return ~0ull;
  uint32_t o = ord_bits(v);
  if (desc) o = ~o;
  return ((uint64_t)o << 32) | (uint32_t)idx;
}

// Envelope of every query point over [i-r, i+r] (dtw.cpp:20-34).
__global__ void __launch_bounds__(256) k_envelope(SetView q, int32_t r, float4* env) {
  const int t = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int n = q.len[t];
  if (i >= n) return;
  const int64_t o = q.off[t];
  float mnx = 1e20f, mxx = -1e20f, mny = 1e20f, mxy = -1e20f;
  const int a = max(0, i - r), b = min(n - 1, i + r);
  for (int k = a; k <= b; ++k) {
    float px = q.x[o + k], py = q.y[o + k];
    mnx = fminf(mnx, px); mxx = fmaxf(mxx, px);
    mny = fminf(mny, py); mxy = fmaxf(mxy, py);
  }
  env[o + i] = make_float4(mnx, mxx, mny, mxy);
}

hipError_t launch_envelope(hipStream_t st, SetView q, int32_t max_len, int32_t r, float4* env) {
  hipLaunchKernelGGL(k_envelope, dim3((max_len + 255) / 256, q.count), dim3(256), 0, st, q, r, env);
  return hipGetLastError();
}

// REF filter of the MBR measures, one lane per (query, DB trajectory): the
// squared MBR gap, the R-tree's comparable box distance (hausdorff.cpp:162,
// euclidean.cpp:93-96); Euclidean pairs of unequal length are never
// candidates.
template <int MEAS>
__global__ void __launch_bounds__(256) k_filter_ref(SetView q, int32_t q0, const float4* env, SetView db,
                                                    float eps, float* scores) {
  static_assert(MEAS == TS_HAUSDORFF || MEAS == TS_EUCLIDEAN, "point measures run k_filter_ref_tiled");
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= db.count) return;
  const int qi = q0 + blockIdx.y;
  const int n = q.len[qi];
  const int m = db.len[j];
  float out;
  if (MEAS == TS_EUCLIDEAN && m != n) {
    out = __int_as_float(0x7f800000);
  } else {
    float4 a = q.mbr[qi], b = db.mbr[j];
    float gx = fmaxf(0.0f, fmaxf(b.x - a.y, a.x - b.y));
    float gy = fmaxf(0.0f, fmaxf(b.z - a.w, a.z - b.w));
    out = gx * gx + gy * gy;
  }
  scores[(int64_t)blockIdx.y * db.count + j] = out;
}

// REF filter of the point measures over the DB's length-bucketed tiles: a
// wave takes one tile (64 DB trajectories of nearly equal length, lane l
// the trajectory in slot l) against QT queries of the batch. Each step reads
// point v of all 64 trajectories with one coalesced load per coordinate and
// folds it into the QT pairs of every lane, in order of v per pair (the
// reference's left fold, dtw.cpp:45-58), so each DB point is read once per
// QT queries instead of once per query, with no scattered per-lane streams.
// Envelope reads are per query and lane (one broadcast line when the tile's
// lengths are equal). Square roots are sqrtf's correctly rounded value
// (sqrt_cr), the slow path only for pairs with a tiny coordinate.
template <int MEAS, int QT>
__global__ void __launch_bounds__(256) k_filter_ref_tiled(SetView q, int32_t q0, int32_t nq, const float4* env,
                                                          SetView db, TileView tv, float eps, float* scores) {
  const int lane = threadIdx.x & (kTile - 1);
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= tv.ntiles) return;  // wave-uniform
  const int qb = blockIdx.y * QT;
  const int id = tv.id[(int64_t)tile * kTile + lane];
  const int m = id >= 0 ? db.len[id] : 0;
  const int L = tv.len[tile];
  const float* tx = tv.x + tv.off[tile] + lane;
  const float* ty = tv.y + tv.off[tile] + lane;
  int n[QT], idx[QT], acc[QT], qd[QT], rm[QT];
  float s[QT];
  const float4* e[QT];
  bool tiny = id >= 0 && MEAS != TS_LCSS && db.tiny[id] != 0;
#pragma unroll
  for (int j = 0; j < QT; ++j) {
    const int qi = q0 + min(qb + j, nq - 1);  // past the batch: repeats its last query, never stored
    n[j] = q.len[qi];
    e[j] = env + q.off[qi];
    const int mm = max(m, 1);
    qd[j] = n[j] / mm;
    rm[j] = n[j] - qd[j] * mm;
    idx[j] = 0;
    acc[j] = 0;
    s[j] = 0.0f;
    if (MEAS != TS_LCSS) tiny = tiny || q.tiny[qi] != 0;
  }
  const bool wave_tiny = __any(tiny);
  for (int v = 0; v < L; ++v) {
    const float px = tx[(int64_t)v * kTile], py = ty[(int64_t)v * kTile];
    const bool live = v < m, step = v + 1 < m;
#pragma unroll
    for (int j = 0; j < QT; ++j) {
      const float4 E = e[j][idx[j]];
      if (MEAS == TS_LCSS) {
        const bool in = px > E.x - eps && px < E.y + eps && py > E.z - eps && py < E.w + eps;
        s[j] += (live && in) ? 1.0f : 0.0f;  // an exact count below 2^24
      } else {
        const float dx = fmaxf(0.0f, fmaxf(E.x - px, px - E.y));
        const float dy = fmaxf(0.0f, fmaxf(E.z - py, py - E.w));
        float d[1] = {dx * dx + dy * dy};
        sqrt_cr<1>(d, wave_tiny);
        const float dv = live ? d[0] : 0.0f;  // adding or maxing 0 leaves s unchanged (s >= 0)
        if (MEAS == TS_DTW) s[j] += dv;
        else s[j] = fmaxf(s[j], dv);
      }
      // index map v * n / m, stepped by quotient and remainder up to the
      // trajectory's last point (the envelope read above stays inside the
      // query's n entries on every lane)
      int ni = idx[j] + qd[j], na = acc[j] + rm[j];
      if (na >= m) { na -= m; ++ni; }
      idx[j] = step ? ni : idx[j];
      acc[j] = step ? na : acc[j];
    }
  }
  if (id < 0) return;
  const int64_t N = db.count;
#pragma unroll
  for (int j = 0; j < QT; ++j) {
    if (qb + j < nq) {
      const float out = (MEAS == TS_LCSS) ? s[j] / (float)min(n[j], m) : s[j];
      scores[(int64_t)(qb + j) * N + id] = out;
    }
  }
}

constexpr int kRefQT = 8;

hipError_t launch_filter_ref(hipStream_t st, int32_t measure, SetView q, int32_t q0, int32_t nq,
                             const float4* env, SetView db, const TileView* tiles, float eps, float* scores) {
  if (measure == TS_DTW || measure == TS_FRECHET || measure == TS_LCSS) {
    if (!tiles) return hipErrorInvalidValue;
    if (tiles->ntiles == 0 || nq == 0) return hipSuccess;
    const dim3 grid((tiles->ntiles + 3) / 4, (nq + kRefQT - 1) / kRefQT);
    switch (measure) {
      case TS_DTW:
        hipLaunchKernelGGL((k_filter_ref_tiled<TS_DTW, kRefQT>), grid, dim3(256), 0, st, q, q0, nq, env, db, *tiles,
                           eps, scores);
        break;
      case TS_FRECHET:
        hipLaunchKernelGGL((k_filter_ref_tiled<TS_FRECHET, kRefQT>), grid, dim3(256), 0, st, q, q0, nq, env, db,
                           *tiles, eps, scores);
        break;
      default:
        hipLaunchKernelGGL((k_filter_ref_tiled<TS_LCSS, kRefQT>), grid, dim3(256), 0, st, q, q0, nq, env, db, *tiles,
                           eps, scores);
    }
    return hipGetLastError();
  }
  dim3 grid((db.count + 255) / 256, nq);
  switch (measure) {
    case TS_HAUSDORFF: hipLaunchKernelGGL(k_filter_ref<TS_HAUSDORFF>, grid, dim3(256), 0, st, q, q0, env, db, eps, scores); break;
    case TS_EUCLIDEAN: hipLaunchKernelGGL(k_filter_ref<TS_EUCLIDEAN>, grid, dim3(256), 0, st, q, q0, env, db, eps, scores); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------- EXACT bounds
// Admissible bound of one pair from summaries (see oracle.cpp summary_bound
// for the derivation; the f32 DTW sum is discounted by (L+8)*2^-23).
// Distances: lower bound; LCSS: upper bound of the similarity.
// Returns +inf (distances) / -inf (LCSS) when the pair is not comparable.
struct QSum {
  float4 mbr, ends;
  int n;
  int64_t off;
};

// LCSS box tests against c's eps-inflated box cm (same comparisons as the
// per-point test px - cm.y < eps && cm.x - px < eps && ...).
__device__ __forceinline__ bool lcss_box_out(const float4 b, const float4 cm, float eps) {
  return b.x - cm.y >= eps || cm.x - b.y >= eps || b.z - cm.w >= eps || cm.z - b.w >= eps;
}
__device__ __forceinline__ bool lcss_box_in(const float4 b, const float4 cm, float eps) {
  return b.y - cm.y < eps && cm.x - b.x < eps && b.w - cm.w < eps && cm.z - b.z < eps;
}

// Points of the trajectory at `off` (n points) inside cm inflated by eps.
__device__ int lcss_box_count(const SetView& q, int64_t off, int n, const float4 cm, float eps) {
  const float4* sb = q.sblk + (off >> 3);
  const float4* bl = q.blk + (off >> 3);
  const int nb = (n + 7) >> 3;
  int cnt = 0;
  for (int g = 0; g < nb; g += 8) {
    const float4 s = sb[g];
    if (lcss_box_out(s, cm, eps)) continue;
    if (lcss_box_in(s, cm, eps)) {
      cnt += min(n, 8 * g + 64) - 8 * g;
      continue;
    }
    const int ge = min(nb, g + 8);
    for (int b = g; b < ge; ++b) {
      const float4 bb = bl[b];
      if (lcss_box_out(bb, cm, eps)) continue;
      const int np = min(n - 8 * b, 8);
      if (lcss_box_in(bb, cm, eps)) {
        cnt += np;
        continue;
      }
      for (int i = 0; i < np; ++i) {
        const float px = q.x[off + 8 * b + i], py = q.y[off + 8 * b + i];
        cnt += (px - cm.y < eps && cm.x - px < eps && py - cm.w < eps && cm.z - py < eps) ? 1 : 0;
      }
    }
  }
  return cnt;
}

__device__ __forceinline__ float pair_bound(const MeasureArgs& ma, const QSum& qs, const SetView& q,
                                            const float4 cm, const float4 ce, int m) {
  const float4 a = qs.mbr;
  float gx = fmaxf(0.0f, fmaxf(cm.x - a.y, a.x - cm.y));
  float gy = fmaxf(0.0f, fmaxf(cm.z - a.w, a.z - cm.w));
  float g2 = gx * gx + gy * gy;
  const int n = qs.n;
  switch (ma.measure) {
    case TS_DTW: {
      float dx = qs.ends.x - ce.x, dy = qs.ends.y - ce.y;
      float e = sqrt_lo(dx * dx + dy * dy);
      int L = max(n, m);
      if (L >= 2) {
        float ex = qs.ends.z - ce.z, ey = qs.ends.w - ce.w;
        e = e + sqrt_lo(ex * ex + ey * ey);
      }
      float s = e + (float)max(0, L - 2) * sqrt_lo(g2);
      return s * (1.0f - (float)(L + 8) * 1.1920929e-7f);
    }
    case TS_FRECHET: {
      float dx = qs.ends.x - ce.x, dy = qs.ends.y - ce.y;
      float ex = qs.ends.z - ce.z, ey = qs.ends.w - ce.w;
      return fmaxf(sqrt_lo(g2), fmaxf(sqrt_lo(dx * dx + dy * dy), sqrt_lo(ex * ex + ey * ey)));
    }
    case TS_HAUSDORFF: return sqrt_lo(g2);
    case TS_EUCLIDEAN:
      if (m != n) return __int_as_float(0x7f800000);
      return (float)(sqrt((double)n * (double)g2) * (1.0 - 1e-6));
    case TS_LCSS: {
      // Points of q that can match any point of c lie inside c's
      // eps-inflated MBR. The count descends q's box hierarchy (MBR,
      // 64-point superblocks, 8-point blocks, points): a box wholly outside
      // adds none of its points and one wholly inside adds all of them.
      // f32 subtraction is monotone, so each box test agrees with the test
      // of every point in the box and the count equals the per-point count.
      const float eps = ma.eps;
      if (lcss_box_out(a, cm, eps)) return 0.0f;
      const int cnt = lcss_box_in(a, cm, eps) ? n : lcss_box_count(q, qs.off, n, cm, eps);
      int mn = min(n, m);
      return (float)min(cnt, mn) / (float)mn;
    }
  }
  return 0.0f;
}

// pair_bound with the measure known at compile time (the switch folds).
template <int MEAS>
__device__ __forceinline__ float pair_bound_m(const MeasureArgs& ma, const QSum& qs, const SetView& q,
                                              const float4 cm, const float4 ce, int m) {
  MeasureArgs mc = ma;
  mc.measure = MEAS;
  return pair_bound(mc, qs, q, cm, ce, m);
}

// Seed ranking, where no pair's bound is kept: seeds are refined exactly and
// any exact value bounds the k-th best from above, so their order need not
// come from an admissible bound. DTW: the bound's terms with plain square
// roots and no rounding discount; Frechet: the largest squared term (the
// square root is monotone).
template <int MEAS>
__device__ __forceinline__ float seed_rank(const QSum& qs, const float4 cm, const float4 ce, int m) {
  const float4 a = qs.mbr;
  const float gx = fmaxf(0.0f, fmaxf(cm.x - a.y, a.x - cm.y));
  const float gy = fmaxf(0.0f, fmaxf(cm.z - a.w, a.z - cm.w));
  const float g2 = gx * gx + gy * gy;
  const float dx = qs.ends.x - ce.x, dy = qs.ends.y - ce.y, ex = qs.ends.z - ce.z, ey = qs.ends.w - ce.w;
  if (MEAS == TS_FRECHET) return fmaxf(g2, fmaxf(dx * dx + dy * dy, ex * ex + ey * ey));
  const int L = max(qs.n, m);
  return __builtin_amdgcn_sqrtf(dx * dx + dy * dy) + (L >= 2 ? __builtin_amdgcn_sqrtf(ex * ex + ey * ey) : 0.0f) +
         (float)max(0, L - 2) * __builtin_amdgcn_sqrtf(g2);
}

// The compaction's test of a distance bound, b <= tq. DTW and Frechet first
// try to reject on the MBR-gap term alone, compared squared (no square
// root): the bound is at least (L-2) sqrt_lo(g2) disc (DTW, L = max(n, m))
// or sqrt_lo(g2) (Frechet). With g2 >= 2^-96 (where sqrt_lo is not 0) and
// tq >= 2^-40 (tq^2 normal), a pair whose bound is at most tq has
// ((L-2) disc)^2 g2 (DTW) or g2 (Frechet) at most tq^2 (1 + 2^-17) as f32
// evaluates it (sqrt_lo(g2) >= sqrt(g2) (1 - 2^-20), and each product or sum
// rounds by at most 2^-24), which is below r2 = tq^2 (1 + 2^-16) in f32: the
// shortcut rejects only pairs the bound rejects. r2 = +inf turns it off.
template <int MEAS>
__device__ __forceinline__ bool bound_passes(const MeasureArgs& ma, const QSum& qs, const SetView& q, const float4 cm,
                                             const float4 ce, int m, float tq, float r2) {
  if (MEAS == TS_DTW || MEAS == TS_FRECHET) {
    const float4 a = qs.mbr;
    const float gx = fmaxf(0.0f, fmaxf(cm.x - a.y, a.x - cm.y));
    const float gy = fmaxf(0.0f, fmaxf(cm.z - a.w, a.z - cm.w));
    const float g2 = gx * gx + gy * gy;
    float lhs = g2;
    if (MEAS == TS_DTW) {
      const int L = max(qs.n, m);
      const float c = (float)max(0, L - 2) * (1.0f - (float)(L + 8) * 1.1920929e-7f);
      lhs = c * c * g2;
    }
    if (g2 >= 0x1p-96f && lhs > r2) return false;
  }
  const float b = pair_bound_m<MEAS>(ma, qs, q, cm, ce, m);
  return isfinite(b) && (b <= tq);
}

// Queries per seeding block (the keep-the-best form).
constexpr int kSeedQB = 8;

// DB trajectories per seeding block and thread: 16 when the block keeps the
// best key of each thread (s <= 256: the chunk's blocks fit the chip in one
// round at the 64x100k scale), 64 when the batch still gives 4 blocks per CU
// at that (a quarter of the blocks, so of the per-block sorts: the 4096 x 1M
// DTW seeding went 8.4 -> 5.5 ms), 8 when it sorts every key of the chunk.
constexpr int kSeedPerThread = 16, kSeedPerThreadWide = 64, kSeedPerThreadWhole = 8;
int32_t bound_chunk_items(int32_t s) { return 256 * (s <= 256 ? kSeedPerThread : kSeedPerThreadWhole); }
int32_t bound_chunk_items(int32_t s, int32_t nq, int32_t n_db) {
  if (s > 256) return 256 * kSeedPerThreadWhole;
  const int64_t wide_blocks = (int64_t)((n_db + 256 * kSeedPerThreadWide - 1) / (256 * kSeedPerThreadWide)) *
                              ((nq + kSeedQB - 1) / kSeedQB);
  return 256 * (wide_blocks >= 4ll * device_cu_count() ? kSeedPerThreadWide : kSeedPerThread);
}

// Bitonic sort of `nseg` consecutive segments of `seg` keys each (seg a
// power of two), every segment ascending: compare partners and directions
// below the last stage depend only on an index's low bits, so all segments
// sort together; the last stage merges each segment ascending.
__device__ void bitonic_sort_segments(uint64_t* s, int seg, int nseg, int tid, int nthreads) {
  const int n = seg * nseg;
  for (int k = 2; k <= seg; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < n; i += nthreads) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const uint64_t a = s[i], b = s[ixj];
          const bool asc = k == seg || (i & k) == 0;
          if ((a > b) == asc) { s[i] = b; s[ixj] = a; }
        }
      }
      __syncthreads();
    }
  }
}

// Seeds: good candidates per (query, DB chunk of bound_chunk_items(s)
// trajectories). Any real pair's exact value bounds the k-th best from above,
// so seeds need not be the exact top-s by bound: each thread keeps the best
// key of its kSeedPerThread trajectories and the block keeps the best `s` of
// those (s <= 256), for QB queries at once: every DB summary (36 B) is read
// once per QB queries. Larger s sorts the whole chunk, one query per block.
// RANK: seed_rank instead of the bound (DTW, Frechet; no bstore).
template <bool WHOLE, int MEAS, int QB, int PER, bool RANK>
__global__ void __launch_bounds__(256) k_bound_chunks(MeasureArgs ma, SetView q, int32_t q0, int32_t nq, SetView db,
                                                      int32_t s, uint64_t* keys, float* bstore, int32_t* zero,
                                                      int32_t nzero) {
  static_assert(!WHOLE || QB == 1, "whole-chunk sorts take one query per block");
  constexpr int NS = WHOLE ? 256 * PER : 256;
  __shared__ uint64_t sk[NS * QB];
  // counters the batch's later kernels expect to start at zero
  if (blockIdx.x == 0 && blockIdx.y == 0 && (int)threadIdx.x < nzero) zero[threadIdx.x] = 0;
  const int qr0 = blockIdx.y * QB;  // first query of the block, relative to q0
  QSum qs[QB];
#pragma unroll
  for (int j = 0; j < QB; ++j) {
    const int qi = q0 + min(qr0 + j, nq - 1);  // past the batch: repeats its last query, never stored
    qs[j] = QSum{q.mbr[qi], q.ends[qi], q.len[qi], q.off[qi]};
  }
  const bool desc = ma.measure == TS_LCSS;
  const int base = blockIdx.x * 256 * PER;
  uint64_t best[QB];
#pragma unroll
  for (int j = 0; j < QB; ++j) best[j] = ~0ull;
  // eight trajectories' summaries in flight at a time
#pragma unroll 1
  for (int t0 = 0; t0 < PER; t0 += 8) {
#pragma unroll
    for (int t = t0; t < t0 + 8; ++t) {
      const int i = t * 256 + threadIdx.x;
      const int jd = base + i;
      const bool ok = jd < db.count;
      const float4 cm = ok ? db.mbr[jd] : make_float4(0.f, 0.f, 0.f, 0.f);
      const float4 ce = ok ? db.ends[jd] : make_float4(0.f, 0.f, 0.f, 0.f);
      const int m = ok ? db.len[jd] : 1;
#pragma unroll
      for (int j = 0; j < QB; ++j) {
        uint64_t key = ~0ull;
        if (ok) {
          const float b = RANK ? seed_rank<MEAS>(qs[j], cm, ce, m) : pair_bound_m<MEAS>(ma, qs[j], q, cm, ce, m);
          if (!RANK && bstore && qr0 + j < nq) bstore[(int64_t)(qr0 + j) * db.count + jd] = b;  // for the count pass
          key = make_key(b, jd, desc);
        }
        if (WHOLE) sk[i] = key;
        else best[j] = key < best[j] ? key : best[j];
      }
    }
  }
  if (!WHOLE) {
#pragma unroll
    for (int j = 0; j < QB; ++j) sk[j * NS + threadIdx.x] = best[j];
  }
  __syncthreads();
  bitonic_sort_segments(sk, NS, QB, threadIdx.x, 256);
#pragma unroll
  for (int j = 0; j < QB; ++j) {
    if (qr0 + j >= nq) break;
    uint64_t* out = keys + ((int64_t)(qr0 + j) * gridDim.x + blockIdx.x) * s;
    for (int t = threadIdx.x; t < s; t += 256) out[t] = t < NS ? sk[j * NS + t] : ~0ull;
  }
}

hipError_t launch_bound_chunks(hipStream_t st, MeasureArgs m, int32_t k, SetView q, int32_t q0, int32_t nq,
                               SetView db, int32_t s, int32_t chunk, uint64_t* keys, float* bstore, int32_t* zero,
                               int32_t nzero) {
  (void)k;
  if (chunk != bound_chunk_items(s, nq, db.count)) return hipErrorInvalidValue;
  const int32_t gx = (db.count + chunk - 1) / chunk;
  const dim3 grid_qb(gx, (nq + kSeedQB - 1) / kSeedQB);
  const bool rank = !bstore && (m.measure == TS_DTW || m.measure == TS_FRECHET);
#define TS_CHUNKS_R(MEAS, RANK)                                                                                  \
  if (s > 256)                                                                                                   \
    hipLaunchKernelGGL((k_bound_chunks<true, MEAS, 1, kSeedPerThreadWhole, RANK>), dim3(gx, nq), dim3(256), 0, st, \
                       m, q, q0, nq, db, s, keys, bstore, zero, nzero);                                          \
  else if (chunk == 256 * kSeedPerThreadWide)                                                                    \
    hipLaunchKernelGGL((k_bound_chunks<false, MEAS, kSeedQB, kSeedPerThreadWide, RANK>), grid_qb, dim3(256), 0,  \
                       st, m, q, q0, nq, db, s, keys, bstore, zero, nzero);                                      \
  else                                                                                                           \
    hipLaunchKernelGGL((k_bound_chunks<false, MEAS, kSeedQB, kSeedPerThread, RANK>), grid_qb, dim3(256), 0, st,  \
                       m, q, q0, nq, db, s, keys, bstore, zero, nzero);
#define TS_CHUNKS(MEAS)                   \
  if constexpr (MEAS == TS_DTW || MEAS == TS_FRECHET) { \
    if (rank) {                           \
      TS_CHUNKS_R(MEAS, true)             \
    } else {                              \
      TS_CHUNKS_R(MEAS, false)            \
    }                                     \
  } else {                                \
    TS_CHUNKS_R(MEAS, false)              \
  }
  switch (m.measure) {
    case TS_DTW: TS_CHUNKS(TS_DTW) break;
    case TS_FRECHET: TS_CHUNKS(TS_FRECHET) break;
    case TS_LCSS: TS_CHUNKS(TS_LCSS) break;
    case TS_HAUSDORFF: TS_CHUNKS(TS_HAUSDORFF) break;
    case TS_EUCLIDEAN: TS_CHUNKS(TS_EUCLIDEAN) break;
    default: return hipErrorInvalidValue;
  }
#undef TS_CHUNKS
#undef TS_CHUNKS_R
  return hipGetLastError();
}

// Threshold compaction, QB queries per block and kCompactJ DB
// trajectories per thread: each DB summary is loaded once and tested against
// every query of the block. A pair survives when its bound can still reach
// the query's k-th value tau: distances bound <= tau; LCSS ub >= tau and,
// because zero-similarity ties are broken by index, (ub > 0 or j < k).
// Two passes, no global atomics (one hot counter per query serialises in
// L2): COUNT reduces each block's survivors per query in LDS and stores them
// in bcounts[q * gridDim.x + block], and each wave's pass mask (one bit per
// DB trajectory) in bits[q * words + word]; after k_block_prefix has turned
// the counts into exclusive offsets within each query's region, WRITE reads
// the masks back (no bound is evaluated twice) and places the block's
// survivors at off[q] + bcounts[...] + an LDS cursor (slot = position).
// The count pass takes 16 queries per block (half the DB summary reads of 8:
// 4096 x 1M DTW 4.97 -> 4.47 ms), the write pass 8 (2.83 -> 2.53 ms).
constexpr int kCompactQCount = 16, kCompactQWrite = 8, kCompactJ = 2;

int32_t compact_blocks(int32_t n_db) { return (n_db + 256 * kCompactJ - 1) / (256 * kCompactJ); }
int64_t compact_mask_words(int32_t n_db) { return (int64_t)compact_blocks(n_db) * kCompactJ * 4; }

template <bool WRITE, int MEAS, int QB = WRITE ? kCompactQWrite : kCompactQCount>
__global__ void __launch_bounds__(256) k_bound_compact(MeasureArgs ma, int32_t k, SetView q, int32_t q0, int32_t nq,
                                                       SetView db, const float* tau, const int64_t* off,
                                                       const float* bvals, int32_t* bcounts, uint64_t* bits,
                                                       Item* items, int64_t cap) {
  __shared__ QSum sq[QB];
  __shared__ float st[QB], sr2[QB];
  __shared__ int bc[QB];
  __shared__ int64_t bbase[QB];
  const int qb = blockIdx.y * QB;
  const int nqb = min(QB, nq - qb);
  const int gx = gridDim.x, bx = blockIdx.x;
  if (threadIdx.x < nqb) {
    const int qi = q0 + qb + threadIdx.x;
    sq[threadIdx.x] = QSum{q.mbr[qi], q.ends[qi], q.len[qi], q.off[qi]};
    const float tq = tau[qb + threadIdx.x];
    st[threadIdx.x] = tq;
    sr2[threadIdx.x] = tq >= 0x1p-40f ? tq * tq * (1.0f + 0x1p-16f) : __int_as_float(0x7f800000);  // see bound_passes
    bc[threadIdx.x] = 0;
    if (WRITE) bbase[threadIdx.x] = off[qb + threadIdx.x] + bcounts[(int64_t)(qb + threadIdx.x) * gx + bx];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t words = (int64_t)gx * kCompactJ * 4;
  for (int t = 0; t < kCompactJ; ++t) {
    const int j = (blockIdx.x * kCompactJ + t) * 256 + threadIdx.x;
    const int64_t word = (int64_t)(blockIdx.x * kCompactJ + t) * 4 + (threadIdx.x >> 6);
    const bool valid = j < db.count;
    const int jj = valid ? j : 0;
    float4 cm = make_float4(0.0f, 0.0f, 0.0f, 0.0f), ce = cm;
    int m = 0;
    if (!WRITE && !bvals) {
      cm = db.mbr[jj];
      ce = db.ends[jj];
      m = db.len[jj];
    }
    // every load of the block's queries up front (independent, so their
    // latencies overlap)
    uint64_t wbits[QB];
    float sb[QB];
#pragma unroll
    for (int u = 0; u < QB; ++u) {
      if (WRITE) wbits[u] = u < nqb ? bits[(int64_t)(qb + u) * words + word] : 0ull;
      else sb[u] = (bvals && valid && u < nqb) ? bvals[(int64_t)(qb + u) * db.count + j] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < QB; ++u) {
      if (u >= nqb) break;
      uint64_t bal;
      bool pass;
      if (WRITE) {
        bal = wbits[u];
        pass = (bal >> lane) & 1ull;
      } else {
        pass = false;
        if (valid) {
          const float tq = st[u];
          if (MEAS == TS_LCSS) {
            const float b = bvals ? sb[u] : pair_bound_m<MEAS>(ma, sq[u], q, cm, ce, m);
            pass = (b >= tq) && (b > 0.0f || j < k);
          } else if (bvals) {
            pass = isfinite(sb[u]) && (sb[u] <= tq);
          } else {
            pass = bound_passes<MEAS>(ma, sq[u], q, cm, ce, m, tq, sr2[u]);
          }
        }
        bal = __ballot(pass);
        if (lane == 0) bits[(int64_t)(qb + u) * words + word] = bal;
      }
      if (bal == 0) continue;
      const int leader = __ffsll((unsigned long long)bal) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(&bc[u], __popcll(bal));  // LDS
      if (WRITE) {
        base = __shfl(base, leader);
        if (pass) {
          const int64_t pos = bbase[u] + base + __popcll(bal & ((1ull << lane) - 1ull));
          if (pos < cap) items[pos] = Item{q0 + qb + u, j, (int32_t)pos, 0};
        }
      }
    }
  }
  if (!WRITE) {
    __syncthreads();
    if (threadIdx.x < nqb) bcounts[(int64_t)(qb + threadIdx.x) * gx + bx] = bc[threadIdx.x];
  }
}

// Per query (one block each): exclusive prefix of its gx block counts in
// place, and the query's total into counts[q].
__global__ void __launch_bounds__(1024) k_block_prefix(int32_t* bcounts, int32_t gx, int32_t* counts) {
  __shared__ int32_t part[1024];
  int32_t* row = bcounts + (int64_t)blockIdx.x * gx;
  const int per = (gx + 1023) / 1024;
  const int b = threadIdx.x * per;
  int32_t sum = 0;
  for (int i = b; i < min(gx, b + per); ++i) sum += row[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const int32_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int32_t run = part[threadIdx.x] - sum;
  for (int i = b; i < min(gx, b + per); ++i) {
    const int32_t c = row[i];
    row[i] = run;
    run += c;
  }
  if (threadIdx.x == 1023) counts[blockIdx.x] = part[1023];
  // counts[nq] and counts[nq + 1] are the tight filter's and its
  // pre-check's survivor counters, cleared here
  if (blockIdx.x == 0 && threadIdx.x < 2) counts[gridDim.x + threadIdx.x] = 0;
}

hipError_t launch_block_prefix(hipStream_t st, int32_t* bcounts, int32_t nq, int32_t gx, int32_t* counts) {
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(k_block_prefix, dim3(nq), dim3(1024), 0, st, bcounts, gx, counts);
  return hipGetLastError();
}

hipError_t launch_bound_compact(hipStream_t st, MeasureArgs m, int32_t k, SetView q, int32_t q0, int32_t nq,
                                SetView db, const float* tau, const int64_t* off, const float* bvals, int32_t* counts,
                                uint64_t* bits, Item* items, int64_t cap) {
  const dim3 grid_w(compact_blocks(db.count), (nq + kCompactQWrite - 1) / kCompactQWrite);
  const dim3 grid_c(compact_blocks(db.count), (nq + kCompactQCount - 1) / kCompactQCount);
#define TS_COMPACT(MEAS)                                                                                          \
  if (items)                                                                                                      \
    hipLaunchKernelGGL((k_bound_compact<true, MEAS>), grid_w, dim3(256), 0, st, m, k, q, q0, nq, db, tau, off,    \
                       bvals, counts, bits, items, cap);                                                          \
  else                                                                                                            \
    hipLaunchKernelGGL((k_bound_compact<false, MEAS>), grid_c, dim3(256), 0, st, m, k, q, q0, nq, db, tau, off,   \
                       bvals, counts, bits, items, cap);
  switch (m.measure) {
    case TS_DTW: TS_COMPACT(TS_DTW) break;
    case TS_FRECHET: TS_COMPACT(TS_FRECHET) break;
    case TS_LCSS: TS_COMPACT(TS_LCSS) break;
    case TS_HAUSDORFF: TS_COMPACT(TS_HAUSDORFF) break;
    case TS_EUCLIDEAN: TS_COMPACT(TS_EUCLIDEAN) break;
    default: return hipErrorInvalidValue;
  }
#undef TS_COMPACT
  return hipGetLastError();
}

// ------------------------------------------------------ tight bound stage
// Second filter over the pairs that survived the summary bound, O(n+m) per
// pair and one wave per pair. Every warping path / coupling visits every row
// and every column, and a cell's computed cost is never below its point's
// computed distance to the other trajectory's MBR (f32 subtraction is
// monotone). Hence, with d(p, B) the point-to-box distance:
//   DTW       >= max( d(q0,c0) + sum_{0<i<n-1} d(q_i, B_c) + d(q_n-1,c_m-1),
//                     the same over c's points against B_q ),
//             discounted for the f32 sums (dtw_disc: tree order here,
//             running order in the DP); without a band, B_c is the nearest
//             of the boxes that tile c (tile_pass2), with a band the box of
//             c's points inside row i's band (band_side_sum), and likewise;
//   Frechet   >= max( max_i d(q_i,B_c), max_j d(c_j,B_q), endpoint costs );
//   Hausdorff >= max( max_i d(q_i,B_c), max_j d(c_j,B_q) );
//             and then, with B_c the nearest of the 64-point tiles of c
//             instead of its MBR, the same forms again (tile_side_*);
//   Euclidean >= sqrt(sum_i d(q_i,B_c)^2) (double, discounted 1e-6);
//   LCSS      <= min(#q points in c's eps-box, #c points in q's eps-box,
//                    min(n,m)) / min(n,m); with a window delta, the box of
//                    the other trajectory's points within delta of the
//                    point's index (window_side_count).
// min/max do not round, so the Frechet/Hausdorff/LCSS forms are exact.
__device__ __forceinline__ float wave_sum(float v) {
  for (int sh = 32; sh > 0; sh >>= 1) v += __shfl_xor(v, sh);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int sh = 32; sh > 0; sh >>= 1) v = fmaxf(v, __shfl_xor(v, sh));
  return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
  for (int sh = 32; sh > 0; sh >>= 1) v += __shfl_xor(v, sh);
  return v;
}

// Sum/max over the points of `a` (length na) of their distance to box `b`,
// skipping the first and last point when `skip_ends`.
template <int MEAS>
__device__ void side_bound(const float* ax, const float* ay, int na, float4 b, float eps, bool skip_ends, float& acc,
                           double& accd, int& cnt) {
  const int lane = threadIdx.x & 63;
  for (int i = lane; i < na; i += 64) {
    const float px = ax[i], py = ay[i];
    if (MEAS == TS_LCSS) {
      cnt += (px - b.y < eps && b.x - px < eps && py - b.w < eps && b.z - py < eps) ? 1 : 0;
    } else if (MEAS == TS_EUCLIDEAN) {
      accd += (double)pt_box(px, py, b);
    } else {
      const float d2 = pt_box(px, py, b);
      if (MEAS == TS_DTW) {
        if (!(skip_ends && (i == 0 || i == na - 1))) acc += sqrt_lo(d2);
      } else {
        acc = fmaxf(acc, d2);
      }
    }
  }
}

// Banded DTW form: row i of `a` (na points) has a path cell only in the
// columns j of `b` (nb points) with |i*nb - j*na| <= W (W = w*max(na,nb)),
// so it costs at least the distance from a_i to the box of b's points there;
// the union of the 64-point superblock and 8-point block boxes covering that
// window contains it. Sum over the points of a except the first and last,
// 64 rows at a time; returns early (the partial sum) once that partial,
// plus `base` and scaled by `disc`, exceeds `limit`.
__device__ float band_side_sum(const float* ax, const float* ay, int na, const float4* bblk, const float4* bsblk,
                               int nb, int64_t W, float base, float disc, float limit) {
  const int lane = threadIdx.x & 63;
  float total = 0.0f;
  for (int i0 = 1; i0 < na - 1; i0 += 64) {
    const int i = i0 + lane;
    float acc = 0.0f;
    if (i < na - 1) {
      const int64_t lo = (int64_t)i * nb - W;
      const int jlo = lo <= 0 ? 0 : (int)((lo + na - 1) / na);
      const int jhi = (int)min<int64_t>(nb - 1, ((int64_t)i * nb + W) / na);
      int kb = jlo >> 3;
      const int ke = jhi >> 3;
      float4 bx = bblk[kb];
      for (; kb + 7 <= ke; kb += 8) {  // superblock kb covers blocks kb..kb+7
        const float4 t = bsblk[kb];
        bx.x = fminf(bx.x, t.x); bx.y = fmaxf(bx.y, t.y); bx.z = fminf(bx.z, t.z); bx.w = fmaxf(bx.w, t.w);
      }
      for (; kb <= ke; ++kb) {
        const float4 t = bblk[kb];
        bx.x = fminf(bx.x, t.x); bx.y = fmaxf(bx.y, t.y); bx.z = fminf(bx.z, t.z); bx.w = fmaxf(bx.w, t.w);
      }
      acc = sqrt_lo(pt_box(ax[i], ay[i], bx));
    }
    total += wave_sum(acc);
    if ((total + base) * disc > limit) break;  // uniform: total is the same in every lane
  }
  return total;
}

// Sum over the wave (all lanes active) by DPP: prefix sums within rows of 16,
// rows combined by row_bcast 15 and 31; lane 63's total returned to every lane.
template <int CTRL, int ROWS>
__device__ __forceinline__ float add_dpp(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, false));
}
__device__ __forceinline__ float wave_sum_dpp(float v) {
  v = add_dpp<0x111, 0xf>(v);
  v = add_dpp<0x112, 0xf>(v);
  v = add_dpp<0x114, 0xf>(v);
  v = add_dpp<0x118, 0xf>(v);
  v = add_dpp<0x142, 0xa>(v);
  v = add_dpp<0x143, 0xc>(v);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Boxes that tile a trajectory of n points, into the wave's LDS slice `dst`
// (one per lane, at most 64): its 64-point superblocks (every 8th entry of a
// set's sblk), or for trajectories of more than 4096 points the unions of
// g = 2, 4, ... consecutive superblocks (a coarser tiling; every box still
// lies inside the MBR). Returns the number of boxes.
constexpr int kTileBoxes = 64;
// Superblocks per tile box of an n-point trajectory.
__device__ __forceinline__ int tile_group(int n) {
  const int ns = (((n + 7) >> 3) + 7) >> 3;
  int g = 1;
  while (ns > g * kTileBoxes) g <<= 1;
  return g;
}
__device__ int stage_tiles(const float4* sblk, int n, float4* dst) {
  const int lane = threadIdx.x & 63;
  const int ns = (((n + 7) >> 3) + 7) >> 3;
  const int g = tile_group(n);
  const int nb = (ns + g - 1) / g;
  if (lane < nb) {
    float4 b = sblk[8 * lane * g];
    for (int u = lane * g + 1; u < min(ns, (lane + 1) * g); ++u) {
      const float4 t = sblk[8 * u];
      b.x = fminf(b.x, t.x); b.y = fmaxf(b.y, t.y); b.z = fminf(b.z, t.z); b.w = fmaxf(b.w, t.w);
    }
    dst[lane] = b;
  }
  return nb;
}

// Tiled form of both sides of an unbanded DTW pair in one pass: the interior
// points of a (all but the first and last) against the nearest of the boxes
// that tile b, summed into sa, and those of b against a's into sb (row i of a
// has a path cell somewhere in b, in a row of its own). The nearest tile box
// lies inside the MBR, so each side also bounds the MBR form from above. The
// boxes sit in the wave's LDS slice (abox, bbox:
// nsa and nsb of them), read by all lanes at once; both trajectories' next
// 64-point chunk is loaded before this chunk's boxes are scanned (indices
// clamped to the last point, so every load stays inside the trajectory).
// After each chunk both partial sums are reduced; the pair fails as soon as
// the larger, plus the endpoint costs e and scaled by disc, exceeds t: the
// partial sums of non-negative terms never exceed the full sums. The f32
// sums are in a different order than a serial sum; disc covers it.
__device__ bool tile_pass2(const float* ax, const float* ay, int na, const float* bx, const float* by, int nb,
                           const float4* abox, int nsa, const float4* bbox, int nsb, float e, float disc, float t,
                           float* bound) {
  const int lane = threadIdx.x & 63;
  const int L = max(na, nb);
  float sa = 0.0f, sb = 0.0f;
  float px = ax[min(lane, na - 1)], py = ay[min(lane, na - 1)];
  float rx = bx[min(lane, nb - 1)], ry = by[min(lane, nb - 1)];
  for (int i0 = 0; i0 < L; i0 += 64) {
    const int i = i0 + lane;
    const int ja = min(i + 64, na - 1), jb = min(i + 64, nb - 1);
    const float npx = ax[ja], npy = ay[ja], nrx = bx[jb], nry = by[jb];
    float da = 3.4e38f, db = 3.4e38f;
    if (i0 < na - 1) {
#pragma unroll 4
      for (int s = 0; s < nsb; ++s) da = fminf(da, pt_box(px, py, bbox[s]));
    }
    if (i0 < nb - 1) {
#pragma unroll 4
      for (int s = 0; s < nsa; ++s) db = fminf(db, pt_box(rx, ry, abox[s]));
    }
    sa += wave_sum_dpp(i >= 1 && i < na - 1 ? sqrt_lo(da) : 0.0f);
    sb += wave_sum_dpp(i >= 1 && i < nb - 1 ? sqrt_lo(db) : 0.0f);
    if (!((fmaxf(sa, sb) + e) * disc <= t)) return false;  // uniform: the sums are the same in every lane
    px = npx;
    py = npy;
    rx = nrx;
    ry = nry;
  }
  *bound = (fmaxf(sa, sb) + e) * disc;
  return true;
}

// Tiled max form (Frechet, Hausdorff): every point of a is matched to some
// point of b, so max_i of the squared distance from a_i to the nearest tile
// box of b bounds the squared value from below. Stops early once the running
// maximum exceeds limit2.
__device__ float tile_side_max(const float* ax, const float* ay, int na, const float4* bsblk, int nb, float limit2) {
  const int lane = threadIdx.x & 63;
  const int nblk = (nb + 7) >> 3;
  float mx = 0.0f;
  for (int i0 = 0; i0 < na; i0 += 64) {
    const int i = i0 + lane;
    if (i < na) {
      const float px = ax[i], py = ay[i];
      float best = pt_box(px, py, bsblk[0]);
      for (int kb = 8; kb < nblk; kb += 8) best = fminf(best, pt_box(px, py, bsblk[kb]));
      mx = fmaxf(mx, best);
    }
    if (wave_max(mx) > limit2) break;
  }
  return wave_max(mx);
}

// Tiled LCSS count: points of a inside the eps-inflated box of some tile of b.
__device__ int tile_side_count(const float* ax, const float* ay, int na, const float4* bsblk, int nb, float eps) {
  const int lane = threadIdx.x & 63;
  const int nblk = (nb + 7) >> 3;
  int cnt = 0;
  for (int i = lane; i < na; i += 64) {
    const float px = ax[i], py = ay[i];
    bool in = false;
    for (int kb = 0; kb < nblk && !in; kb += 8) {
      const float4 bx = bsblk[kb];
      in = px - bx.y < eps && bx.x - px < eps && py - bx.w < eps && bx.z - py < eps;
    }
    cnt += in ? 1 : 0;
  }
  for (int sh = 32; sh > 0; sh >>= 1) cnt += __shfl_xor(cnt, sh);
  return cnt;
}

// LCSS window form: point i of `a` can only match points j of `b` with
// |i - j| <= delta, so it counts only if it lies inside the eps-inflated box
// of b's points there (superblocks + blocks covering the window; f32
// subtraction is monotone, so every real match passes this test).
__device__ int window_side_count(const float* ax, const float* ay, int na, const float4* bblk, const float4* bsblk,
                                 int nb, int delta, float eps) {
  const int lane = threadIdx.x & 63;
  int cnt = 0;
  for (int i = lane; i < na; i += 64) {
    const int jlo = max(0, i - delta), jhi = min(nb - 1, i + delta);
    if (jlo > jhi) continue;
    int kb = jlo >> 3;
    const int ke = jhi >> 3;
    float4 bx = bblk[kb];
    for (; kb + 7 <= ke; kb += 8) {
      const float4 t = bsblk[kb];
      bx.x = fminf(bx.x, t.x); bx.y = fmaxf(bx.y, t.y); bx.z = fminf(bx.z, t.z); bx.w = fmaxf(bx.w, t.w);
    }
    for (; kb <= ke; ++kb) {
      const float4 t = bblk[kb];
      bx.x = fminf(bx.x, t.x); bx.y = fmaxf(bx.y, t.y); bx.z = fminf(bx.z, t.z); bx.w = fmaxf(bx.w, t.w);
    }
    const float px = ax[i], py = ay[i];
    cnt += (px - bx.y < eps && bx.x - px < eps && py - bx.w < eps && bx.z - py < eps) ? 1 : 0;
  }
  for (int sh = 32; sh > 0; sh >>= 1) cnt += __shfl_xor(cnt, sh);
  return cnt;
}

// Survivors are staged in an LDS buffer per block and appended to `out` with
// one atomic per block at the end (a single hot counter per pair serialises
// in L2); a full buffer sends the rest directly.
constexpr int kTightBuf = 256;

// The factor that discounts a DTW bound of pairs up to L points per side for
// f32 rounding: the DP's running sum along a path of at most 2L-1 cells can
// fall short of the exact sum by (L-1)*2^-23 of it, the bound's sums (64-lane
// reductions chained over L/64 chunks, then the endpoint costs) can exceed
// theirs by (L/64 + 9)*2^-24, and the final product rounds once more.
__device__ __forceinline__ float dtw_disc(int L) {
  return 1.0f - (float)(L + (L >> 6) + 72) * 1.1920929e-7f;
}

// Squared gap between two boxes (minx, maxx, miny, maxy): for points p of
// one and b of the other, the f32 squared distance from p to b is at least
// this (the f32 operations are monotone).
__device__ __forceinline__ float box_gap2(float4 a, float4 b) {
  const float gx = fmaxf(0.0f, fmaxf(b.x - a.y, a.x - b.y));
  const float gy = fmaxf(0.0f, fmaxf(b.z - a.w, a.z - b.w));
  return gx * gx + gy * gy;
}

// Tiles of a trajectory against the other trajectory's MBR (one wave): every
// interior point (not the first or last) of an n-point trajectory costs at
// least its tile's gap to the other MBR, so the sum over tiles of interior
// points times gap bounds the side's point sum of the tiled form from below.
// `boxes`: tiles of `span` points each (the last shorter), nb of them.
__device__ __forceinline__ float tiles_vs_mbr(const float4* boxes, int nb, int span, int n, float4 other) {
  const int lane = threadIdx.x & 63;
  float v = 0.0f;
  for (int t = lane; t < nb; t += 64) {
    const int lo = max(t * span, 1), hi = min(t * span + span, n - 1);
    if (hi > lo) v += (float)(hi - lo) * sqrt_lo(box_gap2(boxes[t], other));
  }
  return wave_sum_dpp(v);
}

// Largest superblock count (both trajectories) of a pair that k_tight_pre
// tests (see there).
constexpr int kPreBoxes = 48;

// TILED: unbanded DTW, every pair by the one-pass tiled form. pre: the pairs
// have passed k_tight_pre, whose superblock test equals the staged-tile test
// below for pairs it tested (tiles are superblocks up to 4096 points).
template <int MEAS, bool TILED>
__global__ void __launch_bounds__(256) k_tight_filter(MeasureArgs ma, int32_t k, SetView q, int32_t q0, SetView db,
                                                      const float* tau, const Item* items, const int32_t* nitems,
                                                      int64_t max_items, Item* out, int32_t* nout, float* rej,
                                                      float* tb, const int32_t* seed_idx, int32_t s, bool pre) {
  static_assert(!TILED || MEAS == TS_DTW, "the tiled pass is DTW's");
  __shared__ Item sbuf[kTightBuf];
  __shared__ int scnt, sbase;
  // each wave's slice: the query's tile boxes, then the DB trajectory's
  __shared__ float4 sbox[TILED ? 4 : 1][2 * kTileBoxes];
  if (threadIdx.x == 0) scnt = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t n_items = nitems ? min<int64_t>(max_items, (int64_t)*nitems) : max_items;
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); w < n_items; w += nw) {
    const Item it = items[w];
    if (it.d < 0) {
      if (lane == 0) {
        rej[it.slot] = __int_as_float(0x7fc00000);
        if (tb) tb[it.slot] = __int_as_float(0x7fc00000);
      }
      continue;
    }
    const int n = q.len[it.q], m = db.len[it.d];
    const int64_t qo = q.off[it.q], co = db.off[it.d];
    const float4 qb = q.mbr[it.q], cb = db.mbr[it.d];
    const float4 qe = q.ends[it.q], ce = db.ends[it.d];
    const float t = tau[it.q - q0];
    bool pass;
    float bval = 0.0f;  // the bound compared with t (DTW and Frechet)
    if (MEAS == TS_LCSS) {
      int c1 = 0, c2 = 0;
      float f = 0.0f;
      double fd = 0.0;
      side_bound<MEAS>(q.x + qo, q.y + qo, n, cb, ma.eps, false, f, fd, c1);
      side_bound<MEAS>(db.x + co, db.y + co, m, qb, ma.eps, false, f, fd, c2);
      for (int sh = 32; sh > 0; sh >>= 1) {
        c1 += __shfl_xor(c1, sh);
        c2 += __shfl_xor(c2, sh);
      }
      const int mn = min(n, m);
      const float ub = (float)min(min(c1, c2), mn) / (float)mn;
      pass = (ub >= t) && (ub > 0.0f || it.d < k);
      if (pass && ma.delta < 0 && ub > 0.0f) {
        // count against the tiles of the other trajectory instead of its MBR
        const int w1 = tile_side_count(q.x + qo, q.y + qo, n, db.sblk + (co >> 3), m, ma.eps);
        const int w2 = tile_side_count(db.x + co, db.y + co, m, q.sblk + (qo >> 3), n, ma.eps);
        const float ubt = (float)min(min(w1, w2), mn) / (float)mn;
        pass = (ubt >= t) && (ubt > 0.0f || it.d < k);
      }
      if (pass && ma.delta >= 0 && ub > 0.0f) {
        // with an index window, count against the window's box instead
        const int w1 = window_side_count(q.x + qo, q.y + qo, n, db.blk + (co >> 3), db.sblk + (co >> 3), m,
                                         ma.delta, ma.eps);
        const int w2 = window_side_count(db.x + co, db.y + co, m, q.blk + (qo >> 3), q.sblk + (qo >> 3), n,
                                         ma.delta, ma.eps);
        const float ubw = (float)min(min(w1, w2), mn) / (float)mn;
        pass = (ubw >= t) && (ubw > 0.0f || it.d < k);
      }
    } else if (MEAS == TS_EUCLIDEAN) {
      float f = 0.0f;
      double s = 0.0;
      int c = 0;
      if (n != m) {
        if (lane == 0) rej[it.slot] = __int_as_float(0x7fc00000);
        continue;
      }
      side_bound<MEAS>(q.x + qo, q.y + qo, n, cb, ma.eps, false, f, s, c);
      s = wave_sum_d(s);
      pass = (float)(sqrt(s) * (1.0 - 1e-6)) <= t;
    } else if (TILED) {
      // the tiled form of both sides (it implies the MBR form). Rows 0 and
      // n-1 hold the cells (0,0) and (n-1,m-1), distinct unless n = m = 1.
      float4* wbox = sbox[threadIdx.x >> 6];
      const int nsa = stage_tiles(q.sblk + (qo >> 3), n, wbox);
      const int nsc = stage_tiles(db.sblk + (co >> 3), m, wbox + kTileBoxes);
      __builtin_amdgcn_wave_barrier();
      const float dx0 = qe.x - ce.x, dy0 = qe.y - ce.y, dx1 = qe.z - ce.z, dy1 = qe.w - ce.w;
      const float e0 = sqrt_lo(dx0 * dx0 + dy0 * dy0);
      const float e = n >= 2 || m >= 2 ? e0 + sqrt_lo(dx1 * dx1 + dy1 * dy1) : e0;
      const float disc = dtw_disc(max(n, m));
      pass = true;
      if (n >= 3 && m >= 3) {
        // Cheaper forms first, each implied by the point form of its side
        // against the other's MBR: the staged tiles against the other MBR
        // (no point read), then the 8-point block boxes against it. On the
        // 4096 x 1M DTW call they leave about 25% and 14% of the pairs to
        // the point pass. Rows 0 and n-1 hold the cells (0,0) and
        // (n-1,m-1), priced by e.
        if (!(pre && ((n + 63) >> 6) + ((m + 63) >> 6) <= kPreBoxes)) {
          const float ta = tiles_vs_mbr(wbox, nsa, 64 * tile_group(n), n, cb);
          const float tc = tiles_vs_mbr(wbox + kTileBoxes, nsc, 64 * tile_group(m), m, qb);
          pass = (fmaxf(ta, tc) + e) * disc <= t;
        }
        if (pass) {
          const float ba = tiles_vs_mbr(q.blk + (qo >> 3), (n + 7) >> 3, 8, n, cb);
          const float bc = tiles_vs_mbr(db.blk + (co >> 3), (m + 7) >> 3, 8, m, qb);
          pass = (fmaxf(ba, bc) + e) * disc <= t;
        }
      }
      if (pass)
        pass = tile_pass2(q.x + qo, q.y + qo, n, db.x + co, db.y + co, m, wbox, nsa, wbox + kTileBoxes, nsc, e,
                          disc, t, &bval);
      __builtin_amdgcn_wave_barrier();  // every lane's box reads before the next pair's writes
    } else if (MEAS == TS_DTW) {
      float a1 = 0.0f, a2 = 0.0f;
      double fd = 0.0;
      int c = 0;
      side_bound<MEAS>(q.x + qo, q.y + qo, n, cb, ma.eps, n >= 2, a1, fd, c);
      side_bound<MEAS>(db.x + co, db.y + co, m, qb, ma.eps, m >= 2, a2, fd, c);
      a1 = wave_sum(a1);
      a2 = wave_sum(a2);
      const float dx0 = qe.x - ce.x, dy0 = qe.y - ce.y, dx1 = qe.z - ce.z, dy1 = qe.w - ce.w;
      const float e0 = sqrt_lo(dx0 * dx0 + dy0 * dy0), e1 = sqrt_lo(dx1 * dx1 + dy1 * dy1);
      // Row form: rows 0 and n-1 contain (0,0) and (n-1,m-1); with n == 1 the
      // single row holds both when m >= 2. Column form likewise.
      const int L = max(n, m);
      const float disc = dtw_disc(L);
      auto bound = [&](float s1, float s2) {
        const float r1 = n >= 2 ? s1 + e0 + e1 : (m >= 2 ? e0 + e1 : e0);
        const float r2 = m >= 2 ? s2 + e0 + e1 : (n >= 2 ? e0 + e1 : e0);
        return fmaxf(r1, r2) * disc;
      };
      bval = bound(a1, a2);
      pass = bval <= t;
      if (pass && ma.band != 0 && n >= 3 && m >= 3) {
        // The whole-MBR form passed: the banded form (row windows of the
        // band) is tighter and costs O(window / 64) per point. Each side alone
        // bounds the value; the row side runs first and may end early.
        const int64_t w = ma.band > 0 ? ma.band : ((int64_t)-ma.band * L + 99) / 100;
        const float b1 = band_side_sum(q.x + qo, q.y + qo, n, db.blk + (co >> 3), db.sblk + (co >> 3), m, w * L,
                                       e0 + e1, disc, t);
        pass = (b1 + e0 + e1) * disc <= t;
        if (pass) {
          const float b2 = band_side_sum(db.x + co, db.y + co, m, q.blk + (qo >> 3), q.sblk + (qo >> 3), n, w * L,
                                         e0 + e1, disc, t);
          bval = fmaxf(bval, bound(b1, b2));
          pass = bval <= t;
        }
      }
    } else {
      float a1 = 0.0f, a2 = 0.0f;
      double fd = 0.0;
      int c = 0;
      side_bound<MEAS>(q.x + qo, q.y + qo, n, cb, ma.eps, false, a1, fd, c);
      side_bound<MEAS>(db.x + co, db.y + co, m, qb, ma.eps, false, a2, fd, c);
      float b2 = fmaxf(wave_max(a1), wave_max(a2));
      if (MEAS == TS_FRECHET) {
        const float dx0 = qe.x - ce.x, dy0 = qe.y - ce.y, dx1 = qe.z - ce.z, dy1 = qe.w - ce.w;
        b2 = fmaxf(b2, fmaxf(dx0 * dx0 + dy0 * dy0, dx1 * dx1 + dy1 * dy1));
      }
      pass = sqrtf(b2) <= t;
      if (pass) {
        // nearest tile box instead of the whole MBR (sqrtf is monotone, so
        // comparing squares against the threshold's square loses nothing
        // only where t*t rounds; the test stays on sqrtf)
        const float lim = t * t * 1.0000005f;
        float c1 = tile_side_max(q.x + qo, q.y + qo, n, db.sblk + (co >> 3), m, lim);
        if (sqrtf(fmaxf(b2, c1)) <= t) {
          const float c2 = tile_side_max(db.x + co, db.y + co, m, q.sblk + (qo >> 3), n, lim);
          c1 = fmaxf(c1, c2);
        }
        bval = sqrtf(fmaxf(b2, c1));
        pass = bval <= t;
      }
    }
    // Seed pairs (already refined exactly, their values merged into the
    // top-k later) leave the candidates.
    if (pass && seed_idx) {
      bool hit = false;
      for (int u = lane; u < s; u += 64) hit |= seed_idx[(int64_t)(it.q - q0) * s + u] == it.d;
      pass = !__any(hit);
    }
    // tb[slot]: a survivor's bound, NaN for the others
    if (tb && lane == 0) tb[it.slot] = pass ? bval : __int_as_float(0x7fc00000);
    // a rejected pair's value slot holds NaN (no refine writes it)
    if (!pass && lane == 0) rej[it.slot] = __int_as_float(0x7fc00000);
    if (pass && lane == 0) {
      const int p = atomicAdd(&scnt, 1);
      if (p < kTightBuf) sbuf[p] = it;
      else out[atomicAdd(nout, 1)] = it;
    }
  }
  __syncthreads();
  const int n = min(scnt, kTightBuf);
  if (threadIdx.x == 0) sbase = n > 0 ? atomicAdd(nout, n) : 0;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) out[sbase + i] = sbuf[i];
}

// Lane-per-pair pre-check ahead of the tight filter (unbanded DTW, Frechet),
// which reads no point: each lane takes one candidate and tests both
// trajectories' 64-point superblock boxes (every 8th entry of a set's sblk)
// against the other trajectory's MBR. Every point lies in its superblock's
// box, so the box's gap to the MBR bounds the point's distance to the MBR
// from below (f32 monotone), and that in turn bounds its distance to the
// nearest tile box the tight filter uses:
//   DTW:     the sum over a side's superblocks of (its interior points) x
//            sqrt_lo(gap^2), plus the endpoint costs e, discounted by
//            dtw_disc (a serial sum of at most kPreBoxes terms, inside the
//            discount's (L/64 + 9) 2^-24 allowance for the bound's sums);
//   Frechet: sqrtf of the largest squared gap of either side's superblocks
//            and of the endpoint costs (max does not round).
// A pair either form puts above its query's threshold is rejected as the
// tight filter would (NaN into rej[slot], and into tb[slot] when given); the
// others, dead items, short pairs (DTW: n or m < 3, whose interior the forms
// leave out) and pairs of more than kPreBoxes superblocks go on to `out`, in
// order within each 256-item round of a block (one atomic per round).
template <int MEAS>
__global__ void __launch_bounds__(256) k_tight_pre(SetView q, int32_t q0, SetView db, const float* tau,
                                                   const Item* items, const int32_t* nitems, int64_t max_items,
                                                   Item* out, int32_t* nout, float* rej, float* tb) {
  __shared__ int wcnt[4];
  __shared__ int sbase;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t n_items = nitems ? min<int64_t>(max_items, (int64_t)*nitems) : max_items;
  for (int64_t r0 = (int64_t)blockIdx.x * 256; r0 < n_items; r0 += (int64_t)gridDim.x * 256) {
    const int64_t i = r0 + threadIdx.x;
    bool keep = false;
    Item it{0, -1, 0, 0};
    if (i < n_items) {
      it = items[i];
      keep = true;
      if (it.d >= 0) {
        const int n = q.len[it.q], m = db.len[it.d];
        const int nsa = (n + 63) >> 6, nsc = (m + 63) >> 6;
        if (nsa + nsc <= kPreBoxes && (MEAS == TS_FRECHET || (n >= 3 && m >= 3))) {
          const int64_t qo = q.off[it.q], co = db.off[it.d];
          const float4 qb = q.mbr[it.q], cb = db.mbr[it.d];
          const float4 qe = q.ends[it.q], ce = db.ends[it.d];
          const float t = tau[it.q - q0];
          const float4* qs = q.sblk + (qo >> 3);
          const float4* cs = db.sblk + (co >> 3);
          const float dx0 = qe.x - ce.x, dy0 = qe.y - ce.y, dx1 = qe.z - ce.z, dy1 = qe.w - ce.w;
          if (MEAS == TS_DTW) {
            float ta = 0.0f, tc = 0.0f;
            for (int u = 0; u < nsa; ++u) {
              const int lo = max(64 * u, 1), hi = min(64 * u + 64, n - 1);
              ta += (float)(hi - lo) * sqrt_lo(box_gap2(qs[8 * u], cb));
            }
            for (int u = 0; u < nsc; ++u) {
              const int lo = max(64 * u, 1), hi = min(64 * u + 64, m - 1);
              tc += (float)(hi - lo) * sqrt_lo(box_gap2(cs[8 * u], qb));
            }
            const float e = sqrt_lo(dx0 * dx0 + dy0 * dy0) + sqrt_lo(dx1 * dx1 + dy1 * dy1);
            keep = (fmaxf(ta, tc) + e) * dtw_disc(max(n, m)) <= t;
          } else {
            float g2 = fmaxf(dx0 * dx0 + dy0 * dy0, dx1 * dx1 + dy1 * dy1);
            for (int u = 0; u < nsa; ++u) g2 = fmaxf(g2, box_gap2(qs[8 * u], cb));
            for (int u = 0; u < nsc; ++u) g2 = fmaxf(g2, box_gap2(cs[8 * u], qb));
            keep = sqrtf(g2) <= t;
          }
          if (!keep) {
            rej[it.slot] = __int_as_float(0x7fc00000);
            if (tb) tb[it.slot] = __int_as_float(0x7fc00000);
          }
        }
      }
    }
    const uint64_t bal = __ballot(keep);
    if (lane == 0) wcnt[wv] = __popcll(bal);
    __syncthreads();
    if (threadIdx.x == 0) {
      const int tot = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
      sbase = tot > 0 ? atomicAdd(nout, tot) : 0;
    }
    __syncthreads();
    if (keep) {
      int pos = sbase + __popcll(bal & ((1ull << lane) - 1ull));
      for (int w = 0; w < wv; ++w) pos += wcnt[w];
      out[pos] = it;
    }
    __syncthreads();  // wcnt and sbase are rewritten next round
  }
}

hipError_t launch_tight_pre(hipStream_t st, MeasureArgs m, SetView q, int32_t q0, SetView db, const float* tau,
                            const Item* items, const int32_t* nitems, int64_t max_items, Item* out, int32_t* nout,
                            float* rej, float* tb) {
  const int64_t rounds = (max_items + 255) / 256;
  const dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>(rounds, 16ll * device_cu_count())));
  if (m.measure == TS_DTW && m.band == 0)
    hipLaunchKernelGGL(k_tight_pre<TS_DTW>, grid, dim3(256), 0, st, q, q0, db, tau, items, nitems, max_items, out,
                       nout, rej, tb);
  else if (m.measure == TS_FRECHET)
    hipLaunchKernelGGL(k_tight_pre<TS_FRECHET>, grid, dim3(256), 0, st, q, q0, db, tau, items, nitems, max_items,
                       out, nout, rej, tb);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// At most grid_waves waves, and no more blocks than fit on the device at once
// (each block's share of the pairs is fixed at launch).
template <int MEAS, bool TILED>
void launch_tight(hipStream_t st, MeasureArgs m, int32_t k, SetView q, int32_t q0, SetView db, const float* tau,
                  const Item* items, const int32_t* nitems, int64_t max_items, Item* out, int32_t* nout, float* rej,
                  int grid_waves, float* tb, const int32_t* seed_idx, int32_t s, bool pre) {
  static int per_cu = 0;
  if (!per_cu && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_tight_filter<MEAS, TILED>, 256, 0) !=
                      hipSuccess ||
                  per_cu < 1))
    per_cu = 1;
  const dim3 grid(std::max(1, std::min((grid_waves + 3) / 4, per_cu * device_cu_count())));
  hipLaunchKernelGGL((k_tight_filter<MEAS, TILED>), grid, dim3(256), 0, st, m, k, q, q0, db, tau, items, nitems,
                     max_items, out, nout, rej, tb, seed_idx, s, pre);
}

hipError_t launch_tight_filter(hipStream_t st, MeasureArgs m, int32_t k, SetView q, int32_t q0, SetView db,
                               const float* tau, const Item* items, const int32_t* nitems, int64_t max_items,
                               Item* out, int32_t* nout, float* rej, int grid_waves, float* tb,
                               const int32_t* seed_idx, int32_t s, bool pre) {
  if ((tb || seed_idx) && m.measure != TS_DTW && m.measure != TS_FRECHET) return hipErrorInvalidValue;
  switch (m.measure) {
#define TS_TIGHT(MEAS, TILED)                                                                              \
  launch_tight<MEAS, TILED>(st, m, k, q, q0, db, tau, items, nitems, max_items, out, nout, rej, grid_waves, tb, \
                            seed_idx, s, pre)
    case TS_DTW:
      if (m.band == 0) TS_TIGHT(TS_DTW, true);
      else TS_TIGHT(TS_DTW, false);
      break;
    case TS_FRECHET: TS_TIGHT(TS_FRECHET, false); break;
    case TS_LCSS: TS_TIGHT(TS_LCSS, false); break;
    case TS_HAUSDORFF: TS_TIGHT(TS_HAUSDORFF, false); break;
    case TS_EUCLIDEAN: TS_TIGHT(TS_EUCLIDEAN, false); break;
#undef TS_TIGHT
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace ts
