// dp_mw.hip -- wavefront DP in which one (query, DB) pair spans several
// wavefronts: the NW waves of a block share a pair, with the rows between
// strips handed from wave to wave through an LDS halo.
//
// Same recurrences and arithmetic as dp.hip (DTW dtw.cpp:64-100, Frechet
// frechet.cpp:83-112 on squared distances, LCSS lcss.cpp:87-103; DTW bands
// and LCSS windows), so values are bit-identical to the one-wave kernels.
//
// Schedule (no counterpart in the reference, whose block-per-pair kernels
// put two block barriers on every anti-diagonal, dtw.cu:66-93):
//  * Rows run along the longer trajectory in strips of 64 (one row per lane);
//    strip s is run by wave s mod NW, lane l owning row 64s+l. Each strip
//    covers the chunk range [cs, ce) of U columns that holds its band (all
//    chunks without a band); lane l sweeps it one chunk per step, one step
//    behind lane l-1, whose outputs arrive by DPP (wave_shr:1).
//  * Strip s starts at step T_s = max(T_{s-1} + kMwLag + (cs_s - cs_{s-1}),
//    T_{s-NW} + (ce_{s-NW} - cs_{s-NW})): late enough that lane 63 of strip
//    s-1 wrote every halo chunk lane 0 of strip s reads at least kMwLag - 63
//    steps earlier, and after its wave finished strip s-NW. The block writes
//    the table of (T_s, cs, ce) in LDS when it fetches the pair.
//  * The block meets at a barrier every kMwB steps; every step count is the
//    same on all waves (the pair's total), so barriers always match, and no
//    wave ever waits for a flag. A halo chunk written in step t is read in
//    step t + kMwB - 1 or later (with the lag below), i.e. after a barrier
//    that follows the write.
//  * The halo is two buffers in LDS: strip s writes buffer s&1 at the chunk's
//    offset from cs_s; strip s+1 reads it, and strip s+2 overwrites it only
//    after strip s+1's reads of the same offsets (at least one barrier
//    later). Lane 0 reads the next step's halo chunk at the end of a step,
//    after that step's barrier.
//  * Points are read from global memory (the row point once per strip, the
//    columns' U points every step), leaving LDS to the halo and the table.
// The block handles one pair after another from the device work counter
// (persistent). Many pairs per launch are throughput work for the one-wave
// kernels; this kernel cuts a long pair's latency by up to NW, for refines
// with few pairs.
#include <algorithm>
#include <atomic>

#include "ts_internal.h"

namespace ts {

// Steps between block barriers, and the least distance between strips'
// starts (the halo written at step t is read at step t + kMwLag - 63 - 1).
constexpr int kMwB = 8;
constexpr int kMwLag = 64 + kMwB;

struct MwArgs {
  SetView q, db;
  const Item* items;
  const int32_t* nitems;
  const int32_t* ntail;
  int64_t max_items;
  int32_t* next;
  float* out;
  int32_t hcap;  // floats per halo buffer (a multiple of U, at least the widest strip range)
  int32_t smax;  // strip table entries
  int32_t band, delta;
  float eps;
  // early abandoning (DTW, banded DTW, Frechet): as in dp.hip's DpArgs
  const float* tau;
  int32_t tau_q0;
  unsigned long long* cells;
};

struct MwStrip {
  int32_t t0, cs, ce, pad;
};

// Strip s's chunk range [clo, chi] (strips of 64 rows), as dp.hip's
// strip_range.
template <bool BAND, bool DELTA, int U>
__device__ __forceinline__ void mw_range(int s, int R, int C, int wb, int delta, int& clo, int& chi) {
  if (!BAND && !DELTA) {
    clo = 0;
    chi = (C - 1) / U;
    return;
  }
  const int r0 = s * 64, r1 = min(s * 64 + 63, R - 1);
  int lo, hi;
  if (BAND) {
    const uint32_t a0 = (uint32_t)r0 * (uint32_t)C, a1 = (uint32_t)r1 * (uint32_t)C;
    lo = max(0, (int)((a0 + (uint32_t)R - 1u) / (uint32_t)R) - wb);
    hi = min(C - 1, (int)(a1 / (uint32_t)R) + wb);
  } else {
    lo = max(0, r0 - delta);
    hi = min(C - 1, r1 + delta);
  }
  clo = lo / U;
  chi = hi / U;
}

// The U column points of chunk kc: one or two aligned 16-byte loads per
// coordinate (sets pad trajectories with their last point to 8).
template <int U>
__device__ __forceinline__ void load_chunk(const float* Bx, const float* By, int kc, float (&cx)[U], float (&cy)[U]) {
#pragma unroll
  for (int j = 0; j < U; j += 4) {
    const float4 x4 = *(const float4*)(Bx + kc * U + j), y4 = *(const float4*)(By + kc * U + j);
    cx[j] = x4.x; cx[j + 1] = x4.y; cx[j + 2] = x4.z; cx[j + 3] = x4.w;
    cy[j] = y4.x; cy[j + 1] = y4.y; cy[j + 2] = y4.z; cy[j + 3] = y4.w;
  }
}

// U consecutive floats from LDS (16-byte aligned).
template <int U>
__device__ __forceinline__ void load_lds(const float* p, float (&v)[U]) {
#pragma unroll
  for (int j = 0; j < U; j += 4) {
    const float4 t = *(const float4*)(p + j);
    v[j] = t.x; v[j + 1] = t.y; v[j + 2] = t.z; v[j + 3] = t.w;
  }
}

// Last chunk holding an in-band cell of row r (r*C < 2^31 for banded pairs).
template <int U>
__device__ __forceinline__ int mw_band_last_chunk(int r, int R, int C, int wb) {
  return min(C - 1, (int)(((uint32_t)r * (uint32_t)C) / (uint32_t)R) + wb) / U;
}

__device__ __forceinline__ int mw_band_half_width(int band, int R, int C) {
  const int w = band > 0 ? band : (-band * R + 99) / 100;
  return min(w, C);
}

// LDS: [0, 64) bytes header (item index, step total, the pair's drop flag;
// bytes 32..63 a border chunk of 8 values outside the table), the strip
// table, then the two halo buffers.
__host__ __device__ constexpr size_t mw_lds_bytes(int smax, int hcap) {
  return 64 + (size_t)smax * sizeof(MwStrip) + 2 * (size_t)hcap * sizeof(float);
}

template <int MEAS, int U, int NW, bool BAND, bool DELTA>
__global__ void __launch_bounds__(256) k_dp_mw(MwArgs a) {
  constexpr float BND = (MEAS == TS_LCSS) ? 0.0f : kInf;
  constexpr bool kAbandon = MEAS != TS_LCSS;
  const float kNoTau = __int_as_float(0x7f800000);
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  extern __shared__ float4 lds_mw4[];
  int* hdr = (int*)lds_mw4;                       // [0] item, [1] step total, [2] dropped
  float* brow = (float*)lds_mw4 + 8;              // [8] the border value
  MwStrip* tbl = (MwStrip*)((char*)lds_mw4 + 64);
  float* halo = (float*)(tbl + a.smax);           // [2][hcap]
  const int hcap = a.hcap, delta = a.delta;

  int64_t n_items = a.max_items;
  if (a.nitems) n_items = min<int64_t>(n_items, (int64_t)*a.nitems);
  const int64_t n_head = n_items;
  if (a.ntail) n_items = min<int64_t>(a.max_items, n_head + (int64_t)*a.ntail);

  if (threadIdx.x < 8) brow[threadIdx.x] = BND;
  for (;;) {
    __syncthreads();  // the previous pair's table and halo are no longer read
    if (threadIdx.x == 0) hdr[0] = atomicAdd(a.next, 1);
    __syncthreads();
    const int it_idx = hdr[0];
    if (it_idx >= n_items) break;  // the same on every thread of the block
    const Item it = a.items[it_idx < n_head ? (int64_t)it_idx : a.max_items - 1 - (it_idx - n_head)];
    if (it.d < 0) continue;
    const int n = a.q.len[it.q], m = a.db.len[it.d];
    const int64_t qo = a.q.off[it.q], dof = a.db.off[it.d];
    const float *Ax, *Ay, *Bx, *By;
    int R, C;
    if (n >= m) {
      Ax = a.q.x + qo; Ay = a.q.y + qo; Bx = a.db.x + dof; By = a.db.y + dof; R = n; C = m;
    } else {
      Ax = a.db.x + dof; Ay = a.db.y + dof; Bx = a.q.x + qo; By = a.q.y + qo; R = m; C = n;
    }
    const bool tiny = MEAS == TS_DTW && (a.q.tiny[it.q] | a.db.tiny[it.d]) != 0;
    const int lastch = (C - 1) / U;
    const int wb = BAND ? mw_band_half_width(a.band, R, C) : 0;
    const int bw = wb * R;
    const int res_row = DELTA ? min(R - 1, C - 1 + delta) : R - 1;
    const int S = res_row / 64 + 1;
    // strip ranges, then their start steps (a short sequential scan)
    for (int s = threadIdx.x; s < S; s += blockDim.x) {
      int clo, chi;
      mw_range<BAND, DELTA, U>(s, R, C, wb, delta, clo, chi);
      tbl[s].cs = clo;
      tbl[s].ce = chi + 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int total = 0;
      for (int s = 0; s < S; ++s) {
        int t0 = 0;
        if (s > 0) t0 = tbl[s - 1].t0 + kMwLag + (tbl[s].cs - tbl[s - 1].cs);
        if (s >= NW) t0 = max(t0, tbl[s - NW].t0 + (tbl[s - NW].ce - tbl[s - NW].cs));
        tbl[s].t0 = t0;
        total = max(total, t0 + 63 + (tbl[s].ce - tbl[s].cs));
      }
      hdr[1] = total;
      hdr[2] = 0;
    }
    __syncthreads();
    const int total = hdr[1];
    const float tau_p = (kAbandon && a.tau) ? a.tau[it.q - a.tau_q0] : kNoTau;

    // ---- per-lane state: the strip this lane is in (or will be in next),
    // its chunk at the current step, its row point and the chunk's points
    int s = wv;
    MwStrip st = s < S ? tbl[s] : MwStrip{INT32_MAX / 2, 0, 0, 0};
    // the previous strip's chunk range (where this strip's halo comes from)
    int pcs = 0, pce = 0;
    if (s > 0 && s < S) {
      pcs = tbl[s - 1].cs;
      pce = tbl[s - 1].ce;
    }
    int row = s * 64 + lane;
    float ax = 0.0f, ay = 0.0f;
    if (s < S) {
      ax = Ax[min(row, R - 1)];
      ay = Ay[min(row, R - 1)];
    }
    // early abandoning: the last chunk holding a (band) cell of the lane's
    // row, and the running minimum of the row
    int rlast = BAND ? mw_band_last_chunk<U>(min(row, R - 1), R, C, wb) : lastch;
    float rowmin = kNoTau;
    int k = st.cs - st.t0 - lane;  // step 0 (at most cs: no switch yet)
    float cx[U], cy[U];
    load_chunk<U>(Bx, By, min(max(k, 0), lastch), cx, cy);
    float o[U], hv[U];
#pragma unroll
    for (int j = 0; j < U; ++j) { o[j] = BND; hv[j] = BND; }
    float up_last = BND, hdiag = BND;
    // lane 0 of strip 0 reads the border above row 0 (hv = BND); lane 0 of
    // every later strip gets its first chunk's operands at the end of the
    // step before the strip starts (T_s >= kMwLag > 0).

    for (int t = 0; t < total; ++t) {
      // ---- this step: strip, chunk, row (the next step's replace them below)
      const int sc = s, kk = k, rowc = row, cs_c = st.cs, rlc = rlast;
      const bool active = sc < S && kk >= cs_c && kk < st.ce;
      const int kc = min(max(kk, 0), lastch);  // idle lanes read a valid chunk
      // ---- point terms: squared distances (DTW, Frechet) or matches (LCSS)
      float pt[U];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const float dx = ax - cx[j], dy = ay - cy[j];
        if (MEAS == TS_LCSS) {
          pt[j] = (fabsf(dx) < a.eps && fabsf(dy) < a.eps) ? 1.0f : 0.0f;
        } else {
          pt[j] = dx * dx + dy * dy;
        }
      }
      // ---- the next step's strip, chunk and operands, read now into the
      // registers the point terms are done with, so the reads have the rest
      // of the step to arrive in. The empty asm takes the point terms and
      // orders memory: no compiler pass moves their arithmetic below the
      // reads (which would keep both chunks live and copy one into the other
      // every step), nor the reads above it.
#pragma unroll
      for (int j = 0; j < U; ++j) asm volatile("" : : "v"(pt[j]) : "memory");
      __builtin_amdgcn_sched_barrier(0);
      k = kk + 1;
      if (k >= st.ce) {
        // past this strip: the wave's next one (at most one switch per step)
        s += NW;
        st = s < S ? tbl[s] : MwStrip{INT32_MAX / 2, 0, 0, 0};
        row = s * 64 + lane;
        if (s < S) {
          ax = Ax[min(row, R - 1)];
          ay = Ay[min(row, R - 1)];
          pcs = tbl[s - 1].cs;
          pce = tbl[s - 1].ce;
          if (BAND) rlast = mw_band_last_chunk<U>(min(row, R - 1), R, C, wb);
        }
        k = t + 1 - st.t0 - lane + st.cs;
      }
      load_chunk<U>(Bx, By, min(max(k, 0), lastch), cx, cy);
      __builtin_amdgcn_sched_barrier(0);
      // ---- U cells (as in dp.hip's k_dp_lean_r)
      float up[U], on[U];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        up[j] = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(hv[j]), __float_as_int(o[j]),
                                                           0x138 /* wave_shr:1 */, 0xf, 0xf, false));
      }
      const bool first = kk == cs_c;
      float d = first ? ((rowc == 0) ? (cs_c == 0 ? 0.0f : BND) : (lane == 0 ? hdiag : BND)) : up_last;
      if (MEAS == TS_LCSS && first && cs_c == 0) d = 0.0f;
      float l = first ? BND : o[U - 1];
      if (MEAS == TS_DTW) sqrt_cr<U>(pt, tiny);
      int bd = 0;
      if (BAND) bd = rowc * C - kc * U * R;  // kc == kk on active lanes
#pragma unroll
      for (int j = 0; j < U; ++j) {
        float v;
        if (MEAS == TS_DTW) {
          v = pt[j] + fminf(fminf(up[j], l), d);
          if (BAND) {
            const int bj = bd - j * R;
            if ((bj < 0 ? -bj : bj) > bw) v = kInf;
          }
        } else if (MEAS == TS_FRECHET) {
          v = fmaxf(pt[j], fminf(fminf(up[j], l), d));
        } else {
          bool match = pt[j] != 0.0f;
          float uu = up[j], ll = l;
          if (DELTA) {
            const int dd = rowc - (kc * U + j);
            match = match && (dd < 0 ? -dd : dd) <= delta;
            if (dd - 1 > delta || 1 - dd > delta) uu = d;  // (row-1, col) outside the window
            if (dd + 1 > delta || -1 - dd > delta) ll = d;  // (row, col-1) outside the window
          }
          v = match ? d + 1.0f : fmaxf(uu, ll);
        }
        on[j] = v;
        d = up[j];
        l = v;
      }
      up_last = up[U - 1];
#pragma unroll
      for (int j = 0; j < U; ++j) o[j] = on[j];
      // ---- the row between strips, and the pair's result
      if (lane == 63 && active) {
        float* hb = halo + (sc & 1) * hcap + (kk - cs_c) * U;
#pragma unroll
        for (int j = 0; j < U; ++j) hb[j] = on[j];
      }
      // ---- early abandoning: row minima never fall from row to row (see
      // dp.hip); a strip's last row, not the pair's last, whose minimum at
      // its last (band) chunk exceeds tau drops the pair: the flag is seen
      // at the next barrier by every thread, at most kMwB - 1 steps later,
      // and the result cell of a later strip comes at least kMwLag - 63 =
      // kMwB + 1 steps later, so the result is never written for a dropped
      // pair (the write also tests the flag)
      if (kAbandon) {
        float mn = on[0];
#pragma unroll
        for (int j = 1; j < U; ++j) mn = fminf(mn, on[j]);
        rowmin = fminf(kk == cs_c ? kNoTau : rowmin, mn);
        if (lane == 63 && active && kk == rlc && rowc + 1 < R &&
            (MEAS == TS_FRECHET ? sqrtf(rowmin) : rowmin) > tau_p && atomicExch(&hdr[2], 1) == 0) {
          a.out[it.slot] = __int_as_float(0x7fc00000);
          // (integer arithmetic: the device min template is floating point)
          const uint64_t per_row = (uint64_t)(BAND && 2 * wb + 1 < C ? 2 * wb + 1 : C);
          atomicAdd(a.cells, (unsigned long long)(0ull - (uint64_t)(R - 1 - rowc) * per_row));
        }
      }
      if (active && rowc == res_row && kk == lastch && !(kAbandon && *(volatile int*)&hdr[2])) {
        const int rc = (C - 1) - lastch * U;
        float v = on[0];
#pragma unroll
        for (int j = 1; j < U; ++j)
          if (rc == j) v = on[j];
        if (MEAS == TS_FRECHET) v = sqrtf(v);
        if (MEAS == TS_LCSS) v = v / (float)C;
        a.out[it.slot] = v;
      }
      if ((t + 1) % kMwB == 0) {
        __syncthreads();
        if (kAbandon && *(volatile int*)&hdr[2]) break;  // dropped: the same on every thread
      }
      // ---- the halo chunk of step t+1 (strip s, chunk k): the previous
      // strip's row there, else the border (above row 0, or above the
      // previous strip's range: outside the band). Every lane reads it, so
      // it dies at the shift above; only lane 0's is used.
      const bool from_halo = s > 0 && s < S && k >= pcs && k < pce;
      const float* hb = halo + ((s - 1) & 1) * hcap;
      load_lds<U>(from_halo ? hb + (k - pcs) * U : brow, hv);
      // lane 0 at a strip's start: the diagonal of its first cell, column
      // cs*U-1 of the row above, when the previous strip computed it
      if (lane == 0 && s < S && k == st.cs) {
        const int dc = st.cs * U - 1;
        hdiag = (s > 0 && dc >= 0 && dc >= pcs * U && dc < pce * U && dc < C) ? hb[dc - pcs * U] : BND;
      }
    }
  }
}

// LDS a block of this kernel may take: gfx950 gives a workgroup up to
// 160 KiB (two halo rows of 8192 floats and the strip table need 66 KB);
// the least over the devices dp_mw_setup has seen (0: none yet, 64 KiB).
static std::atomic<int32_t> g_mw_lds_max{0};

template <int MEAS, int U, int NW, bool BAND, bool DELTA>
static hipError_t mw_allow_lds(int32_t bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_dp_mw<MEAS, U, NW, BAND, DELTA>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

template <int MEAS, bool BAND, bool DELTA>
static hipError_t mw_allow_lds_m(int32_t bytes) {
  hipError_t e = mw_allow_lds<MEAS, 4, 2, BAND, DELTA>(bytes);
  if (e == hipSuccess) e = mw_allow_lds<MEAS, 4, 4, BAND, DELTA>(bytes);
  if (e == hipSuccess) e = mw_allow_lds<MEAS, 8, 2, BAND, DELTA>(bytes);
  if (e == hipSuccess) e = mw_allow_lds<MEAS, 8, 4, BAND, DELTA>(bytes);
  return e;
}

ts_status dp_mw_setup(int device) {
  int lim = 0;
  TS_HIP(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  const int32_t bytes = std::min<int32_t>(lim, 163840);
  hipError_t e = mw_allow_lds_m<TS_DTW, false, false>(bytes);
  if (e == hipSuccess) e = mw_allow_lds_m<TS_DTW, true, false>(bytes);
  if (e == hipSuccess) e = mw_allow_lds_m<TS_FRECHET, false, false>(bytes);
  if (e == hipSuccess) e = mw_allow_lds_m<TS_LCSS, false, false>(bytes);
  if (e == hipSuccess) e = mw_allow_lds_m<TS_LCSS, false, true>(bytes);
  TS_HIP(e);
  int32_t cur = g_mw_lds_max.load();
  while ((cur == 0 || bytes < cur) && !g_mw_lds_max.compare_exchange_weak(cur, bytes)) {
  }
  return TS_OK;
}

int32_t dp_mw_lds_max() {
  const int32_t v = g_mw_lds_max.load();
  return v ? v : 65536;
}

// Lanes per pair of the multi-wave shapes: 128 (2 waves) or 256 (4).
template <int MEAS, int U, bool BAND, bool DELTA>
static hipError_t launch_mw_t(hipStream_t st, const MwArgs& a, int nw, int grid_waves) {
  const size_t lds = mw_lds_bytes(a.smax, a.hcap);
  if (lds > (size_t)dp_mw_lds_max()) return hipErrorInvalidValue;
  int blocks = std::max(1, grid_waves / nw);
  if (nw == 2)
    hipLaunchKernelGGL((k_dp_mw<MEAS, U, 2, BAND, DELTA>), dim3(blocks), dim3(128), lds, st, a);
  else
    hipLaunchKernelGGL((k_dp_mw<MEAS, U, 4, BAND, DELTA>), dim3(blocks), dim3(256), lds, st, a);
  return hipGetLastError();
}

template <int MEAS, bool BAND, bool DELTA>
static hipError_t launch_mw_m(hipStream_t st, const MwArgs& a, int nw, int u, int grid_waves) {
  return u == 4 ? launch_mw_t<MEAS, 4, BAND, DELTA>(st, a, nw, grid_waves)
                : launch_mw_t<MEAS, 8, BAND, DELTA>(st, a, nw, grid_waves);
}

// Halo floats per buffer for pairs whose shorter length is at most cmax_c
// and whose visited width is at most wmax_c: the widest strip range, the
// whole row without a band; a band's or window's range spans at most
// 63 + width + 1 columns plus a chunk of rounding at each end, and never
// more than the row. One chunk of margin.
int32_t dp_mw_hcap(int u, const MeasureArgs& m, int32_t cmax_c, int32_t wmax_c) {
  int64_t cols = cmax_c;
  if (dp_ranged(m)) cols = std::min<int64_t>(cmax_c, (int64_t)wmax_c + 64 + 2 * u);
  return (int32_t)(((cols + u - 1) / u + 1) * u);
}

size_t dp_mw_lds_bytes(int32_t rmax_c, int32_t hcap) { return mw_lds_bytes((rmax_c + 63) / 64 + 1, hcap); }

hipError_t launch_dp_mw(hipStream_t st, MeasureArgs m, int nw, int u, SetView q, SetView db, const Item* items,
                        const int32_t* nitems, const int32_t* ntail, int64_t max_items, int32_t* next, float* out,
                        int32_t hcap, int32_t rmax_c, int grid_waves, const float* tau, int32_t tau_q0,
                        unsigned long long* cells) {
  if ((nw != 2 && nw != 4) || (u != 4 && u != 8)) return hipErrorInvalidValue;
  MwArgs a{q,       db,      items, nitems, ntail, max_items, next,   out,  hcap, (rmax_c + 63) / 64 + 1,
           m.band,  m.delta, m.eps, tau,    tau_q0, cells};
  switch (m.measure) {
    case TS_DTW:
      return m.band != 0 ? launch_mw_m<TS_DTW, true, false>(st, a, nw, u, grid_waves)
                         : launch_mw_m<TS_DTW, false, false>(st, a, nw, u, grid_waves);
    case TS_FRECHET: return launch_mw_m<TS_FRECHET, false, false>(st, a, nw, u, grid_waves);
    case TS_LCSS:
      return m.delta >= 0 ? launch_mw_m<TS_LCSS, false, true>(st, a, nw, u, grid_waves)
                          : launch_mw_m<TS_LCSS, false, false>(st, a, nw, u, grid_waves);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ts
