// dp.hip -- wavefront DP refine for DTW, discrete Frechet and LCSS.
//
// Recurrences (reference CPU paths):
//   DTW      D[i][j] = d(a_i,b_j) + min(D[i-1][j], D[i][j-1], D[i-1][j-1]),
//            D[-1][-1] = 0, other borders 1e20f           (dtw.cpp:64-100)
//   Frechet  F[i][j] = max(d, min(F[i-1][j], F[i][j-1], F[i-1][j-1]))
//                                                          (frechet.cpp:83-112)
//   LCSS     L[i][j] = L[i-1][j-1] + 1 if |dx| < eps and |dy| < eps,
//            else max(L[i-1][j], L[i][j-1]); borders 0     (lcss.cpp:87-103)
// Frechet runs on squared distances with one sqrtf at the end: sqrtf is
// correctly rounded and monotone, so the result is bit-identical. LCSS counts
// are held as exact small integers in f32 registers.
//
// Execution model (no counterpart in the reference, whose kernels use one
// block per pair and two block barriers per anti-diagonal, dtw.cu:66-93):
//  * A wave64 is split into 64/W lane groups (W = 16, 32 or 64).
//    Each group runs one pair at a time; lane l of the group owns DP rows
//    l, l+W, ... along the longer trajectory and sweeps its row U cells per
//    step, one step behind lane l-1, so the anti-diagonal lives in registers.
//  * The U cells above arrive from lane l-1 by DPP shifts (row_shr:1 for
//    W = 16, wave_shr:1 otherwise) of the U values it produced on the
//    previous step; the first cell's diagonal is the last value that arrived
//    on the step before. Lane 0 of a group reads the row above from an LDS
//    halo row written by lane W-1 during the previous strip of W rows.
//  * Columns run along the shorter trajectory, U per step; a row has at
//    least W (k_dp) or W+1 (k_dp_lean) chunks so the halo is written before
//    it is read; shorter rows are padded with idle chunks.
//  * Three kernels. k_dp_lean runs unbanded pairs with a per-step path cut to
//    the cells and straight-line bookkeeping, staging the pair's columns and
//    rows in LDS when they fit (see above k_dp_lean). k_dp_lean_r runs DTW
//    bands and LCSS windows with the same step: each strip visits only its
//    band's chunk range, taken from a strip table in LDS, and the halo is an
//    LDS ring. k_dp runs the bands and windows whose ring and table do not
//    fit a wave's LDS share, with columns from global memory.
//  * Strips and pairs follow each other without draining the pipeline: after
//    its last row a lane starts its first row of the next pair, which the
//    group's lane 0 fetched from a device-wide atomic work counter. The
//    kernel is persistent: it drains the whole work list in one launch.
//  * Work lists are split into buckets by the shorter length C. Each bucket
//    uses the widest group that needs no idle chunks (C/U >= W), which
//    minimises a pair's latency (R*C/(W*U) steps) when pairs are few, and
//    its own halo length, so short pairs do not pay for long rows' LDS.
#include <algorithm>
#include <climits>
#include <type_traits>

#include "ts_internal.h"

namespace ts {

struct DpArgs {
  SetView q, db;
  const Item* items;
  const int32_t* nitems;  // items [0, *nitems) ...
  const int32_t* ntail;   // ... then, when set, *ntail items stored backwards from items[max_items - 1]
  int64_t max_items;
  int32_t* next;
  float* out;
  int32_t cmax;   // LDS halo floats per group (and staged column capacity)
  int32_t hmask;  // halo index mask: -1, or ring length - 1
  int32_t rmax;   // k_dp_lean: staged row buffer floats per coordinate
  int32_t ccap;   // k_dp_lean_r: staged column capacity in points (the halo is a ring)
  int32_t smax;   // k_dp_lean_r: strip table entries
  int32_t band;
  int32_t delta;
  float eps;
  // Early abandoning (k_dp_lean, DTW and Frechet): a pair stops once a
  // finished row's minimum exceeds its query's threshold tau[q - tau_q0]
  // (null: never); its value slot then holds NaN and `cells` (the launch's
  // cell counter) loses the cells of the rows it did not finish.
  const float* tau;
  int32_t tau_q0;
  unsigned long long* cells;
  // k_dp_lean, DTW: entries of each lane group's suffix-bound table (a
  // multiple of 4; 0: no table), see lean_suffix_table.
  int32_t sfx;
};

// v of the previous lane of the group; the group's lane 0 gets first0.
// W = 16: row_shr:1, whose 16-lane rows are the groups; W = 64: wave_shr:1.
// In both, a lane without a source lane keeps the DPP's old operand, so
// first0 lands in lane 0 with no select. W = 32: wave_shr:1, and lanes 0
// and 32 select first0.
template <int W>
__device__ __forceinline__ float from_prev_lane(float v, float first0, int lg) {
  if constexpr (W == 32) {
    const float s = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
    return lg == 0 ? first0 : s;
  } else {
    constexpr int ctrl = (W == 16) ? 0x111 /* row_shr:1 */ : 0x138 /* wave_shr:1 */;
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(first0), __float_as_int(v), ctrl, 0xf, 0xf, false));
  }
}

// from_prev_lane over a step's U values. W = 32 takes one instruction per
// value: v_cndmask_b32 with a DPP (wave_shr:1, out-of-range lanes read 0)
// first operand, VCC holding lanes 0 and 32, selects first0 there and the
// shifted value elsewhere. The s_nop gives the 2 wait states a DPP read
// needs after a VALU write of its source.
#define TS_CND_DPP(D, S) "v_cndmask_b32_dpp %" #D ", %" #S ", %" #D ", vcc wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
#define TS_CND_HEAD "s_mov_b32 vcc_lo, 1\n\ts_mov_b32 vcc_hi, 1\n\ts_nop 1\n\t"
template <int W, int U>
__device__ __forceinline__ void from_prev_lanes(float (&up)[U], const float (&o)[U], const float (&first0)[U], int lg) {
  if constexpr (W == 32 && U == 8) {
#pragma unroll
    for (int j = 0; j < U; ++j) up[j] = first0[j];
    asm(TS_CND_HEAD TS_CND_DPP(0, 8) TS_CND_DPP(1, 9) TS_CND_DPP(2, 10) TS_CND_DPP(3, 11) TS_CND_DPP(4, 12)
            TS_CND_DPP(5, 13) TS_CND_DPP(6, 14) TS_CND_DPP(7, 15)
        : "+v"(up[0]), "+v"(up[1]), "+v"(up[2]), "+v"(up[3]), "+v"(up[4]), "+v"(up[5]), "+v"(up[6]), "+v"(up[7])
        : "v"(o[0]), "v"(o[1]), "v"(o[2]), "v"(o[3]), "v"(o[4]), "v"(o[5]), "v"(o[6]), "v"(o[7])
        : "vcc");
  } else if constexpr (W == 32 && U == 4) {
#pragma unroll
    for (int j = 0; j < U; ++j) up[j] = first0[j];
    asm(TS_CND_HEAD TS_CND_DPP(0, 4) TS_CND_DPP(1, 5) TS_CND_DPP(2, 6) TS_CND_DPP(3, 7)
        : "+v"(up[0]), "+v"(up[1]), "+v"(up[2]), "+v"(up[3])
        : "v"(o[0]), "v"(o[1]), "v"(o[2]), "v"(o[3])
        : "vcc");
  } else if constexpr (W == 32 && U == 2) {
    up[0] = first0[0];
    up[1] = first0[1];
    asm(TS_CND_HEAD TS_CND_DPP(0, 2) TS_CND_DPP(1, 3) : "+v"(up[0]), "+v"(up[1]) : "v"(o[0]), "v"(o[1]) : "vcc");
  } else {
#pragma unroll
    for (int j = 0; j < U; ++j) up[j] = from_prev_lane<W>(o[j], first0[j], lg);
  }
}
#undef TS_CND_DPP
#undef TS_CND_HEAD

template <int U>
__device__ __forceinline__ void load_vec(const float* p, float (&v)[U]) {
  if constexpr (U == 8) {
    const float4 t = *(const float4*)p, u = *(const float4*)(p + 4);
    v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    v[4] = u.x; v[5] = u.y; v[6] = u.z; v[7] = u.w;
  } else if constexpr (U == 4) {
    const float4 t = *(const float4*)p;
    v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
  } else if constexpr (U == 2) {
    const float2 t = *(const float2*)p;
    v[0] = t.x; v[1] = t.y;
  } else {
    v[0] = *p;
  }
}

template <int U>
__device__ __forceinline__ void store_vec(float* p, const float (&v)[U]) {
  if constexpr (U == 8) {
    *(float4*)p = make_float4(v[0], v[1], v[2], v[3]);
    *(float4*)(p + 4) = make_float4(v[4], v[5], v[6], v[7]);
  } else if constexpr (U == 4) {
    *(float4*)p = make_float4(v[0], v[1], v[2], v[3]);
  } else if constexpr (U == 2) {
    *(float2*)p = make_float2(v[0], v[1]);
  } else {
    *p = v[0];
  }
}

// Band half-width in rows of a pair (R rows, C columns): band cells, or
// -band percent of R when negative; a width of C or more admits every cell,
// so it is clamped there (w*R then stays below 2^31).
__device__ __forceinline__ int band_half_width(int band, int R, int C) {
  const int w = band > 0 ? band : (-band * R + 99) / 100;
  return min(w, C);
}

// LDS per lane group of k_dp, in floats: the halo row (or ring).
__host__ __device__ constexpr int dp_group_floats(int cmax) { return cmax; }

// Chunk range [clo, chi] (chunks of U columns) that holds every in-band cell
// of strip s (rows s*W .. s*W+W-1). DTW band: |row*C - col*R| <= bw, so row
// r spans columns ceil((rC-bw)/R) .. floor((rC+bw)/R); LCSS window: r-delta
// .. r+delta; otherwise the whole row.
// With bw = wb*R the band's limits are ceil(r0*C/R) - wb and
// floor(r1*C/R) + wb (wb an integer), and r*C < R*C < 2^31 (the host refuses
// longer banded pairs), so 32-bit divisions give them.
template <bool BAND, bool DELTA, int W, int U>
__device__ __forceinline__ void strip_range(int s, int R, int C, int wb, int delta, int lastch, int& clo, int& chi) {
  if (!BAND && !DELTA) {
    clo = 0;
    chi = lastch;
    return;
  }
  const int r0 = s * W, r1 = min(s * W + W - 1, R - 1);
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

template <int MEAS, int W, int U, bool BAND, bool DELTA>
__global__ void __launch_bounds__(256) k_dp(DpArgs a) {
  static_assert(BAND || DELTA, "unbanded pairs run k_dp_lean");
  constexpr int G = 64 / W;
  constexpr float BND = (MEAS == TS_LCSS) ? 0.0f : kInf;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const int g = lane / W, lg = lane & (W - 1);
  const int cmax = a.cmax;
  // Halo slot of a column: with a band the halo is a ring that holds one
  // strip's range (its length is at least the longest range plus a chunk,
  // so a slot is rewritten only after lane 0 of the next strip read it).
  const int hmask = a.hmask;
  extern __shared__ float4 lds_dyn4[];
  // LDS: the groups' fetched item slots, one border row of 8 values per wave
  // (what lane 0 reads above row 0 and outside a band), then the group
  // areas, all on 16-byte boundaries (cmax is a multiple of 4).
  const int hdr = ((nwaves * G + 3) >> 2) << 2;
  int* slot_item = (int*)lds_dyn4 + wv * G + g;
  float* brow = (float*)lds_dyn4 + hdr + wv * 8;
  float* gbase = (float*)lds_dyn4 + hdr + nwaves * 8 + (size_t)(wv * G + g) * dp_group_floats(cmax);
  if (lane < 8) brow[lane] = BND;
  float* bnd = gbase;

  int64_t n_items = a.max_items;
  if (a.nitems) n_items = min<int64_t>(n_items, (int64_t)*a.nitems);
  const int64_t n_head = n_items;
  if (a.ntail) n_items = min<int64_t>(a.max_items, n_head + (int64_t)*a.ntail);

  // rem: steps left in this lane's current pair (its first pair starts after
  // lg idle steps). Lane 0 of the group starts each pair exactly lg steps
  // before this lane does, i.e. on the step where rem == lg.
  int row = lg, ch = -lg, rem = lg;
  bool done = false, live = false;
  int R = 1, C = 1, S = 1, slot = 0, lastch = 0;
  // Current strip: chunks [cstart, cend); previous strip's range (lane 0
  // reads its halo only inside it).
  int strip = 0, cstart = 0, cend = W, pclo = 0, pcend = 0;
  int res_row = 0, res_col = 0;
  const float *Ax = a.q.x, *Ay = a.q.y, *Bx = a.q.x, *By = a.q.y;
  float ax = 0.0f, ay = 0.0f, nax = 0.0f, nay = 0.0f;
  // descriptor of the group's next pair (see the fetch below)
  const float *fAx = a.q.x, *fAy = a.q.y, *fBx = a.q.x, *fBy = a.q.y;
  int fR = 1, fC = 1, f_slot = 0;
  bool f_ok = false, f_live = false;
  bool f_tiny = false, pair_tiny = false;  // fetched / current pair has a tiny coordinate
  float f_ax = 0.0f, f_ay = 0.0f;
  int bd = 0, bw = 0, wb = 0;
  int f_rem = 0;  // steps of the fetched pair (ranged: summed over its strips by the group)
  const int delta = a.delta;
  float o[U], cx[U], cy[U], hv[U];
#pragma unroll
  for (int j = 0; j < U; ++j) { o[j] = 0.0f; hv[j] = 0.0f; cx[j] = 0.0f; cy[j] = 0.0f; }
  float up_last = 0.0f, hdiag = 0.0f;

  for (;;) {
    // LDS is shared by the lanes of this wave only; the barrier keeps the
    // compiler from moving LDS accesses across steps.
    __builtin_amdgcn_wave_barrier();
    // ---- the group's lane 0 starts a pair: fetch it.
    // The whole group reads the pair's descriptor here, once, and keeps it
    // in registers (f*) until each lane starts the pair: lanes start on
    // consecutive steps, so per-lane reads would stall W steps in a row on
    // dependent global loads. The next fetch comes after lane 0 finishes
    // this pair, i.e. after every lane has started it.
    if (!done && rem == lg) {
      if (lg == 0) *slot_item = atomicAdd(a.next, 1);
      __builtin_amdgcn_wave_barrier();
      const int it_idx = *slot_item;
      f_ok = it_idx < n_items;
      f_live = false;
      if (f_ok) {
        const Item it = a.items[it_idx < n_head ? (int64_t)it_idx : a.max_items - 1 - (it_idx - n_head)];
        f_live = it.d >= 0;
        if (f_live) {
          const int n = a.q.len[it.q], m = a.db.len[it.d];
          const int64_t qo = a.q.off[it.q], dof = a.db.off[it.d];
          if (n >= m) {
            fAx = a.q.x + qo; fAy = a.q.y + qo; fBx = a.db.x + dof; fBy = a.db.y + dof; fR = n; fC = m;
          } else {
            fAx = a.db.x + dof; fAy = a.db.y + dof; fBx = a.q.x + qo; fBy = a.q.y + qo; fR = m; fC = n;
          }
          f_slot = it.slot;
          f_tiny = MEAS == TS_DTW && (a.q.tiny[it.q] | a.db.tiny[it.d]) != 0;
          // this lane's first row point
          f_ax = fAx[min(lg, fR - 1)];
          f_ay = fAy[min(lg, fR - 1)];
        }
      }
      {
        // the fetched pair's steps, its strips shared out over the group
        const int R2 = f_live ? fR : 1, C2 = f_live ? fC : 1;
        const int lastch2 = (C2 - 1) / U;
        const int wb2 = BAND ? band_half_width(a.band, R2, C2) : 0;
        const int S2 = DELTA ? min(R2 - 1, C2 - 1 + delta) / W + 1 : (R2 + W - 1) / W;
        int part = 0;
        for (int s2 = lg; s2 < S2; s2 += W) {
          int clo, chi, nlo, nhi;
          strip_range<BAND, DELTA, W, U>(s2, R2, C2, wb2, delta, lastch2, clo, chi);
          nlo = clo;
          if (s2 + 1 < S2) strip_range<BAND, DELTA, W, U>(s2 + 1, R2, C2, wb2, delta, lastch2, nlo, nhi);
          part += max(chi - clo + 1, W + (nlo - clo));
        }
#pragma unroll
        for (int sh = W / 2; sh > 0; sh >>= 1) part += __shfl_xor(part, sh);  // within the group
        f_rem = part;
      }
    }
    // ---- this lane starts the pair its group's lane 0 fetched lg steps ago
    if (!done && rem == 0) {
      if (!f_ok) {
        done = true;
      } else {
        live = f_live;
        pair_tiny = f_live && f_tiny;
        if (live) {
          Ax = fAx; Ay = fAy; Bx = fBx; By = fBy; R = fR; C = fC;
          slot = f_slot;
          nax = f_ax;
          nay = f_ay;
        } else {
          R = 1; C = 1;  // a dead item: computed on stale operands, never stored
          nax = 0.0f;
          nay = 0.0f;
        }
        lastch = (C - 1) / U;
        // band < 0: a half-width of -band percent of the longer length.
        if (BAND) {
          wb = band_half_width(a.band, R, C);
          bw = wb * R;
        }
        // The result cell: the corner, or for an LCSS window whose band
        // misses it, the band's last cell in the last column.
        res_col = C - 1;
        res_row = DELTA ? min(R - 1, C - 1 + delta) : R - 1;
        // Strips of the pair (an LCSS window stops at the result's row), and
        // its steps: every strip lasts at least W steps plus the shift of the
        // next strip's range, so the halo chunk a lane 0 needs was written at
        // least one step earlier.
        S = DELTA ? res_row / W + 1 : (R + W - 1) / W;
        int clo0, chi0;
        strip_range<BAND, DELTA, W, U>(0, R, C, wb, delta, lastch, clo0, chi0);
        rem = f_rem;
        strip = 0;
        int nlo = clo0, nhi = chi0;
        if (S > 1) strip_range<BAND, DELTA, W, U>(1, R, C, wb, delta, lastch, nlo, nhi);
        cstart = clo0;
        cend = clo0 + max(chi0 - clo0 + 1, W + (nlo - clo0));
        pclo = 0;
        pcend = 0;
        row = lg;
        ch = cstart;
        const int lc0 = U * min(cstart, lastch);
        load_vec<U>(Bx + lc0, cx);
        load_vec<U>(By + lc0, cy);
        load_vec<U>(brow, hv);  // lane 0 starts on row 0: the border row above it
      }
    }
    if (__all(done)) break;

    // ---- start of a row (its point was loaded one row ahead)
    if (ch == cstart && !done) {
      ax = nax;
      ay = nay;
      const int nr = min(row + W, R - 1);
      nax = Ax[nr];
      nay = Ay[nr];
      if (BAND) bd = row * C - cstart * U * R;
      // Lane 0's diagonal for the strip's first cell: the halo row's value
      // left of the range when the previous strip computed it.
      const int dc = cstart * U - 1;
      hdiag = (dc >= 0 && dc >= pclo * U && dc < pcend * U && dc < C) ? bnd[dc & hmask] : BND;
    }

    // ---- U cells
    float up[U], on[U];
    const bool first = ch == cstart;
#pragma unroll
    for (int j = 0; j < U; ++j) up[j] = from_prev_lane<W>(o[j], hv[j], lg);
    float d = first ? ((row == 0) ? (cstart == 0 ? 0.0f : BND) : (lg == 0 ? hdiag : BND)) : up_last;
    if (MEAS == TS_LCSS && first && cstart == 0) d = 0.0f;
    float l = first ? BND : o[U - 1];
    float dist[U];
    if (MEAS == TS_DTW) {
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const float dx = ax - cx[j], dy = ay - cy[j];
        dist[j] = dx * dx + dy * dy;
      }
      sqrt_cr<U>(dist, __any(pair_tiny));
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const float dx = ax - cx[j], dy = ay - cy[j];
      float v;
      if (MEAS == TS_DTW) {
        v = dist[j] + fminf(fminf(up[j], l), d);
        if (BAND) {
          const int bj = bd - j * R;
          if ((bj < 0 ? -bj : bj) > bw) v = kInf;
        }
      } else if (MEAS == TS_FRECHET) {
        v = fmaxf(dx * dx + dy * dy, fminf(fminf(up[j], l), d));
      } else {
        bool match = fabsf(dx) < a.eps && fabsf(dy) < a.eps;
        float uu = up[j], ll = l;
        if (DELTA) {
          // In-band cells read an out-of-band neighbour as their diagonal
          // (L is constant along a column below the band and along a row
          // right of it); the diagonal of an in-band cell is in band.
          const int dd = row - (ch * U + j);
          match = match && (dd < 0 ? -dd : dd) <= delta;
          if (dd - 1 > delta || 1 - dd > delta) uu = d;   // (row-1, col) outside
          if (dd + 1 > delta || -1 - dd > delta) ll = d;   // (row, col-1) outside
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

    // ---- halo row for the next strip, and the pair's result
    if (lg == W - 1 && ch >= 0 && ch <= lastch && !done) store_vec<U>(bnd + ((ch * U) & hmask), on);
    if (!done && live && row == res_row && ch == res_col / U) {
      const int rc = res_col - ch * U;
      float v = on[0];
#pragma unroll
      for (int j = 1; j < U; ++j)
        if (rc == j) v = on[j];
      if (MEAS == TS_FRECHET) v = sqrtf(v);
      if (MEAS == TS_LCSS) v = v / (float)C;
      a.out[slot] = v;
    }

    // ---- advance
    --rem;
    ++ch;
    if (BAND) bd -= U * R;
    if (ch == cend) {
      // next strip of this lane
      ++strip;
      row += W;
      pclo = cstart;
      pcend = cend;
      int nlo = 0, nhi = 0, nnlo = 0, nnhi = 0;
      strip_range<BAND, DELTA, W, U>(strip, R, C, wb, delta, lastch, nlo, nhi);
      nnlo = nlo;
      if (strip + 1 < S) strip_range<BAND, DELTA, W, U>(strip + 1, R, C, wb, delta, lastch, nnlo, nnhi);
      cstart = nlo;
      cend = nlo + max(nhi - nlo + 1, W + (nnlo - nlo));
      ch = cstart;
    }
    // ---- operands of the next step: columns of chunk ch (global memory),
    // and the halo chunk ch, written by lane W-1 at least one step ago or
    // earlier in this step, which precedes this read in program order.
    const int lc = U * min(max(ch, 0), lastch);
    load_vec<U>(Bx + lc, cx);
    load_vec<U>(By + lc, cy);
    // Lane 0 above row 0, or above the previous strip's range (DTW cells
    // there are out of band, i.e. infinite; an in-window LCSS cell there
    // reads its diagonal instead): the border row.
    const bool border = row == 0 || ch >= pcend;
    load_vec<U>(border ? brow : bnd + (lc & hmask), hv);
  }
}

// Unbanded DP with both trajectories staged in LDS: the k_dp schedule with
// its per-step control cut down. A lone wave issues one instruction every
// four cycles whatever its kind, so a step's latency is its instruction
// count; the common step is the cells plus straight-line bookkeeping:
//  * the pair's result, pair fetch, pair start and the exit test sit behind
//    one wave-uniform branch taken only on the steps where some lane needs
//    one; a lane's result is stored from its outputs of the step before,
//    found by its countdown of steps left in the pair;
//  * every LDS operand of step t+1 is read early in step t (LDS latency
//    hides behind the cells). Both trajectories are staged as (x, y)
//    pairs, so one 16-byte read brings two points;
//  * lane W-1 writes every chunk of its row to the halo, idle ones
//    included. A strip lasts D >= W+1 steps: lane 0 of the next strip reads
//    chunk k early in step D+k-1 after the strip started, after lane W-1's
//    write of it in step W-1+k, and the next write of chunk k comes in step
//    D+W-1+k, after the read;
//  * row changes are selects. Rows are single-buffered: a lane uses a row
//    point read in the step before a strip starts, and every lane's last
//    strip of a pair starts before lane 0 refills the buffer. Rows read
//    from global memory need no select: the point read in the step before
//    is that of the lane's current row (its own pair's, set at the start).
template <int U>
__device__ __forceinline__ void load_pts(const float2* p, float (&x)[U], float (&y)[U]) {
  if constexpr (U == 1) {
    const float2 t = *p;
    x[0] = t.x;
    y[0] = t.y;
  } else {
#pragma unroll
    for (int j = 0; j < U; j += 2) {
      const float4 t = *(const float4*)(p + j);
      x[j] = t.x; y[j] = t.y; x[j + 1] = t.z; y[j + 1] = t.w;
    }
  }
}

// Copies n points of a trajectory (sets pad it to whole 16-byte vectors)
// into LDS as (x, y) pairs, the group's W lanes 4 points each per round.
template <int W>
__device__ __forceinline__ void stage_pts(float2* dst, const float* gx, const float* gy, int n, int lg) {
  for (int c = lg; c < ((n + 3) >> 2); c += W) {
    const float4 x = *(const float4*)(gx + 4 * c), y = *(const float4*)(gy + 4 * c);
    *(float4*)(dst + 4 * c) = make_float4(x.x, y.x, x.y, y.y);
    *(float4*)(dst + 4 * c + 2) = make_float4(x.z, y.z, x.w, y.w);
  }
}

// LDS per lane group of k_dp_lean, in floats: the halo row, with SC the two
// column buffers of (x, y) points, with SR the row buffer, then the suffix
// table (sfx entries).
__host__ __device__ constexpr int lean_group_floats(int cmax, int rmax, bool sc, bool sr, int sfx) {
  return cmax + (sc ? 4 * cmax : 0) + (sr ? 2 * rmax : 0) + sfx;
}

// Sum of v over the W lanes of a lane group, in its lane W-1 (other lanes
// hold partial sums): Hillis-Steele steps within each 16-lane row by DPP
// row_shr 1, 2, 4, 8, then rows combined by row_bcast 15 (and 31 for W = 64)
// into the rows that hold a group's last lane. Only the group's own lanes
// need be active: the broadcasts write only rows whose source row is in the
// same group.
template <int CTRL, int ROWS>
__device__ __forceinline__ float add_dpp_rows(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, false));
}
template <int W>
__device__ __forceinline__ float group_sum_last(float v) {
  v = add_dpp_rows<0x111, 0xf>(v);
  v = add_dpp_rows<0x112, 0xf>(v);
  v = add_dpp_rows<0x114, 0xf>(v);
  v = add_dpp_rows<0x118, 0xf>(v);
  if constexpr (W >= 32) v = add_dpp_rows<0x142, 0xa>(v);
  if constexpr (W == 64) v = add_dpp_rows<0x143, 0xc>(v);
  return v;
}

// Suffix bounds of an unbanded DTW pair for early abandoning, built by the
// group at the pair's fetch. Every warping path visits each row i of A (the
// R-point row trajectory) at some cell, whose point cost is at least
// lb_i = sqrt_lo of the squared distance from a_i to the nearest 64-point
// box of B (the column trajectory; the set's superblock boxes, staged here
// into the group's halo row, which no lane reads during a fetch: the pair
// lane W-1 still runs is in its last strip, whose halo writes are never
// read, and the new pair's first strip reads the border row). Entry s of
// the table is the sum of lb_i over the rows of strips s+1, s+2, ...: the
// path's cost after row sW+W-1 is at least that, so the pair's value is at
// least the row's minimum plus it. Strips are summed from the last, each by
// a DPP reduction into lane W-1, in f32 over non-negative terms, so every
// entry is at most the exact sum times (1 + 2^-24)^(R-1) whatever the
// order; the test's discount covers that and the DP's own rounding. S - 1
// entries for S strips; each lane loads its row point of the next strip
// before scanning the boxes for this one. (8-point block boxes, found
// through the superblocks, cut the refine's cells another 10% on the
// 4096 x 1M DTW call but cost more to build than they saved.)
template <int W>
__device__ __forceinline__ void lean_suffix_table(float* tbl, float4* boxes, const float4* bsblk, int C,
                                                  const float* ax, const float* ay, int R, int lg) {
  const int nbox = (((C + 7) >> 3) + 7) >> 3;
  for (int t = lg; t < nbox; t += W) boxes[t] = bsblk[8 * t];
  __builtin_amdgcn_wave_barrier();
  const int S = (R + W - 1) / W;
  float acc = 0.0f;
  int i = (S - 1) * W + lg;
  float px = ax[min(i, R - 1)], py = ay[min(i, R - 1)];
  for (int s = S - 1; s >= 1; --s) {
    const int in = i - W;
    const float nx = ax[min(in, R - 1)], ny = ay[min(in, R - 1)];
    float m = pt_box(px, py, boxes[0]);
    for (int t = 1; t < nbox; ++t) m = fminf(m, pt_box(px, py, boxes[t]));
    acc += group_sum_last<W>(i < R ? sqrt_lo(m) : 0.0f);
    if (lg == W - 1) tbl[s - 1] = acc;
    i = in;
    px = nx;
    py = ny;
  }
  __builtin_amdgcn_wave_barrier();
}

// Next event of a lane of the lean kernels, tested at the top of its loop:
// the group's fetch at rem == lg, the pair's start at rem == 0, the result at
// rem + 1 == res_rem. rem falls by one a step, so the first of those it meets
// is the largest below the count it leaves the event branch with; the loop
// tests rem == ev instead of the three conditions. INT_MIN when none is left.
__device__ __forceinline__ int next_event(int rem, int lg, int res_rem) {
  int ev = INT_MIN;
  if (0 < rem) ev = 0;
  if (lg < rem) ev = max(ev, lg);
  if (res_rem != INT_MIN && res_rem - 1 < rem) ev = max(ev, res_rem - 1);
  return ev;
}

// One step of k_dp_lean's loop for every lane of the wave; TINY: whether
// the square roots test for inputs below 2^-96 (uniform). The column chunk
// and row point only feed the cells' point terms (squared distances, LCSS
// matches), which come first; the next step's chunk and row point are then
// read into the same registers, so the reads have the rest of the step (the
// square roots, the recurrence, the halo traffic) to arrive in. Scheduling
// barriers keep the reads below the point terms and above the rest.
#define TS_LEAN_STEP(TINY)                                                                          \
  do {                                                                                              \
    if (!SR || wrapped) {                                                                           \
      ax = rp.x;                                                                                    \
      ay = rp.y;                                                                                    \
    }                                                                                               \
    /* ---- point terms */                                                                          \
    float pt[U];                                                                                    \
    _Pragma("unroll")                                                                               \
    for (int j = 0; j < U; ++j) {                                                                   \
      const float dx = ax - cx[j], dy = ay - cy[j];                                                 \
      if (MEAS == TS_LCSS) {                                                                        \
        pt[j] = (fabsf(dx) < a.eps && fabsf(dy) < a.eps) ? 1.0f : 0.0f;                             \
      } else {                                                                                      \
        pt[j] = dx * dx + dy * dy;                                                                  \
      }                                                                                             \
    }                                                                                               \
    /* ---- the next step's chunk, row and operands (idle chunks re-read */                         \
    /* the last one: sets pad trajectories to 8 points) */                                          \
    __builtin_amdgcn_sched_barrier(0);                                                              \
    const int kh = k, rowh = row;                                                                   \
    wrapped = k + 1 == D;                                                                           \
    k = wrapped ? 0 : k + 1;                                                                        \
    row = wrapped ? row + W : row;                                                                  \
    if (SR) {                                                                                       \
      rp = rows[min(row, R - 1)];                                                                   \
    } else {                                                                                        \
      const uint32_t ri = (uint32_t)min(row, R - 1);                                                \
      rp = make_float2(Ax[ri], Ay[ri]);                                                             \
    }                                                                                               \
    if (SC) {                                                                                       \
      load_pts<U>(cb + k * U, cx, cy);                                                              \
    } else {                                                                                        \
      const uint32_t ci = (uint32_t)(U * min(k, lastch));                                           \
      load_vec<U>(Bx + ci, cx);                                                                     \
      load_vec<U>(By + ci, cy);                                                                     \
    }                                                                                               \
    __builtin_amdgcn_sched_barrier(0);                                                              \
                                                                                                    \
    /* ---- U cells */                                                                              \
    float up[U], on[U];                                                                             \
    from_prev_lanes<W, U>(up, o, hv, lg);                                                           \
    const bool first = kh == 0;                                                                     \
    float d = first ? dfirst : up_last;                                                             \
    float l = first ? BND : o[U - 1];                                                               \
    if (MEAS == TS_DTW) sqrt_cr<U>(pt, (TINY));                                                     \
    _Pragma("unroll")                                                                               \
    for (int j = 0; j < U; ++j) {                                                                   \
      float v;                                                                                      \
      if (MEAS == TS_DTW) {                                                                         \
        v = pt[j] + fminf(fminf(up[j], l), d);                                                      \
      } else if (MEAS == TS_FRECHET) {                                                              \
        v = fmaxf(pt[j], fminf(fminf(up[j], l), d));                                                \
      } else {                                                                                      \
        v = pt[j] != 0.0f ? d + 1.0f : fmaxf(up[j], l);                                             \
      }                                                                                             \
      on[j] = v;                                                                                    \
      d = up[j];                                                                                    \
      l = v;                                                                                        \
    }                                                                                               \
    up_last = up[U - 1];                                                                            \
    _Pragma("unroll")                                                                               \
    for (int j = 0; j < U; ++j) o[j] = on[j];                                                       \
                                                                                                    \
    /* ---- the next step's halo chunk (see above for its timing), before */                        \
    /* this step's halo store */                                                                    \
    const bool top = row == 0;                                                                      \
    load_vec<U>(top ? brow : bnd + k * U, hv);                                                      \
    dfirst = (MEAS == TS_LCSS || top) ? 0.0f : BND;                                                 \
    if (lg == W - 1) store_vec<U>(bnd + kh * U, on);                                                \
    /* ---- early abandoning: row minima never fall from one row to the */                          \
    /* next (a cell is its point cost, >= 0, plus or max'ed with a */                               \
    /* neighbour's value, and f32 rounding is monotone), and the pair's */                          \
    /* value is at least each of them; when lane W-1's row, not the last, */                        \
    /* ends above tau at its last real chunk, the group drops the pair (NaN */                      \
    /* in its slot) and fetches the next one as if the pair had ended. The */                       \
    /* test comes before the row's idle chunks: during those a lane of the */                       \
    /* next strip may already hold the pair's result (its chunk lastch comes */                     \
    /* D > W-1 steps later than lane W-1's, so not before). Lane W-1 alone */                       \
    /* tests: it starts a pair last (no lane still owes the previous pair's */                      \
    /* result) and never tests in the pair's last strip (where lane 0 may */                        \
    /* already run the next pair). */                                                               \
    if (MEAS != TS_LCSS) {                                                                          \
      float mn = on[0];                                                                             \
      _Pragma("unroll")                                                                             \
      for (int j = 1; j < U; ++j) mn = fminf(mn, on[j]);                                            \
      rowmin = fminf(kh == 0 ? kNoTau : rowmin, mn);                                                \
      bool stop = false;                                                                            \
      if (lg == W - 1 && kh == lastch && rowh + 1 < R) {                                            \
        stop = (MEAS == TS_DTW ? rowmin : sqrtf(rowmin)) > tau_l;                                   \
        /* DTW: the row's minimum plus the rows below it (suffix table) */                          \
        if (MEAS == TS_DTW && disc_l > 0.0f)                                                        \
          stop = stop || (rowmin + sfxt[(uint32_t)rowh / W]) * disc_l > tau_l;                               \
      }                                                                                             \
      const uint64_t ab = __ballot(stop);                                                           \
      if (__builtin_expect(ab != 0, 0) && ((ab >> (g * W + W - 1)) & 1ull)) {                       \
        if (lg == W - 1) {                                                                          \
          a.out[slot] = __int_as_float(0x7fc00000);                                                 \
          atomicAdd(a.cells, (unsigned long long)(0ull - (uint64_t)(R - 1 - rowh) * (uint64_t)C));  \
        }                                                                                           \
        rem = lg + 1;                                                                               \
        ev = lg;                                                                                    \
        res_rem = kNoResult;                                                                        \
        tau_l = kNoTau; /* no second drop before the next pair starts */                            \
      }                                                                                             \
    }                                                                                               \
    --rem;                                                                                          \
  } while (0)

// SC / SR: columns / rows staged in LDS; otherwise read from global memory
// (long trajectories: the launch then needs several waves per SIMD to hide
// the reads, which the throughput regime has). SPLIT (DTW launches of 12 or
// more waves per CU): the loop runs the steps between events in two copies,
// with and without the tiny test, instead of testing the flag every step;
// that is faster at those occupancies and slower below them.
// Register budget: 128 VGPRs, so 4 waves fit a SIMD (launches of up to 16
// waves per CU).
template <int MEAS, int W, int U, bool SC, bool SR, bool SPLIT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_dp_lean(DpArgs a) {
  constexpr int G = 64 / W;
  constexpr float BND = (MEAS == TS_LCSS) ? 0.0f : kInf;
  bool tiny = false;  // a lane's pair has a trajectory with a tiny coordinate (uniform)
  constexpr int kNoResult = INT_MIN;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const int g = lane / W, lg = lane & (W - 1);
  const int cmax = a.cmax, rmax = a.rmax;
  extern __shared__ float4 lds_dyn4[];
  const int hdr = ((nwaves * G + 3) >> 2) << 2;
  int* slot_item = (int*)lds_dyn4 + wv * G + g;
  float* brow = (float*)lds_dyn4 + hdr + wv * 8;
  float* gbase =
      (float*)lds_dyn4 + hdr + nwaves * 8 + (size_t)(wv * G + g) * lean_group_floats(cmax, rmax, SC, SR, a.sfx);
  float* sfxt = gbase + lean_group_floats(cmax, rmax, SC, SR, 0);  // [a.sfx] suffix table
  if (lane < 8) brow[lane] = BND;
  float* bnd = gbase;                                         // [cmax] halo row
  float2* cols = (float2*)(gbase + cmax);                     // SC: [2][cmax] column points
  float2* rows = (float2*)(gbase + (SC ? 5 * cmax : cmax));   // SR: [rmax] row points

  int64_t n_items = a.max_items;
  if (a.nitems) n_items = min<int64_t>(n_items, (int64_t)*a.nitems);
  const int64_t n_head = n_items;
  if (a.ntail) n_items = min<int64_t>(a.max_items, n_head + (int64_t)*a.ntail);

  // Current pair of this lane. Before its first pair lane lg idles lg steps
  // on chunks 0..lg-1 (D = W+1, so k does not wrap).
  int R = 1, C = 1, lastch = 0, D = W + 1, k = 0, row = lg, rem = lg, res_rem = kNoResult, slot = 0;
  int ev = lg;  // the first event: the group's first fetch
  int gpar = 0;
  const float2* cb = cols;  // SC: this lane's column buffer
  // !SC / !SR: the pair's column / row trajectory in global memory
  const float *Bx = a.q.x, *By = a.q.y, *Ax = a.q.x, *Ay = a.q.y;
  bool done = false, wrapped = false;
  // The group's next pair, read once by the whole group at the fetch.
  int fR = 1, fC = 1, f_slot = 0;
  const float *fAx = a.q.x, *fAy = a.q.y, *fBx = a.q.x, *fBy = a.q.y;
  bool f_ok = false, f_live = false;
  bool f_tiny = false, pair_tiny = false;  // fetched / current pair has a tiny coordinate
  // fetched / current pair's abandoning threshold, and the running minimum of
  // the lane's row
  const float kNoTau = __int_as_float(0x7f800000);
  float f_tau = kNoTau, tau_l = kNoTau, rowmin = kNoTau;
  // fetched / current pair's suffix-bound discount (DTW; 0: no suffix test)
  float f_disc = 0.0f, disc_l = 0.0f;
  float o[U], cx[U], cy[U], hv[U];
#pragma unroll
  for (int j = 0; j < U; ++j) { o[j] = 0.0f; hv[j] = BND; cx[j] = 0.0f; cy[j] = 0.0f; }
  float ax = 0.0f, ay = 0.0f, up_last = 0.0f;
  float2 rp = make_float2(0.0f, 0.0f);  // row point of the next strip, read a step ahead
  // diagonal of a row's first cell: D[r-1][-1], 0 on row 0 (and for LCSS)
  float dfirst = (MEAS == TS_LCSS || lg == 0) ? 0.0f : BND;

  for (;;) {
    __builtin_amdgcn_wave_barrier();
    if (__any(rem == ev)) {
      // ---- result: the lane's last step covered the result cell
      if (rem + 1 == res_rem) {
        const int rc = (C - 1) - lastch * U;
        float v = o[0];
#pragma unroll
        for (int j = 1; j < U; ++j)
          if (rc == j) v = o[j];
        if (MEAS == TS_FRECHET) v = sqrtf(v);
        if (MEAS == TS_LCSS) v = v / (float)C;
        a.out[slot] = v;
      }
      // ---- fetch: the group's lane 0 finished its pair (rem of lane lg is
      // lg exactly then); the whole group reads the next pair's descriptor
      // and copies both trajectories into LDS.
      if (rem == lg) {
        if (lg == 0) *slot_item = atomicAdd(a.next, 1);
        __builtin_amdgcn_wave_barrier();
        const int it_idx = *slot_item;
        f_ok = it_idx < n_items;
        f_live = false;
        if (f_ok) {
          const Item it = a.items[it_idx < n_head ? (int64_t)it_idx : a.max_items - 1 - (it_idx - n_head)];
          f_live = it.d >= 0;
          if (f_live) {
            const int n = a.q.len[it.q], m = a.db.len[it.d];
            const int64_t qo = a.q.off[it.q], dof = a.db.off[it.d];
            if (n >= m) {
              fAx = a.q.x + qo; fAy = a.q.y + qo; fBx = a.db.x + dof; fBy = a.db.y + dof; fR = n; fC = m;
            } else {
              fAx = a.db.x + dof; fAy = a.db.y + dof; fBx = a.q.x + qo; fBy = a.q.y + qo; fR = m; fC = n;
            }
            f_slot = it.slot;
            f_tiny = MEAS == TS_DTW && (a.q.tiny[it.q] | a.db.tiny[it.d]) != 0;
            f_tau = (MEAS != TS_LCSS && a.tau) ? a.tau[it.q - a.tau_q0] : kNoTau;
            if (SC) stage_pts<W>(cols + gpar * cmax, fBx, fBy, fC, lg);
            if (SR) stage_pts<W>(rows, fAx, fAy, fR, lg);
            // Suffix bounds (DTW pairs with a finite threshold, coordinates
            // in the range where the bound's rounding argument holds, more
            // than one strip and a table that holds them; uniform in the
            // group). A row's minimum plus its table entry, scaled by
            // 1 - (2R + C + 64) 2^-23, exceeding the threshold proves the
            // pair's value does: the value is at least (row minimum + exact
            // suffix) (1 - 2^-24)^(R+C) (the DP's f32 additions after the
            // row), the table entry at most the exact suffix times
            // (1 + 2^-24)^(R-1), and the test's own sum and product round
            // twice more.
            f_disc = 0.0f;
            const int S = (fR + W - 1) / W;
            if (MEAS == TS_DTW && a.sfx > 0 && !f_tiny && f_tau < kNoTau && S >= 2 && S - 1 <= a.sfx &&
                4 * ((((fC + 7) >> 3) + 7) >> 3) <= cmax) {
              const float4* bsblk = n >= m ? a.db.sblk + (dof >> 3) : a.q.sblk + (qo >> 3);
              lean_suffix_table<W>(sfxt, (float4*)bnd, bsblk, fC, fAx, fAy, fR, lg);
              f_disc = 1.0f - (float)(2 * fR + fC + 64) * 0x1p-23f;
            }
          }
        }
        gpar ^= 1;
      }
      // ---- start: this lane begins the pair fetched lg steps ago
      if (rem == 0) {
        if (!f_ok) {
          done = true;
          rem = -2;  // no further events
          res_rem = kNoResult;
          tau_l = kNoTau;
        } else {
          pair_tiny = f_live && f_tiny;
          tau_l = f_live ? f_tau : kNoTau;
          disc_l = f_live ? f_disc : 0.0f;
          R = f_live ? fR : 1;
          C = f_live ? fC : 1;
          slot = f_slot;
          lastch = (C - 1) / U;
          D = max(lastch + 1, W + 1);
          const int S = (R + W - 1) / W;
          rem = S * D;
          // the lane holding row R-1 in the last strip covers the result
          // cell on chunk lastch, when D - lastch steps are left
          res_rem = (f_live && lg == R - 1 - (S - 1) * W) ? D - lastch : kNoResult;
          k = 0;
          row = lg;
          wrapped = false;
          if (f_live) {
            Ax = fAx; Ay = fAy; Bx = fBx; By = fBy;
          }
          if (SC) {
            cb = cols + (gpar ^ 1) * cmax;
            load_pts<U>(cb, cx, cy);
          } else {
            load_vec<U>(Bx, cx);
            load_vec<U>(By, cy);
          }
          if (SR) {
            const float2 p0 = rows[min(lg, R - 1)];
            ax = p0.x;
            ay = p0.y;
          } else {
            ax = Ax[min(lg, R - 1)];
            ay = Ay[min(lg, R - 1)];
            rp = make_float2(ax, ay);
          }
          load_vec<U>(brow, hv);
          dfirst = (MEAS == TS_LCSS || lg == 0) ? 0.0f : BND;
        }
      }
      ev = next_event(rem, lg, res_rem);
      tiny = __any(pair_tiny);
      if (__all(done)) break;
    }
    if constexpr (SPLIT) {
      // Steps until some lane's next event, with the test compiled in or
      // out: the wave's tiny flag only changes at events. (Written as a
      // lambda called twice, which measured faster than two plain loops.)
      const auto steps = [&](auto tc) __attribute__((always_inline)) {
        const bool tiny_now = tc;
        do {
          TS_LEAN_STEP(tiny_now);
          __builtin_amdgcn_wave_barrier();
        } while (!__any(rem == ev));
      };
      if (tiny) {
        steps(std::true_type{});
      } else {
        steps(std::false_type{});
      }
    } else {
      TS_LEAN_STEP(tiny);
    }
  }
}
#undef TS_LEAN_STEP

// Banded DTW and windowed LCSS with the lean step. The strips' chunk ranges
// come from a table the group writes at the fetch (the same strip_range as
// k_dp, each strip lasting max(range, W + shift + 1) steps, so the early
// halo read of chunk k of strip s+1 and lane 0's diagonal at a strip's
// start follow lane W-1's write of them in strip s); a lane reads its next
// strip's entry every step and takes it on a wrap. The table, like the row
// buffer, is single-buffered: every lane has read its last entry before
// lane 0 refills it. Band cells, window cells and the halo ring as in k_dp;
// the halo chunk of a step is its chunk clamped to the row's last one.
// ST: both trajectories staged in LDS (short pairs), else both read from
// global memory.
// Last chunk (of U columns) holding an in-band cell of row r of a banded
// pair: column min(C-1, floor(r*C/R) + wb) (r*C < 2^31, as in strip_range).
template <int U>
__device__ __forceinline__ int band_last_chunk(int r, int R, int C, int wb) {
  return min(C - 1, (int)(((uint32_t)r * (uint32_t)C) / (uint32_t)R) + wb) / U;
}

__host__ __device__ constexpr int lean_r_group_floats(int cmax, int ccap, int rmax, int smax, bool st) {
  return cmax + (st ? 4 * ccap + 2 * rmax : 0) + ((2 * smax + 3) & ~3);  // whole float4s per group
}

template <int MEAS, int W, int U, bool BAND, bool DELTA, bool ST>
__global__ void __launch_bounds__(256) k_dp_lean_r(DpArgs a) {
  static_assert(BAND || DELTA, "ranged pairs only");
  constexpr int G = 64 / W;
  constexpr float BND = (MEAS == TS_LCSS) ? 0.0f : kInf;
  bool tiny = false;  // a lane's pair has a trajectory with a tiny coordinate (uniform)
  constexpr int kNoResult = INT_MIN;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const int g = lane / W, lg = lane & (W - 1);
  const int cmax = a.cmax, hmask = a.hmask, ccap = a.ccap, rmax = a.rmax, smax = a.smax;
  const int delta = a.delta;
  extern __shared__ float4 lds_dyn4[];
  const int hdr = ((nwaves * G + 3) >> 2) << 2;
  int* slot_item = (int*)lds_dyn4 + wv * G + g;
  float* brow = (float*)lds_dyn4 + hdr + wv * 8;
  float* gbase = (float*)lds_dyn4 + hdr + nwaves * 8 +
                 (size_t)(wv * G + g) * lean_r_group_floats(cmax, ccap, rmax, smax, ST);
  if (lane < 8) brow[lane] = BND;
  float* bnd = gbase;                                              // [cmax] halo ring
  float2* cols = (float2*)(gbase + cmax);                          // ST: [2][ccap] column points
  float2* rows = (float2*)(gbase + cmax + 4 * ccap);               // ST: [rmax] row points
  int2* tbl = (int2*)(gbase + cmax + (ST ? 4 * ccap + 2 * rmax : 0));  // [smax] (cstart, cend)

  int64_t n_items = a.max_items;
  if (a.nitems) n_items = min<int64_t>(n_items, (int64_t)*a.nitems);
  const int64_t n_head = n_items;
  if (a.ntail) n_items = min<int64_t>(a.max_items, n_head + (int64_t)*a.ntail);

  // Current pair of this lane; before its first pair lane lg idles lg steps
  // on one strip of W+1 chunks.
  int R = 1, C = 1, lastch = 0, S = 1, rem = lg, res_rem = kNoResult, slot = 0;
  int ev = lg;  // the first event: the group's first fetch
  int strip = 0, k = 0, cs = 0, ce = W + 1, pcs = 0, pce = 0, row = lg;
  int bw = 0, bd = 0;
  int2 nxt = make_int2(0, W + 1);  // the next strip's (cstart, cend)
  // early abandoning (banded DTW): the fetched / current pair's threshold,
  // the band's half-width, the last chunk holding an in-band cell of the
  // lane's row, and the running minimum of the row
  constexpr bool kAbandon = MEAS == TS_DTW && BAND;
  const float kNoTau = __int_as_float(0x7f800000);
  float f_tau = kNoTau, tau_l = kNoTau, rowmin = kNoTau;
  int wbl = 0, rlast = 0;
  int gpar = 0;
  const float2* cb = cols;
  const float *Ax = a.q.x, *Ay = a.q.y, *Bx = a.q.x, *By = a.q.y;
  bool done = false, wrapped = false;
  // the group's next pair (read once by the whole group at the fetch)
  int fR = 1, fC = 1, f_slot = 0, f_total = 0, f_ires = 0, f_lgres = -1;
  const float *fAx = a.q.x, *fAy = a.q.y, *fBx = a.q.x, *fBy = a.q.y;
  bool f_ok = false, f_live = false;
  bool f_tiny = false, pair_tiny = false;  // fetched / current pair has a tiny coordinate
  float o[U], cx[U], cy[U], hv[U];
#pragma unroll
  for (int j = 0; j < U; ++j) { o[j] = 0.0f; hv[j] = BND; cx[j] = 0.0f; cy[j] = 0.0f; }
  float ax = 0.0f, ay = 0.0f, up_last = 0.0f, hdiag = BND;
  float2 rp = make_float2(0.0f, 0.0f);

  for (;;) {
    __builtin_amdgcn_wave_barrier();
    if (__any(rem == ev)) {
      // ---- result: the lane's last step covered the result cell
      if (rem + 1 == res_rem) {
        const int rc = (C - 1) - lastch * U;
        float v = o[0];
#pragma unroll
        for (int j = 1; j < U; ++j)
          if (rc == j) v = o[j];
        if (MEAS == TS_LCSS) v = v / (float)C;
        a.out[slot] = v;
      }
      // ---- fetch (the whole group): descriptor, staged trajectories, the
      // strip table, the pair's step total and its result step
      if (rem == lg) {
        if (lg == 0) *slot_item = atomicAdd(a.next, 1);
        __builtin_amdgcn_wave_barrier();
        const int it_idx = *slot_item;
        f_ok = it_idx < n_items;
        f_live = false;
        if (f_ok) {
          const Item it = a.items[it_idx < n_head ? (int64_t)it_idx : a.max_items - 1 - (it_idx - n_head)];
          f_live = it.d >= 0;
          if (f_live) {
            const int n = a.q.len[it.q], m = a.db.len[it.d];
            const int64_t qo = a.q.off[it.q], dof = a.db.off[it.d];
            if (n >= m) {
              fAx = a.q.x + qo; fAy = a.q.y + qo; fBx = a.db.x + dof; fBy = a.db.y + dof; fR = n; fC = m;
            } else {
              fAx = a.db.x + dof; fAy = a.db.y + dof; fBx = a.q.x + qo; fBy = a.q.y + qo; fR = m; fC = n;
            }
            f_slot = it.slot;
            f_tiny = MEAS == TS_DTW && (a.q.tiny[it.q] | a.db.tiny[it.d]) != 0;
            f_tau = (kAbandon && a.tau) ? a.tau[it.q - a.tau_q0] : kNoTau;
            if (ST) {
              stage_pts<W>(cols + gpar * ccap, fBx, fBy, fC, lg);
              stage_pts<W>(rows, fAx, fAy, fR, lg);
            }
          }
        }
        const int R2 = f_live ? fR : 1, C2 = f_live ? fC : 1;
        const int lastch2 = (C2 - 1) / U;
        const int wb2 = BAND ? band_half_width(a.band, R2, C2) : 0;
        const int res_row2 = DELTA ? min(R2 - 1, C2 - 1 + delta) : R2 - 1;
        const int S2 = DELTA ? res_row2 / W + 1 : (R2 + W - 1) / W;
        const int sres = res_row2 / W;
        int part = 0, before = 0, at_res = 0;
        for (int s2 = lg; s2 < S2; s2 += W) {
          int clo, chi, nlo, nhi;
          strip_range<BAND, DELTA, W, U>(s2, R2, C2, wb2, delta, lastch2, clo, chi);
          nlo = clo;
          if (s2 + 1 < S2) strip_range<BAND, DELTA, W, U>(s2 + 1, R2, C2, wb2, delta, lastch2, nlo, nhi);
          const int dur = max(chi - clo + 1, W + (nlo - clo) + 1);
          tbl[s2] = make_int2(clo, clo + dur);
          part += dur;
          if (s2 < sres) before += dur;
          if (s2 == sres) at_res = lastch2 - clo;  // the result chunk's step within its strip
        }
#pragma unroll
        for (int sh = W / 2; sh > 0; sh >>= 1) {  // within the group
          part += __shfl_xor(part, sh);
          before += __shfl_xor(before, sh);
          at_res += __shfl_xor(at_res, sh);
        }
        f_total = part;
        f_ires = before + at_res;
        f_lgres = f_live ? res_row2 - sres * W : -1;
        gpar ^= 1;
      }
      // ---- start: this lane begins the pair fetched lg steps ago
      if (rem == 0) {
        if (!f_ok) {
          done = true;
          rem = -2;
          res_rem = kNoResult;
          tau_l = kNoTau;
        } else {
          pair_tiny = f_live && f_tiny;
          R = f_live ? fR : 1;
          C = f_live ? fC : 1;
          slot = f_slot;
          lastch = (C - 1) / U;
          if (BAND) {
            wbl = band_half_width(a.band, R, C);
            bw = wbl * R;
          }
          tau_l = f_live ? f_tau : kNoTau;
          if (kAbandon) rlast = band_last_chunk<U>(row, R, C, wbl);
          S = DELTA ? min(R - 1, C - 1 + delta) / W + 1 : (R + W - 1) / W;
          rem = f_total;
          res_rem = lg == f_lgres ? f_total - f_ires : kNoResult;
          strip = 0;
          row = lg;
          wrapped = false;
          const int2 t0 = tbl[0];
          cs = t0.x;
          ce = t0.y;
          pcs = 0;
          pce = 0;
          k = cs;
          nxt = tbl[min(1, S - 1)];
          if (f_live) {
            Ax = fAx; Ay = fAy; Bx = fBx; By = fBy;
          }
          if (BAND) bd = row * C - cs * U * R;
          const int lc = U * min(cs, lastch);
          if (ST) {
            cb = cols + (gpar ^ 1) * ccap;
            load_pts<U>(cb + lc, cx, cy);
            const float2 p0 = rows[min(lg, R - 1)];
            ax = p0.x;
            ay = p0.y;
          } else {
            load_vec<U>(Bx + lc, cx);
            load_vec<U>(By + lc, cy);
            ax = Ax[min(lg, R - 1)];
            ay = Ay[min(lg, R - 1)];
          }
          load_vec<U>(brow, hv);
        }
      }
      ev = next_event(rem, lg, res_rem);
      tiny = __any(pair_tiny);
      if (__all(done)) break;
    }
    if (wrapped) {
      ax = rp.x;
      ay = rp.y;
    }

    // ---- U cells (as in k_dp)
    float up[U], on[U];
#pragma unroll
    for (int j = 0; j < U; ++j) up[j] = from_prev_lane<W>(o[j], hv[j], lg);
    const bool first = k == cs;
    float d = first ? ((row == 0) ? (cs == 0 ? 0.0f : BND) : (lg == 0 ? hdiag : BND)) : up_last;
    if (MEAS == TS_LCSS && first && cs == 0) d = 0.0f;
    float l = first ? BND : o[U - 1];
    float dist[U];
    if (MEAS == TS_DTW) {
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const float dx = ax - cx[j], dy = ay - cy[j];
        dist[j] = dx * dx + dy * dy;
      }
      sqrt_cr<U>(dist, tiny);
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const float dx = ax - cx[j], dy = ay - cy[j];
      float v;
      if (MEAS == TS_DTW) {
        v = dist[j] + fminf(fminf(up[j], l), d);
        if (BAND) {
          const int bj = bd - j * R;
          if ((bj < 0 ? -bj : bj) > bw) v = kInf;
        }
      } else {
        bool match = fabsf(dx) < a.eps && fabsf(dy) < a.eps;
        float uu = up[j], ll = l;
        if (DELTA) {
          const int dd = row - (k * U + j);
          match = match && (dd < 0 ? -dd : dd) <= delta;
          if (dd - 1 > delta || 1 - dd > delta) uu = d;  // (row-1, col) outside
          if (dd + 1 > delta || -1 - dd > delta) ll = d;  // (row, col-1) outside
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

    // ---- early abandoning (banded DTW), as in k_dp_lean: row minima never
    // fall from row to row (cells outside the band hold 1e20f, more than
    // any in-band cell they could stand for); lane W-1's row, not the last,
    // is tested at its last in-band chunk, at least two steps before a lane
    // of a later strip can compute the result cell (a strip lasts at least
    // W + 1 + its successor's shift steps); the group drops the pair (NaN in
    // its slot) and fetches the next one as if the pair had ended.
    if (kAbandon) {
      float mn = on[0];
#pragma unroll
      for (int j = 1; j < U; ++j) mn = fminf(mn, on[j]);
      rowmin = fminf(first ? kNoTau : rowmin, mn);
      const bool stop = lg == W - 1 && k == rlast && row + 1 < R && rowmin > tau_l;
      const uint64_t ab = __ballot(stop);
      if (__builtin_expect(ab != 0, 0) && ((ab >> (g * W + W - 1)) & 1ull)) {
        if (lg == W - 1) {
          a.out[slot] = __int_as_float(0x7fc00000);
          // the unfinished rows' in-band cells, at most min(2w+1, C) a row:
          // the launch's cell count errs low, never high (integer arithmetic
          // throughout: the device min template is floating point)
          const uint64_t per_row = (uint64_t)(2 * wbl + 1 < C ? 2 * wbl + 1 : C);
          atomicAdd(a.cells, (unsigned long long)(0ull - (uint64_t)(R - 1 - row) * per_row));
        }
        rem = lg + 1;
        ev = lg;
        res_rem = kNoResult;
        tau_l = kNoTau;  // no second drop before the next pair starts
      }
    }

    // ---- advance; the next step's operands (before this step's halo store)
    const int kh = k;
    wrapped = k + 1 == ce;
    if (wrapped) {
      ++strip;
      row += W;
      if (kAbandon) rlast = band_last_chunk<U>(min(row, R - 1), R, C, wbl);
      pcs = cs;
      pce = ce;
      cs = nxt.x;
      ce = nxt.y;
      k = cs;
      if (BAND) bd = row * C - cs * U * R;
      // lane 0's diagonal for the new strip's first cell: the halo value
      // left of its range, when the previous strip computed it
      const int dc = cs * U - 1;
      hdiag = (dc >= 0 && dc >= pcs * U && dc < pce * U && dc < C) ? bnd[dc & hmask] : BND;
    } else {
      ++k;
      if (BAND) bd -= U * R;
    }
    nxt = tbl[min(strip + 1, S - 1)];
    const int lc = U * min(k, lastch);
    if (ST) {
      rp = rows[min(row, R - 1)];
      load_pts<U>(cb + lc, cx, cy);
    } else {
      rp = make_float2(Ax[min(row, R - 1)], Ay[min(row, R - 1)]);
      load_vec<U>(Bx + lc, cx);
      load_vec<U>(By + lc, cy);
    }
    load_vec<U>((row == 0 || k >= pce) ? brow : bnd + (lc & hmask), hv);
    if (lg == W - 1 && kh <= lastch) store_vec<U>(bnd + ((kh * U) & hmask), on);
    --rem;
  }
}

// Halo of a launch of shape `shape` over bucket b. Without a band: a whole
// row of the longest C in the bucket (at least W+1 chunks, the shortest
// strip of k_dp_lean). With a band: a row
// of the longest C when that is short, else a ring of 2^k floats covering a
// strip's range, which spans at most W-1+2w+2U columns (band half-width w,
// 2w+1 <= width) or W+1 chunks plus the range's shift, plus the chunk read
// before the strip starts.
DpHalo dp_halo(int shape, int b, const MeasureArgs& m, int32_t cmax_c, int32_t wmax_c) {
  const DpShape& k = kDpShapeList[shape];
  if (k.W > 64) {
    // bucket b bounds the shorter length of unbanded pairs, and the visited
    // width of banded and windowed ones (whose shorter length it does not bound)
    const int32_t c = dp_ranged(m) ? cmax_c : std::min<int32_t>(cmax_c, kDpBucketCap[b]);
    return DpHalo{dp_mw_hcap(k.U, m, c, std::min<int32_t>(wmax_c, kDpBucketCap[b])), -1};
  }
  if (!dp_ranged(m)) {
    const int32_t c = std::min<int32_t>(cmax_c, kDpBucketCap[b]);
    const int32_t chunks = std::max((c + k.U - 1) / k.U, k.W + 1);
    return DpHalo{((chunks * k.U + 3) / 4) * 4, -1};
  }
  const int32_t chunks = std::max((cmax_c + k.U - 1) / k.U, k.W);
  const int32_t full = ((chunks * k.U + 3) / 4) * 4;
  const int32_t wd = std::min<int32_t>(wmax_c, kDpBucketCap[b]);
  const int32_t need = std::max(k.W + wd + 4 * k.U, k.W * k.U + k.W + 4 * k.U);
  int32_t h = 4;
  while (h < need) h *= 2;
  if (h >= full) return DpHalo{full, -1};
  return DpHalo{h, h - 1};
}

// Launches k_dp (LEAN 0: bands and windows) or k_dp_lean (LEAN 1: both trajectories
// staged, 2: columns staged and rows from global memory, 3: both from global
// memory) as a persistent grid of grid_waves waves, blocks of up to 4 waves
// whose group areas fit in 64 KiB of LDS.
template <int MEAS, int W, int U, bool BAND, bool DELTA, int LEAN>
static hipError_t launch_dp_t(hipStream_t st, const DpArgs& a, int grid_waves) {
  constexpr int G = 64 / W;
  const size_t gf = LEAN == 0   ? (size_t)dp_group_floats(a.cmax)
                    : LEAN >= 4 ? (size_t)lean_r_group_floats(a.cmax, a.ccap, a.rmax, a.smax, LEAN == 4)
                                : (size_t)lean_group_floats(a.cmax, a.rmax, LEAN != 3, LEAN == 1, a.sfx);
  const size_t per_wave = (size_t)G * gf * sizeof(float) + G * sizeof(int) + 32;
  int wpb = 4;
  while (wpb > 1 && per_wave * wpb + 64 > 65536) --wpb;
  if (per_wave + 64 > 65536) return hipErrorInvalidValue;
  const size_t lds = (size_t)(((wpb * G + 3) >> 2) << 2) * sizeof(int) + (size_t)wpb * 8 * sizeof(float) +
                     (size_t)wpb * G * gf * sizeof(float);
  int blocks = (grid_waves + wpb - 1) / wpb;
  if (blocks < 1) blocks = 1;
  if constexpr (LEAN >= 4) {
    hipLaunchKernelGGL((k_dp_lean_r<MEAS, W, U, BAND, DELTA, LEAN == 4>), dim3(blocks), dim3(64 * wpb), lds, st, a);
  } else if constexpr (LEAN != 0) {
    static_assert(!BAND && !DELTA, "the lean kernel is unbanded");
    // DTW launches of 12 or more waves per CU (the throughput regime) take
    // the loop split by the tiny flag; fewer waves run faster without it.
    if (MEAS == TS_DTW && grid_waves >= 12 * device_cu_count())
      hipLaunchKernelGGL((k_dp_lean<MEAS, W, U, LEAN != 3, LEAN == 1, MEAS == TS_DTW>), dim3(blocks), dim3(64 * wpb),
                         lds, st, a);
    else
      hipLaunchKernelGGL((k_dp_lean<MEAS, W, U, LEAN != 3, LEAN == 1, false>), dim3(blocks), dim3(64 * wpb), lds, st,
                         a);
  } else {
    hipLaunchKernelGGL((k_dp<MEAS, W, U, BAND, DELTA>), dim3(blocks), dim3(64 * wpb), lds, st, a);
  }
  return hipGetLastError();
}

// Whether a launch of `shape` with this halo fits the 64 KiB of LDS a block
// may use (one wave per block at least; staging is only chosen when it fits
// in 20 KiB per wave).
bool dp_fits(int shape, DpHalo halo, int32_t rmax) {
  if (kDpShapeList[shape].W > 64) return dp_mw_lds_bytes(rmax, halo.len) <= (size_t)dp_mw_lds_max();
  const int G = 64 / kDpShapeList[shape].W;
  return (size_t)G * dp_group_floats(halo.len) * sizeof(float) + G * sizeof(int) + 32 + 64 <= 65536;
}

// Unbanded pairs run k_dp_lean, staging in LDS what fits in 20 KiB per wave
// (room for 8 waves per CU): both trajectories, else the columns, else
// neither. Bands and windows run k_dp, columns from global memory (a pair's
// column trajectory can be far longer than the band the LDS holds).
template <int MEAS, int W, int U, bool BAND, bool DELTA>
static hipError_t launch_dp_s(hipStream_t st, DpArgs a, int grid_waves) {
  constexpr int G = 64 / W;
  if constexpr (BAND || DELTA) {
    // k_dp_lean_r: both trajectories staged, else both from global memory;
    // k_dp when even the halo ring and strip table do not fit
    a.smax = (a.rmax + W - 1) / W;
    a.ccap = ((a.ccap + U - 1) / U) * U;  // whole chunks: the last one is read U points wide
    if ((size_t)G * lean_r_group_floats(a.cmax, a.ccap, a.rmax, a.smax, true) * sizeof(float) <= 20480)
      return launch_dp_t<MEAS, W, U, BAND, DELTA, 4>(st, a, grid_waves);
    if ((size_t)G * lean_r_group_floats(a.cmax, 0, 0, a.smax, false) * sizeof(float) <= 20480)
      return launch_dp_t<MEAS, W, U, BAND, DELTA, 5>(st, a, grid_waves);
    return launch_dp_t<MEAS, W, U, BAND, DELTA, 0>(st, a, grid_waves);
  } else {
    // DTW with thresholds: a suffix table of one entry per strip of the
    // longest row trajectory, up to 256 (longer pairs test row minima only)
    a.sfx = (MEAS == TS_DTW && a.tau) ? std::min(((a.rmax + W - 1) / W + 3) / 4 * 4, 256) : 0;
    if ((size_t)G * lean_group_floats(a.cmax, a.rmax, true, true, a.sfx) * sizeof(float) <= 20480)
      return launch_dp_t<MEAS, W, U, BAND, DELTA, 1>(st, a, grid_waves);
    a.rmax = 0;
    if ((size_t)G * lean_group_floats(a.cmax, 0, true, false, a.sfx) * sizeof(float) <= 20480)
      return launch_dp_t<MEAS, W, U, BAND, DELTA, 2>(st, a, grid_waves);
    return launch_dp_t<MEAS, W, U, BAND, DELTA, 3>(st, a, grid_waves);
  }
}

template <int MEAS, bool BAND, bool DELTA>
static hipError_t launch_dp_b(hipStream_t st, const DpArgs& a, int shape, int grid_waves) {
  switch (shape) {
    case 0: return launch_dp_s<MEAS, 16, 1, BAND, DELTA>(st, a, grid_waves);
    case 1: return launch_dp_s<MEAS, 16, 2, BAND, DELTA>(st, a, grid_waves);
    case 2: return launch_dp_s<MEAS, 16, 4, BAND, DELTA>(st, a, grid_waves);
    case 3: return launch_dp_s<MEAS, 32, 2, BAND, DELTA>(st, a, grid_waves);
    case 4: return launch_dp_s<MEAS, 32, 4, BAND, DELTA>(st, a, grid_waves);
    case 5: return launch_dp_s<MEAS, 64, 2, BAND, DELTA>(st, a, grid_waves);
    case 6: return launch_dp_s<MEAS, 64, 4, BAND, DELTA>(st, a, grid_waves);
    case 7: return launch_dp_s<MEAS, 32, 8, BAND, DELTA>(st, a, grid_waves);
    case 8: return launch_dp_s<MEAS, 64, 8, BAND, DELTA>(st, a, grid_waves);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_dp(hipStream_t st, MeasureArgs m, int shape, SetView q, SetView db, const Item* items,
                     const int32_t* nitems, const int32_t* ntail, int64_t max_items, int32_t* next, float* out,
                     DpHalo halo, int32_t rmax_c, int32_t ccap_c, int grid_waves, const float* tau, int32_t tau_q0,
                     unsigned long long* cells) {
  // row buffer: the longest trajectory of either set, in whole float4s;
  // column buffer (ranged): the longest shorter trajectory, likewise
  const int32_t rmax = ((std::max(rmax_c, 1) + 3) / 4) * 4;
  const int32_t ccap = ((std::max(ccap_c, 1) + 3) / 4) * 4;
  if (kDpShapeList[shape].W > 64)
    return launch_dp_mw(st, m, kDpShapeList[shape].W / 64, kDpShapeList[shape].U, q, db, items, nitems, ntail,
                        max_items, next, out, halo.len, rmax_c, grid_waves, tau, tau_q0, cells);
  DpArgs a{q,          db,      items,      nitems,  ntail,   max_items, next, out,   halo.len, halo.mask, rmax,
           ccap,       0,       m.band,     m.delta, m.eps,   tau,       tau_q0, cells, 0};
  switch (m.measure) {
    case TS_DTW:
      return m.band != 0 ? launch_dp_b<TS_DTW, true, false>(st, a, shape, grid_waves)
                        : launch_dp_b<TS_DTW, false, false>(st, a, shape, grid_waves);
    case TS_FRECHET: return launch_dp_b<TS_FRECHET, false, false>(st, a, shape, grid_waves);
    case TS_LCSS:
      return m.delta >= 0 ? launch_dp_b<TS_LCSS, false, true>(st, a, shape, grid_waves)
                          : launch_dp_b<TS_LCSS, false, false>(st, a, shape, grid_waves);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ts
