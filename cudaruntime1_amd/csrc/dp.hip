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
//  * A wave64 is split into 64/W lane groups (W = 64, 32 or 16). Each group
//    runs one pair at a time; lane l of the group owns DP rows l, l+W, ...
//    and sweeps its row left to right one cell per step, one step behind
//    lane l-1, so the anti-diagonal lives in registers.
//  * The cell above arrives from lane l-1 by a DPP wave shift (v_mov_dpp
//    wave_shr:1) of the value it produced on the previous step; the diagonal
//    is the value that arrived on the step before. Lane 0 of a group reads
//    the row above from an LDS row written by lane W-1 during the previous
//    strip of W rows (the pair's "halo" row).
//  * Rows run along the longer trajectory and columns along the shorter, so
//    the halo delay (columns - W + 1 steps) is positive; pairs shorter than
//    W columns are padded with idle columns.
//  * Strips and pairs follow each other without draining the pipeline: after
//    its last row a lane starts its first row of the next pair, which the
//    group's lane 0 fetched from a device-wide atomic work counter. The
//    kernel is persistent: it drains the whole work list in one launch.
#include "ts_internal.h"

namespace ts {

struct DpArgs {
  SetView q, db;
  const Item* items;
  const int32_t* nitems;
  int64_t max_items;
  int32_t* next;
  float* out;
  int32_t cmax;
  int32_t band;
  int32_t delta;
  float eps;
};

// Value of lane-1 (lane 0 receives 0; it is overridden by the caller).
__device__ __forceinline__ float shift_from_prev_lane(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138 /* wave_shr:1 */, 0xf, 0xf, true));
}

template <int MEAS, int W, bool BAND, bool DELTA>
__global__ void __launch_bounds__(256) k_dp(DpArgs a) {
  constexpr int G = 64 / W;
  constexpr float BND = (MEAS == TS_LCSS) ? 0.0f : kInf;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const int g = lane / W, lg = lane & (W - 1);
  extern __shared__ int lds_dyn[];
  int* slot_item = lds_dyn + wv * G + g;
  float* bnd = (float*)(lds_dyn + ((nwaves * G + 3) & ~3)) + (size_t)(wv * G + g) * a.cmax;
  const int cmax = a.cmax;

  int64_t n_items = a.max_items;
  if (a.nitems) n_items = min<int64_t>(n_items, (int64_t)*a.nitems);

  int row = lg, col = -lg;
  bool done = false;
  int R = 1, C = 1, Cp = W, SW = W, slot = 0;
  const float *Ax = a.q.x, *Ay = a.q.y, *Bx = a.q.x, *By = a.q.y;
  float ax = 0.0f, ay = 0.0f;
  float out_prev = 0.0f, up_prev = 0.0f;
  int64_t bd = 0, bw = 0;
  bool live = false;  // current pair is a real pair (not a skipped slot)
  // Column operands: (cbx, cby) for the current step, (nbx, nby) for the
  // next one; loads are issued two steps before their use.
  float cbx = Bx[0], cby = By[0], nbx = cbx, nby = cby, pbv = 0.0f;
  // Point of the row this lane starts next (loaded one row ahead, so no load
  // result is needed in the step that issues it).
  float nax = 0.0f, nay = 0.0f;

  for (;;) {
    // LDS is shared by the lanes of this wave only; the barrier keeps the
    // compiler from moving halo/slot accesses across steps.
    __builtin_amdgcn_wave_barrier();
    // ---- next pair (lane 0 of the group fetches, the others follow it)
    if (!done && col == 0 && row == lg) {
      int it_idx;
      if (lg == 0) {
        it_idx = atomicAdd(a.next, 1);
        *slot_item = it_idx;
      } else {
        it_idx = *slot_item;
      }
      if (it_idx >= n_items) {
        done = true;
      } else {
        const Item it = a.items[it_idx];
        live = it.d >= 0;
        if (live) {
          const int n = a.q.len[it.q], m = a.db.len[it.d];
          const int64_t qo = a.q.off[it.q], dof = a.db.off[it.d];
          if (n >= m) {
            Ax = a.q.x + qo; Ay = a.q.y + qo; Bx = a.db.x + dof; By = a.db.y + dof; R = n; C = m;
          } else {
            Ax = a.db.x + dof; Ay = a.db.y + dof; Bx = a.q.x + qo; By = a.q.y + qo; R = m; C = n;
          }
          slot = it.slot;
        } else {
          R = 1; C = 1;
        }
        Cp = max(C, W);
        SW = ((R + W - 1) / W) * W;
        if (BAND) bw = (int64_t)a.band * R;
        cbx = Bx[0];
        cby = By[0];
        nbx = Bx[min(1, C - 1)];
        nby = By[min(1, C - 1)];
        nax = Ax[min(lg, R - 1)];
        nay = Ay[min(lg, R - 1)];
      }
    }
    if (__all(done)) break;

    // ---- start of a row
    if (col == 0 && !done) {
      ax = nax;
      ay = nay;
      const int nr = min(row + W, R - 1);
      nax = Ax[nr];
      nay = Ay[nr];
      if (BAND) bd = (int64_t)row * C;
    }

    // ---- one cell
    const float bx = cbx, by = cby;
    const float up_sh = shift_from_prev_lane(out_prev);
    const float up = (lg == 0) ? ((row == 0) ? BND : pbv) : up_sh;
    const float diag = (col == 0) ? ((row == 0) ? 0.0f : BND) : up_prev;
    const float left = (col == 0) ? BND : out_prev;
    const float dx = ax - bx, dy = ay - by;
    float o;
    if (MEAS == TS_DTW) {
      const float cost = sqrtf(dx * dx + dy * dy);
      o = cost + fminf(fminf(up, left), diag);
      if (BAND) {
        const int64_t ad = bd < 0 ? -bd : bd;
        if (ad > bw) o = kInf;
      }
    } else if (MEAS == TS_FRECHET) {
      const float d2 = dx * dx + dy * dy;
      o = fmaxf(d2, fminf(fminf(up, left), diag));
    } else {
      bool match = fabsf(dx) < a.eps && fabsf(dy) < a.eps;
      if (DELTA) match = match && abs(row - col) <= a.delta;
      o = match ? diag + 1.0f : fmaxf(up, left);
    }

    // ---- halo row for the next strip, and the pair's result
    if (lg == W - 1 && col >= 0 && !done) bnd[min(col, cmax - 1)] = o;
    if (!done && live && row == R - 1 && col == C - 1) {
      float v = o;
      if (MEAS == TS_FRECHET) v = sqrtf(o);
      if (MEAS == TS_LCSS) v = o / (float)C;
      a.out[slot] = v;
    }

    // ---- advance
    up_prev = up;
    out_prev = o;
    ++col;
    if (BAND) bd -= R;
    if (col == Cp) {
      col = 0;
      row += W;
      if (row >= SW) row = lg;
    }
    // ---- operands of the coming steps. Within a pair the column point of
    // a step is B[col]; a new pair reloads both stages after its fetch. The
    // halo value of column c was written by lane W-1 at least one step ago
    // (columns >= W) or earlier in this step, which precedes this read in
    // program order.
    int col2 = col + 1;
    if (col2 == Cp) col2 = 0;
    const int ncc = min(max(col2, 0), C - 1);
    cbx = nbx;
    cby = nby;
    nbx = Bx[ncc];
    nby = By[ncc];
    pbv = bnd[min(max(col, 0), cmax - 1)];
  }
}

int dp_lanes_for_class(int cls) { return cls == 0 ? 64 : (cls == 1 ? 32 : 16); }

template <int MEAS, int W, bool BAND, bool DELTA>
static hipError_t launch_dp_t(hipStream_t st, const DpArgs& a, int grid_waves) {
  constexpr int G = 64 / W;
  // Waves per block bounded so that the halo rows fit in 64 KiB of LDS.
  const size_t per_wave = (size_t)G * a.cmax * sizeof(float) + G * sizeof(int);
  int wpb = 4;
  while (wpb > 1 && per_wave * wpb + 64 > 65536) --wpb;
  if (per_wave + 64 > 65536) return hipErrorInvalidValue;
  const size_t lds = (size_t)((wpb * G + 3) & ~3) * sizeof(int) + (size_t)wpb * G * a.cmax * sizeof(float);
  int blocks = (grid_waves + wpb - 1) / wpb;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((k_dp<MEAS, W, BAND, DELTA>), dim3(blocks), dim3(64 * wpb), lds, st, a);
  return hipGetLastError();
}

template <int MEAS, bool BAND, bool DELTA>
static hipError_t launch_dp_w(hipStream_t st, const DpArgs& a, int cls, int grid_waves) {
  switch (cls) {
    case 0: return launch_dp_t<MEAS, 64, BAND, DELTA>(st, a, grid_waves);
    case 1: return launch_dp_t<MEAS, 32, BAND, DELTA>(st, a, grid_waves);
    default: return launch_dp_t<MEAS, 16, BAND, DELTA>(st, a, grid_waves);
  }
}

hipError_t launch_dp(hipStream_t st, MeasureArgs m, int cls, SetView q, SetView db, const Item* items,
                     const int32_t* nitems, int64_t max_items, int32_t* next, float* out, int32_t cmax,
                     int grid_waves) {
  DpArgs a{q, db, items, nitems, max_items, next, out, cmax, m.band, m.delta, m.eps};
  switch (m.measure) {
    case TS_DTW:
      return m.band > 0 ? launch_dp_w<TS_DTW, true, false>(st, a, cls, grid_waves)
                        : launch_dp_w<TS_DTW, false, false>(st, a, cls, grid_waves);
    case TS_FRECHET: return launch_dp_w<TS_FRECHET, false, false>(st, a, cls, grid_waves);
    case TS_LCSS:
      return m.delta >= 0 ? launch_dp_w<TS_LCSS, false, true>(st, a, cls, grid_waves)
                          : launch_dp_w<TS_LCSS, false, false>(st, a, cls, grid_waves);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ts
