// reduce.hip -- exact Hausdorff and lock-step Euclidean refine.
//
// Hausdorff (hausdorff.cpp:11-109, 172-182): H^2 = max(max_a min_b |a-b|^2,
// max_b min_a |a-b|^2) with |a-b|^2 = dx*dx + dy*dy in f32, result sqrtf.
// min/max do not round, so any evaluation order gives the reference value;
// the reference's grid accelerator and early prune are value-preserving.
// One wave per pair: lanes own points of the source side, the target side
// streams through uniform (scalar) loads.
//
// Euclidean (euclidean.cpp:27-40, :113): sqrt of a sequential double sum of
// dx*dx + dy*dy over equal-length trajectories, then cast to float. One lane
// per pair folds in the reference's order, so the double sum is bit-identical.
#include "ts_internal.h"

namespace ts {

__device__ float directed_h2_wave(const float* ax, const float* ay, int na, const float* bx, const float* by,
                                  int nb, int lane) {
  float best = 0.0f;
  for (int i0 = 0; i0 < na; i0 += 64) {
    const int i = i0 + lane;
    const int ii = min(i, na - 1);
    const float px = ax[ii], py = ay[ii];
    float mn = kInf;
    int j = 0;
    for (; j + 4 <= nb; j += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float dx = px - bx[j + u], dy = py - by[j + u];
        mn = fminf(mn, dx * dx + dy * dy);
      }
    }
    for (; j < nb; ++j) {
      const float dx = px - bx[j], dy = py - by[j];
      mn = fminf(mn, dx * dx + dy * dy);
    }
    if (i < na) best = fmaxf(best, mn);
  }
  for (int sh = 32; sh > 0; sh >>= 1) best = fmaxf(best, __shfl_xor(best, sh));
  return best;
}

// Wave-uniform copies: the pair's index, lengths and offsets are the same on
// every lane; readfirstlane moves them to scalar registers so the target loops
// are scalar loops over scalar loads of the target points.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni64(int64_t v) {
  return (int64_t)(((uint64_t)(uint32_t)uni((int)(v >> 32)) << 32) | (uint32_t)uni((int)v));
}

// One wave per work item (no fetch loop): wave w of the grid takes item w.
__global__ void __launch_bounds__(256) k_hausdorff(SetView q, SetView db, const Item* items, const int32_t* nitems,
                                                   int64_t max_items, float* out) {
  const int lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  int64_t n_items = max_items;
  if (nitems) n_items = min<int64_t>(n_items, (int64_t)*nitems);
  if (w >= n_items) return;
  const int64_t it_idx = uni64(w);
  const int iq = uni(items[it_idx].q), id = uni(items[it_idx].d), slot = uni(items[it_idx].slot);
  if (id < 0) return;
  const int n = uni(q.len[iq]), m = uni(db.len[id]);
  const int64_t qo = uni64(q.off[iq]), dof = uni64(db.off[id]);
  const float h1 = directed_h2_wave(q.x + qo, q.y + qo, n, db.x + dof, db.y + dof, m, lane);
  const float h2 = directed_h2_wave(db.x + dof, db.y + dof, m, q.x + qo, q.y + qo, n, lane);
  if (lane == 0) out[slot] = sqrtf(fmaxf(h1, h2));
}

hipError_t launch_hausdorff(hipStream_t st, SetView q, SetView db, const Item* items, const int32_t* nitems,
                            int64_t max_items, float* out) {
  if (max_items == 0) return hipSuccess;
  hipLaunchKernelGGL(k_hausdorff, dim3((unsigned)((max_items + 3) / 4)), dim3(256), 0, st, q, db, items, nitems,
                     max_items, out);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_euclidean(SetView q, SetView db, const Item* items, const int32_t* nitems,
                                                   int64_t max_items, float* out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int64_t n_items = max_items;
  if (nitems) n_items = min<int64_t>(n_items, (int64_t)*nitems);
  if (i >= n_items) return;
  const Item it = items[i];
  if (it.d < 0) return;
  const int n = q.len[it.q], m = db.len[it.d];
  if (n != m) {
    out[it.slot] = __int_as_float(0x7f800000);
    return;
  }
  const float *qx = q.x + q.off[it.q], *qy = q.y + q.off[it.q];
  const float *cx = db.x + db.off[it.d], *cy = db.y + db.off[it.d];
  double s = 0.0;
  for (int j = 0; j < n; ++j) {
    const double dx = (double)qx[j] - (double)cx[j];
    const double dy = (double)qy[j] - (double)cy[j];
    s += dx * dx + dy * dy;
  }
  out[it.slot] = (float)sqrt(s);
}

hipError_t launch_euclidean(hipStream_t st, SetView q, SetView db, const Item* items, const int32_t* nitems,
                            int64_t max_items, float* out) {
  if (max_items == 0) return hipSuccess;
  hipLaunchKernelGGL(k_euclidean, dim3((unsigned)((max_items + 255) / 256)), dim3(256), 0, st, q, db, items, nitems,
                     max_items, out);
  return hipGetLastError();
}

}  // namespace ts
