// reduce.hip -- exact Hausdorff and lock-step Euclidean refine.
//
// Hausdorff (hausdorff.cpp:11-109, 172-182): H^2 = max(max_a min_b |a-b|^2,
// max_b min_a |a-b|^2) with |a-b|^2 = dx*dx + dy*dy in f32, result sqrtf.
// min/max do not round, so any evaluation order gives the reference value;
// the reference's grid accelerator and early prune are value-preserving.
// One wave per pair, pulled from an atomic work counter: lanes own points of
// the source side, the target side streams through uniform (scalar) loads.
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

__global__ void __launch_bounds__(256) k_hausdorff(SetView q, SetView db, const Item* items, const int32_t* nitems,
                                                   int64_t max_items, int32_t* next, float* out) {
  const int lane = threadIdx.x & 63;
  int64_t n_items = max_items;
  if (nitems) n_items = min<int64_t>(n_items, (int64_t)*nitems);
  for (;;) {
    int it_idx = 0;
    if (lane == 0) it_idx = atomicAdd(next, 1);
    it_idx = __shfl(it_idx, 0);
    if (it_idx >= n_items) break;
    const Item it = items[it_idx];
    if (it.d < 0) continue;
    const int n = q.len[it.q], m = db.len[it.d];
    const int64_t qo = q.off[it.q], dof = db.off[it.d];
    const float h1 = directed_h2_wave(q.x + qo, q.y + qo, n, db.x + dof, db.y + dof, m, lane);
    const float h2 = directed_h2_wave(db.x + dof, db.y + dof, m, q.x + qo, q.y + qo, n, lane);
    if (lane == 0) out[it.slot] = sqrtf(fmaxf(h1, h2));
  }
}

hipError_t launch_hausdorff(hipStream_t st, SetView q, SetView db, const Item* items, const int32_t* nitems,
                            int64_t max_items, int32_t* next, float* out, int grid_waves) {
  int blocks = (grid_waves + 3) / 4;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_hausdorff, dim3(blocks), dim3(256), 0, st, q, db, items, nitems, max_items, next, out);
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
