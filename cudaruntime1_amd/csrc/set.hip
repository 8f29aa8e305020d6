// set.hip -- runtime context, errors, device-resident trajectory sets and the
// synthetic random-walk generator.
//
// A set stores its points SoA (x[], y[]), each trajectory contiguous at
// off[i], plus per-trajectory summaries used by the filters: MBR (the
// reference's get_trajectory_mbr, hausdorff.cpp:124-134), first/last point
// and length. The reference instead re-uploads two AoS batches of fixed
// length on every call (dtw.cu:125-138).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <omp.h>

#include "ts_internal.h"

namespace ts {

namespace {
thread_local std::string g_err;
int g_device = -1;
std::mutex g_init_mu;
}  // namespace

void set_error(const std::string& msg) { g_err = msg; }

ts_status fail(ts_status s, const std::string& msg) {
  set_error(msg);
  return s;
}

ts_status hip_fail(hipError_t e, const char* what, const char* file, int line) {
  std::string m = std::string(what) + " failed at " + file + ":" + std::to_string(line) + ": " +
                  hipGetErrorString(e);
  set_error(m);
  return e == hipErrorOutOfMemory ? TS_ERR_OOM : TS_ERR_HIP;
}

int device_cu_count() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) cus = p.multiProcessorCount;
    if (cus <= 0) cus = 256;
  }
  return cus;
}

// Nonzero and below 2^-25 in magnitude.
__device__ __forceinline__ bool tiny_coord(float c) { return c != 0.0f && fabsf(c) < 0x1p-25f; }
// 2^62 or more in magnitude, or not finite: the squared distance to another
// point may be +inf or NaN.
__device__ __forceinline__ bool huge_coord(float c) { return !(fabsf(c) < 0x1p62f); }

// One wave per trajectory: MBR by wave min/max, endpoints from the first and
// last point, whether a point has a tiny (flag bit 1) or huge (bit 2)
// coordinate.
__global__ void __launch_bounds__(256) k_summaries(SetView s, int32_t t0, float4* mbr, float4* ends,
                                                   uint8_t* tiny) {
  const int lane = threadIdx.x & 63;
  const int t = t0 + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= s.count) return;
  const int64_t o = s.off[t];
  const int n = s.len[t];
  float mnx = s.x[o], mxx = mnx, mny = s.y[o], mxy = mny;
  bool tc = false, hc = false;
  for (int i = lane; i < n; i += 64) {
    float px = s.x[o + i], py = s.y[o + i];
    mnx = fminf(mnx, px); mxx = fmaxf(mxx, px);
    mny = fminf(mny, py); mxy = fmaxf(mxy, py);
    tc |= tiny_coord(px) || tiny_coord(py);
    hc |= huge_coord(px) || huge_coord(py);
  }
  const bool any_tc = __any(tc), any_hc = __any(hc);
  for (int sh = 32; sh > 0; sh >>= 1) {
    mnx = fminf(mnx, __shfl_xor(mnx, sh));
    mxx = fmaxf(mxx, __shfl_xor(mxx, sh));
    mny = fminf(mny, __shfl_xor(mny, sh));
    mxy = fmaxf(mxy, __shfl_xor(mxy, sh));
  }
  if (lane == 0) {
    mbr[t] = make_float4(mnx, mxx, mny, mxy);
    ends[t] = make_float4(s.x[o], s.y[o], s.x[o + n - 1], s.y[o + n - 1]);
    tiny[t] = (any_tc ? TS_SET_TINY_COORD : 0) | (any_hc ? TS_SET_HUGE_COORD : 0);
  }
}

static_assert(kPad == 8, "block boxes take trajectories aligned and padded to 8 points");

// Box of each 8-point block (trajectories start on multiples of 8 points
// and are padded with their last point, so a block holds one trajectory's
// points only).
__global__ void __launch_bounds__(256) k_blocks(const float* x, const float* y, int64_t b0, int64_t nblk,
                                                float4* blk) {
  const int64_t b = b0 + (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= nblk) return;
  float mnx = x[8 * b], mxx = mnx, mny = y[8 * b], mxy = mny;
#pragma unroll
  for (int i = 1; i < 8; ++i) {
    const float px = x[8 * b + i], py = y[8 * b + i];
    mnx = fminf(mnx, px); mxx = fmaxf(mxx, px);
    mny = fminf(mny, py); mxy = fmaxf(mxy, py);
  }
  blk[b] = make_float4(mnx, mxx, mny, mxy);
}

hipError_t launch_summaries(hipStream_t st, SetView s, float4* mbr, float4* ends, uint8_t* tiny, int32_t t0) {
  if (s.count <= t0) return hipSuccess;
  hipLaunchKernelGGL(k_summaries, dim3((s.count - t0 + 3) / 4), dim3(256), 0, st, s, t0, mbr, ends, tiny);
  return hipGetLastError();
}

// Superblock g: the union of blocks g..g+7 that belong to g's trajectory
// (fewer near its end). One wave per trajectory.
__global__ void __launch_bounds__(256) k_superblocks(SetView s, int32_t t0, float4* sblk) {
  const int lane = threadIdx.x & 63;
  const int t = t0 + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= s.count) return;
  const int64_t b0 = s.off[t] >> 3;
  const int nb = (s.len[t] + 7) >> 3;
  for (int r = lane; r < nb; r += 64) {
    float4 u = s.blk[b0 + r];
    for (int i = 1; i < 8 && r + i < nb; ++i) {
      const float4 v = s.blk[b0 + r + i];
      u.x = fminf(u.x, v.x); u.y = fmaxf(u.y, v.y); u.z = fminf(u.z, v.z); u.w = fmaxf(u.w, v.w);
    }
    sblk[b0 + r] = u;
  }
}

hipError_t launch_blocks(hipStream_t st, SetView s, int64_t points, float4* blk, float4* sblk, int32_t t0,
                         int64_t point0) {
  const int64_t nblk = points / 8, b0 = point0 / 8;
  if (nblk <= b0 || s.count <= t0) return hipSuccess;
  hipLaunchKernelGGL(k_blocks, dim3((unsigned)((nblk - b0 + 255) / 256)), dim3(256), 0, st, s.x, s.y, b0, nblk, blk);
  hipLaunchKernelGGL(k_superblocks, dim3((s.count - t0 + 3) / 4), dim3(256), 0, st, s, t0, sblk);
  return hipGetLastError();
}

// ------------------------------------------------------------ generator
// SplitMix64 finaliser over a (seed, trajectory, point, lane) counter.
static inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline uint64_t ctr_hash(uint64_t seed, uint64_t a, uint64_t b) {
  return mix64(mix64(mix64(seed) ^ a) ^ (b * 0xD1B54A32D192ED03ull));
}
// Uniform in [0,1) with 53 random bits.
static inline double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

ts_status pad_layout(const ts_point* xy, const int64_t* offsets, int32_t count, int64_t base,
                     std::vector<int32_t>& lens, std::vector<int64_t>& doff, std::vector<float>& hx,
                     std::vector<float>& hy) {
  if (!offsets || !xy || count <= 0) return fail(TS_ERR_INVALID, "set: need count > 0, offsets and points");
  if (offsets[0] != 0) return fail(TS_ERR_INVALID, "set: offsets[0] must be 0");
  lens.resize(count);
  for (int32_t i = 0; i < count; ++i) {
    const int64_t l = offsets[i + 1] - offsets[i];
    if (l < 1 || l > (1 << 24))
      return fail(TS_ERR_INVALID, "set: trajectory " + std::to_string(i) + " has length " + std::to_string(l) +
                                      " (need 1..2^24)");
    lens[i] = (int32_t)l;
  }
  doff.resize(count + 1);
  doff[0] = base;
  for (int32_t i = 0; i < count; ++i) doff[i + 1] = doff[i] + ((lens[i] + kPad - 1) / kPad) * kPad;
  const int64_t P = doff[count] - base;
  hx.resize(P);
  hy.resize(P);
#pragma omp parallel for schedule(static) if (count > 4096)
  for (int32_t i = 0; i < count; ++i) {
    const int64_t src = offsets[i], dst = doff[i] - base;
    const int32_t n = lens[i];
    for (int32_t j = 0; j < n; ++j) { hx[dst + j] = xy[src + j].x; hy[dst + j] = xy[src + j].y; }
    for (int64_t j = n; j < doff[i + 1] - doff[i]; ++j) { hx[dst + j] = xy[src + n - 1].x; hy[dst + j] = xy[src + n - 1].y; }
  }
  return TS_OK;
}

ts_status set_reserve(ts_set* s, int64_t points_cap, int32_t count_cap, bool keep) {
  if (points_cap <= s->cap_points && count_cap <= s->cap_count) return TS_OK;
  points_cap = std::max(points_cap, s->cap_points);
  count_cap = std::max(count_cap, s->cap_count);
  TS_HIP(hipDeviceSynchronize());
  if (!keep) {  // free first: the peak stays at the new size
    (void)hipFree(s->x); (void)hipFree(s->y); (void)hipFree(s->off); (void)hipFree(s->len);
    (void)hipFree(s->mbr); (void)hipFree(s->ends); (void)hipFree(s->blk); (void)hipFree(s->sblk);
    (void)hipFree(s->tiny);
    s->x = s->y = nullptr; s->off = nullptr; s->len = nullptr; s->mbr = s->ends = s->blk = s->sblk = nullptr;
    s->tiny = nullptr;
    s->bytes = 0; s->cap_points = 0; s->cap_count = 0;
    s->count = 0; s->dev_points = 0;
  }
  ts_set n;
  hipError_t e;
#define TS_ALLOC(ptr, nbytes)                                                        \
  e = hipMalloc((void**)&(ptr), (nbytes));                                           \
  if (e != hipSuccess) {                                                             \
    (void)hipFree(n.x); (void)hipFree(n.y); (void)hipFree(n.off); (void)hipFree(n.len); \
    (void)hipFree(n.mbr); (void)hipFree(n.ends); (void)hipFree(n.blk); (void)hipFree(n.sblk); \
    (void)hipFree(n.tiny);                                                           \
    return hip_fail(e, "hipMalloc(set)", __FILE__, __LINE__);                        \
  }                                                                                  \
  n.bytes += (int64_t)(nbytes);
  TS_ALLOC(n.x, points_cap * sizeof(float));
  TS_ALLOC(n.y, points_cap * sizeof(float));
  TS_ALLOC(n.off, (count_cap + 1) * sizeof(int64_t));
  TS_ALLOC(n.len, count_cap * sizeof(int32_t));
  TS_ALLOC(n.mbr, count_cap * sizeof(float4));
  TS_ALLOC(n.ends, count_cap * sizeof(float4));
  TS_ALLOC(n.blk, (points_cap / 8) * sizeof(float4));
  TS_ALLOC(n.sblk, (points_cap / 8) * sizeof(float4));
  TS_ALLOC(n.tiny, count_cap * sizeof(uint8_t));
#undef TS_ALLOC
  if (keep && s->count > 0) {
    const int64_t P = s->dev_points;
    const int32_t C = s->count;
    TS_HIP(hipMemcpy(n.x, s->x, P * sizeof(float), hipMemcpyDeviceToDevice));
    TS_HIP(hipMemcpy(n.y, s->y, P * sizeof(float), hipMemcpyDeviceToDevice));
    TS_HIP(hipMemcpy(n.off, s->off, (C + 1) * sizeof(int64_t), hipMemcpyDeviceToDevice));
    TS_HIP(hipMemcpy(n.len, s->len, C * sizeof(int32_t), hipMemcpyDeviceToDevice));
    TS_HIP(hipMemcpy(n.mbr, s->mbr, C * sizeof(float4), hipMemcpyDeviceToDevice));
    TS_HIP(hipMemcpy(n.ends, s->ends, C * sizeof(float4), hipMemcpyDeviceToDevice));
    TS_HIP(hipMemcpy(n.blk, s->blk, (P / 8) * sizeof(float4), hipMemcpyDeviceToDevice));
    TS_HIP(hipMemcpy(n.sblk, s->sblk, (P / 8) * sizeof(float4), hipMemcpyDeviceToDevice));
    TS_HIP(hipMemcpy(n.tiny, s->tiny, C * sizeof(uint8_t), hipMemcpyDeviceToDevice));
  }
  (void)hipFree(s->x); (void)hipFree(s->y); (void)hipFree(s->off); (void)hipFree(s->len);
  (void)hipFree(s->mbr); (void)hipFree(s->ends); (void)hipFree(s->blk); (void)hipFree(s->sblk);
  (void)hipFree(s->tiny);
  s->x = n.x; s->y = n.y; s->off = n.off; s->len = n.len;
  s->mbr = n.mbr; s->ends = n.ends; s->blk = n.blk; s->sblk = n.sblk; s->tiny = n.tiny;
  n.x = n.y = nullptr; n.off = nullptr; n.len = nullptr; n.mbr = n.ends = n.blk = n.sblk = nullptr;
  n.tiny = nullptr;
  s->bytes = n.bytes;
  s->cap_points = points_cap;
  s->cap_count = count_cap;
  return TS_OK;
}

// Points of tile t (blockIdx.x): slot l of row v takes point min(v, len-1)
// of its trajectory (the last one for v past its end), -1 slots point 0 of
// trajectory 0 (never read as a result).
__global__ void __launch_bounds__(256) k_tile_points(SetView s, const int64_t* toff, const int32_t* tlen,
                                                     const int32_t* tid, float* tx, float* ty) {
  const int t = blockIdx.x;
  const int lane = threadIdx.x & (kTile - 1);
  const int id = tid[(int64_t)t * kTile + lane];
  const int m = id >= 0 ? s.len[id] : 1;
  const int64_t src = id >= 0 ? s.off[id] : 0;
  const int64_t base = toff[t];
  const int L = tlen[t];
  for (int v = threadIdx.x / kTile; v < L; v += 256 / kTile) {
    const int64_t o = src + min(v, m - 1);
    tx[base + (int64_t)v * kTile + lane] = s.x[o];
    ty[base + (int64_t)v * kTile + lane] = s.y[o];
  }
}

void free_tiles(const ts_set* s) {
  ts_set::Tiles& t = s->tiles;
  (void)hipFree(t.x); (void)hipFree(t.y); (void)hipFree(t.off); (void)hipFree(t.len); (void)hipFree(t.id);
  t = ts_set::Tiles{};
}

ts_status set_tiles(const ts_set* s, hipStream_t st, TileView* out) {
  std::lock_guard<std::mutex> lk(s->tiles_mu);
  ts_set::Tiles& t = s->tiles;
  if (t.gen != s->gen) {
    // (length, index) order by a counting sort over the lengths
    const int32_t C = s->count;
    const int32_t maxl = s->max_len;
    std::vector<int64_t> start((size_t)maxl + 2, 0);
    for (int32_t i = 0; i < C; ++i) ++start[s->h_len[i] + 1];
    for (int32_t l = 1; l <= maxl + 1; ++l) start[l] += start[l - 1];
    const int32_t nt = (C + kTile - 1) / kTile;
    std::vector<int32_t> ids((size_t)nt * kTile, -1), tlen(nt, 1);
    for (int32_t i = 0; i < C; ++i) ids[start[s->h_len[i]]++] = i;
    std::vector<int64_t> toff(nt);
    int64_t elems = 0;
    for (int32_t k = 0; k < nt; ++k) {
      const int32_t last = ids[(size_t)std::min<int64_t>((int64_t)k * kTile + kTile, C) - 1];
      tlen[k] = s->h_len[last];
      toff[k] = elems;
      elems += (int64_t)tlen[k] * kTile;
    }
    if (elems > t.cap_elems || nt > t.cap_tiles) {
      free_tiles(s);
      const int64_t ce = std::max<int64_t>(elems, 1);
      const int32_t ct = std::max(nt, 1);
      hipError_t e = hipMalloc((void**)&t.x, ce * sizeof(float));
      if (e == hipSuccess) e = hipMalloc((void**)&t.y, ce * sizeof(float));
      if (e == hipSuccess) e = hipMalloc((void**)&t.off, ct * sizeof(int64_t));
      if (e == hipSuccess) e = hipMalloc((void**)&t.len, ct * sizeof(int32_t));
      if (e == hipSuccess) e = hipMalloc((void**)&t.id, (size_t)ct * kTile * sizeof(int32_t));
      if (e != hipSuccess) {
        (void)hipGetLastError();
        free_tiles(s);
        return fail(TS_ERR_OOM, "tile view allocation of " + std::to_string(2 * ce * 4) + " bytes failed");
      }
      t.cap_elems = ce;
      t.cap_tiles = ct;
      t.bytes = 2 * ce * 4 + (int64_t)ct * (8 + 4 + 4 * kTile);
    }
    t.ntiles = nt;
    TS_HIP(hipMemcpyAsync(t.off, toff.data(), nt * sizeof(int64_t), hipMemcpyHostToDevice, st));
    TS_HIP(hipMemcpyAsync(t.len, tlen.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice, st));
    TS_HIP(hipMemcpyAsync(t.id, ids.data(), (size_t)nt * kTile * sizeof(int32_t), hipMemcpyHostToDevice, st));
    if (nt > 0) {
      hipLaunchKernelGGL(k_tile_points, dim3(nt), dim3(256), 0, st, s->view(), t.off, t.len, t.id, t.x, t.y);
      TS_HIP(hipGetLastError());
    }
    TS_HIP(hipStreamSynchronize(st));
    t.gen = s->gen;
  }
  *out = TileView{t.x, t.y, t.off, t.len, t.id, t.ntiles};
  return TS_OK;
}

ts_status set_load(ts_set* s, const ts_point* xy, const int64_t* offsets, int32_t count, hipStream_t st) {
  std::vector<int32_t> lens;
  std::vector<int64_t> doff;
  std::vector<float> hx, hy;
  TS_TRY(pad_layout(xy, offsets, count, 0, lens, doff, hx, hy));
  const int64_t P = doff[count];
  TS_TRY(set_reserve(s, P, count, false));
  s->count = count;
  s->points = offsets[count];
  s->dev_points = P;
  s->min_len = *std::min_element(lens.begin(), lens.end());
  s->max_len = *std::max_element(lens.begin(), lens.end());
  s->h_off.assign(offsets, offsets + count + 1);
  s->h_len = lens;
  ++s->gen;
  TS_HIP(hipMemcpyAsync(s->x, hx.data(), P * sizeof(float), hipMemcpyHostToDevice, st));
  TS_HIP(hipMemcpyAsync(s->y, hy.data(), P * sizeof(float), hipMemcpyHostToDevice, st));
  TS_HIP(hipMemcpyAsync(s->off, doff.data(), (count + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  TS_HIP(hipMemcpyAsync(s->len, s->h_len.data(), count * sizeof(int32_t), hipMemcpyHostToDevice, st));
  TS_HIP(launch_summaries(st, s->view(), s->mbr, s->ends, s->tiny));
  TS_HIP(launch_blocks(st, s->view(), P, s->blk, s->sblk));
  TS_HIP(hipStreamSynchronize(st));
  return TS_OK;
}

// Appends trajectories behind the set's current ones: device arrays grow (by
// half again at least, keeping their contents) when the new ones do not fit;
// summaries are computed for the new trajectories only.
ts_status set_append(ts_set* s, const ts_point* xy, const int64_t* offsets, int32_t count) {
  std::vector<int32_t> lens;
  std::vector<int64_t> doff;
  std::vector<float> hx, hy;
  const int64_t P0 = s->dev_points;
  const int32_t C0 = s->count;
  TS_TRY(pad_layout(xy, offsets, count, P0, lens, doff, hx, hy));
  const int64_t P = doff[count];
  if ((int64_t)C0 + count > INT32_MAX - 1) return fail(TS_ERR_INVALID, "set: more than 2^31 trajectories");
  const int32_t C = C0 + count;
  if (P > s->cap_points || C > s->cap_count)
    TS_TRY(set_reserve(s, std::max<int64_t>(P, s->cap_points + s->cap_points / 2),
                       (int32_t)std::min<int64_t>(INT32_MAX - 1, std::max<int64_t>(C, (int64_t)s->cap_count * 3 / 2)),
                       true));
  TS_HIP(hipMemcpy(s->x + P0, hx.data(), (P - P0) * sizeof(float), hipMemcpyHostToDevice));
  TS_HIP(hipMemcpy(s->y + P0, hy.data(), (P - P0) * sizeof(float), hipMemcpyHostToDevice));
  TS_HIP(hipMemcpy(s->off + C0 + 1, doff.data() + 1, count * sizeof(int64_t), hipMemcpyHostToDevice));
  TS_HIP(hipMemcpy(s->len + C0, lens.data(), count * sizeof(int32_t), hipMemcpyHostToDevice));
  const int64_t pts0 = s->points;
  s->h_off.reserve(C + 1);
  for (int32_t i = 1; i <= count; ++i) s->h_off.push_back(pts0 + offsets[i]);
  s->h_len.insert(s->h_len.end(), lens.begin(), lens.end());
  s->min_len = std::min(C0 ? s->min_len : INT32_MAX, *std::min_element(lens.begin(), lens.end()));
  s->max_len = std::max(C0 ? s->max_len : 0, *std::max_element(lens.begin(), lens.end()));
  s->count = C;
  s->points = pts0 + offsets[count];
  s->dev_points = P;
  ++s->gen;
  TS_HIP(launch_summaries(nullptr, s->view(), s->mbr, s->ends, s->tiny, C0));
  TS_HIP(launch_blocks(nullptr, s->view(), P, s->blk, s->sblk, C0, P0));
  TS_HIP(hipStreamSynchronize(nullptr));
  return TS_OK;
}

}  // namespace ts

using namespace ts;

extern "C" {

const char* ts_status_string(ts_status s) {
  switch (s) {
    case TS_OK: return "ok";
    case TS_ERR_INVALID: return "invalid argument";
    case TS_ERR_HIP: return "HIP runtime error";
    case TS_ERR_OOM: return "out of device memory";
    case TS_ERR_UNSUPPORTED: return "unsupported";
    case TS_ERR_NO_DEVICE: return "no device";
    case TS_ERR_BUSY: return "busy";
    case TS_ERR_TIMEOUT: return "timeout";
    case TS_ERR_CANCELLED: return "cancelled";
  }
  return "unknown status";
}

const char* ts_last_error(void) { return g_err.c_str(); }

void ts_params_default(ts_params* p, int32_t measure, int32_t mode) {
  if (!p) return;
  p->measure = measure;
  p->mode = mode;
  p->k = 10;
  p->kc = measure == TS_EUCLIDEAN ? 50 : 10;  // euclidean.cpp:90, dtw.cpp:36, hausdorff.cuh:9
  p->r = 5;                                    // dtw.cpp:17
  p->band = 0;
  p->delta = -1;
  p->epsilon = 0.5f;                           // main.cu:170
}

ts_status ts_init(int32_t device) {
  std::lock_guard<std::mutex> lk(g_init_mu);
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) return fail(TS_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= n) return fail(TS_ERR_INVALID, "device index out of range");
  TS_HIP(hipSetDevice(device));
  TS_HIP(hipFree(nullptr));
  // the allocation-free entry point's scratch array on this device
  TS_TRY(noalloc_setup(device));
  // the multi-wave DP's LDS limit (two halo rows of the longest trajectories)
  TS_TRY(dp_mw_setup(device));
  g_device = device;
  return TS_OK;
}

ts_status ts_device_name(char* buf, int32_t buflen) {
  if (!buf || buflen <= 0) return fail(TS_ERR_INVALID, "buffer");
  int dev = 0;
  TS_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  TS_HIP(hipGetDeviceProperties(&p, dev));
  std::snprintf(buf, buflen, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
  return TS_OK;
}

ts_status ts_set_create(const ts_point* xy, const int64_t* offsets, int32_t count, ts_set** out) {
  if (!out) return fail(TS_ERR_INVALID, "ts_set_create: null out");
  *out = nullptr;
  auto s = new ts_set();
  (void)hipGetDevice(&s->device);
  ts_status st = set_load(s, xy, offsets, count, nullptr);
  if (st != TS_OK) {
    ts_set_destroy(s);
    return st;
  }
  *out = s;
  return TS_OK;
}

ts_status ts_set_reload(ts_set* s, const ts_point* xy, const int64_t* offsets, int32_t count) {
  if (!s) return fail(TS_ERR_INVALID, "ts_set_reload: null set");
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (dev != s->device) return fail(TS_ERR_INVALID, "ts_set_reload: set lives on another device");
  return set_load(s, xy, offsets, count, nullptr);
}

ts_status ts_set_append(ts_set* s, const ts_point* xy, const int64_t* offsets, int32_t count) {
  if (!s) return fail(TS_ERR_INVALID, "ts_set_append: null set");
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (dev != s->device) return fail(TS_ERR_INVALID, "ts_set_append: set lives on another device");
  return set_append(s, xy, offsets, count);
}

ts_status ts_set_destroy(ts_set* s) {
  if (!s) return TS_OK;
  (void)hipFree(s->x); (void)hipFree(s->y); (void)hipFree(s->off); (void)hipFree(s->len);
  (void)hipFree(s->mbr); (void)hipFree(s->ends); (void)hipFree(s->blk); (void)hipFree(s->sblk);
  (void)hipFree(s->tiny);
  free_tiles(s);
  delete s;
  return TS_OK;
}

ts_status ts_set_flags(const ts_set* s, int32_t* flags) {
  if (!s || !flags) return fail(TS_ERR_INVALID, "ts_set_flags: null argument");
  *flags = 0;
  if (s->count <= 0) return TS_OK;
  std::vector<uint8_t> t(s->count);
  TS_HIP(hipDeviceSynchronize());
  TS_HIP(hipMemcpy(t.data(), s->tiny, s->count * sizeof(uint8_t), hipMemcpyDeviceToHost));
  for (const uint8_t f : t) *flags |= f;
  return TS_OK;
}

ts_status ts_set_info(const ts_set* s, int32_t* count, int64_t* points, int32_t* min_len, int32_t* max_len) {
  if (!s) return fail(TS_ERR_INVALID, "null set");
  if (count) *count = s->count;
  if (points) *points = s->points;
  if (min_len) *min_len = s->min_len;
  if (max_len) *max_len = s->max_len;
  return TS_OK;
}

ts_status ts_set_bytes(const ts_set* s, int64_t* bytes) {
  if (!s || !bytes) return fail(TS_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->tiles_mu);
  *bytes = s->bytes + s->tiles.bytes;
  return TS_OK;
}

ts_status ts_gen_lengths(uint64_t seed, int32_t count, int32_t lo, int32_t hi, int32_t log_uniform,
                         int32_t* out_len) {
  if (!out_len || count < 0 || lo < 1 || hi < lo) return fail(TS_ERR_INVALID, "ts_gen_lengths");
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < count; ++i) {
    double u = u01(ctr_hash(seed, 0xFFFFFFFFull, (uint64_t)i));
    int64_t l;
    if (log_uniform) {
      double a = std::log((double)lo), b = std::log((double)hi + 1.0);
      l = (int64_t)std::floor(std::exp(a + u * (b - a)));
    } else {
      l = lo + (int64_t)std::floor(u * (double)(hi - lo + 1));
    }
    out_len[i] = (int32_t)std::min<int64_t>(hi, std::max<int64_t>(lo, l));
  }
  return TS_OK;
}

ts_status ts_gen_walks(uint64_t seed, int32_t count, const int64_t* offsets, double sigma, ts_point* out_xy) {
  if (!offsets || !out_xy || count < 0) return fail(TS_ERR_INVALID, "ts_gen_walks");
#pragma omp parallel for schedule(dynamic, 64)
  for (int32_t t = 0; t < count; ++t) {
    const int64_t o = offsets[t], n = offsets[t + 1] - offsets[t];
    double x = 0, y = 0;
    for (int64_t i = 0; i < n; ++i) {
      uint64_t h1 = ctr_hash(seed, (uint64_t)t, (uint64_t)i * 2);
      uint64_t h2 = ctr_hash(seed, (uint64_t)t, (uint64_t)i * 2 + 1);
      if (i == 0) {
        x = u01(h1);
        y = u01(h2);
      } else {
        // Box-Muller: two independent N(0,1) from two uniforms.
        double u1 = 1.0 - u01(h1), u2 = u01(h2);
        double rad = std::sqrt(-2.0 * std::log(u1));
        x += sigma * rad * std::cos(6.283185307179586 * u2);
        y += sigma * rad * std::sin(6.283185307179586 * u2);
      }
      out_xy[o + i].x = (float)x;
      out_xy[o + i].y = (float)y;
    }
  }
  return TS_OK;
}

}  // extern "C"
