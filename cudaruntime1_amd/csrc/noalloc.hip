// noalloc.hip -- exact values of (query, DB) pairs with no device allocation.
//
// ts_pair_values_noalloc is ts_pair_values (engine.hip) for callers that
// cannot call hipMalloc: serving loops that must keep a fixed device
// footprint, harnesses that forbid allocation after start-up. Every device
// buffer of the call comes from `scratch`, one file-scope device array per
// GPU (kScratchBytes, about 2 MB), which the library holds on each GPU it
// runs on until the process ends:
//   the pair indices as uploaded, the refine's work items and values, the
//   DP's per-bucket work lists, its cell / pair counters and its work
//   counters.
// The call copies the caller's index arrays in, refines on the engine's
// stream through the same kernels as ts_pair_values, and copies the values
// out; it allocates no device memory (the engine's own workspace is never
// touched) and waits for its stream before returning.
//
// The array's name is a host-side stand-in in host code: its device address
// comes from hipGetSymbolAddress, once per device with that device current
// (at ts_init, or at the first call on a device set up no other way), and
// only that address is used: in kernel arguments and copies.
//
// All calls on one GPU share that GPU's array, which keeps the last call's
// contents (the counters are cleared by the call's first kernel). Two calls
// that overlapped on one GPU would race on it, so the library serialises
// them with a lock per device.
#include <climits>
#include <mutex>

#include "ts_internal.h"

namespace ts {
namespace {

constexpr int64_t kMaxPairs = TS_NOALLOC_MAX_PAIRS;
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
// Sections of the scratch array, each on a 256-byte boundary.
constexpr size_t kOffIdx = 0;  // int32 qidx[kMaxPairs], then didx[kMaxPairs]
constexpr size_t kOffItems = kOffIdx + align256(2 * kMaxPairs * sizeof(int32_t));         // Item[kMaxPairs]
constexpr size_t kOffVals = kOffItems + align256(kMaxPairs * sizeof(Item));               // float[kMaxPairs]
constexpr size_t kOffLists = kOffVals + align256(kMaxPairs * sizeof(float));              // Item[kDpBuckets][kMaxPairs]
constexpr size_t kOffCells = kOffLists + align256(kDpBuckets * kMaxPairs * sizeof(Item));  // u64[kCellBuf]
constexpr size_t kOffCtr = kOffCells + align256(kCellBuf * sizeof(unsigned long long));   // int32[32]
constexpr size_t kScratchBytes = kOffCtr + align256(32 * sizeof(int32_t));

constexpr int kMaxDevices = 64;
unsigned char* g_scratch[kMaxDevices];  // device address of `scratch` on each device (null: not set up)
std::mutex g_call_mu[kMaxDevices];      // one call at a time per device
std::mutex g_setup_mu;

// Work items (q, d, slot = position) from the uploaded index arrays; the
// first kCellBuf threads clear the cell / pair counters.
__global__ void k_noalloc_items(const int32_t* qd, int32_t n, Item* items, unsigned long long* cells) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) items[i] = Item{qd[i], qd[kMaxPairs + i], (int32_t)i, 0};
  if (i < kCellBuf) cells[i] = 0ull;
}

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

ts_status noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(TS_ERR_INVALID, "device index out of range for scratch");
  std::lock_guard<std::mutex> lk(g_setup_mu);
  if (g_scratch[device]) return TS_OK;
  void* p = nullptr;
  TS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(TS_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
  g_scratch[device] = (unsigned char*)p;
  return TS_OK;
}

}  // namespace ts

using namespace ts;

extern "C" ts_status ts_pair_values_noalloc(ts_engine* e, const ts_set* db, const ts_set* qs, const ts_params* p,
                                            const int32_t* qidx, const int32_t* didx, int64_t npairs, float* out) {
  if (!e || !db || !qs || !p || !qidx || !didx || !out || npairs < 0)
    return fail(TS_ERR_INVALID, "ts_pair_values_noalloc: null argument or negative count");
  if (npairs > kMaxPairs)
    return fail(TS_ERR_INVALID, "ts_pair_values_noalloc: " + std::to_string(npairs) +
                                    " pairs exceed TS_NOALLOC_MAX_PAIRS (" + std::to_string(kMaxPairs) +
                                    "); split the call or use ts_pair_values");
  TS_TRY(validate_params(p, db, qs));
  for (int64_t i = 0; i < npairs; ++i)
    if (qidx[i] < 0 || qidx[i] >= qs->count || didx[i] < 0 || didx[i] >= db->count)
      return fail(TS_ERR_INVALID, "ts_pair_values_noalloc: pair index out of range");
  if (npairs == 0) return TS_OK;
  const int dev = engine_device(e);
  if (dev < 0 || dev >= kMaxDevices) return fail(TS_ERR_INVALID, "engine device out of range for scratch");
  if (!g_scratch[dev]) {
    // set-up on first use of a device ts_init did not set up: needs it current
    int cur = -1;
    TS_HIP(hipGetDevice(&cur));
    if (cur != dev) return fail(TS_ERR_INVALID, "ts_pair_values_noalloc: call ts_init on the engine's device first");
    TS_TRY(noalloc_setup(dev));
  }
  std::lock_guard<std::mutex> lk(g_call_mu[dev]);
  unsigned char* base = g_scratch[dev];
  int32_t* qd = (int32_t*)(base + kOffIdx);
  Item* items = (Item*)(base + kOffItems);
  float* vals = (float*)(base + kOffVals);
  const RefineBufs bufs{(unsigned long long*)(base + kOffCells), (int32_t*)(base + kOffCtr),
                        (Item*)(base + kOffLists), kMaxPairs};
  hipStream_t st = engine_stream(e);
  const size_t nb = (size_t)npairs * sizeof(int32_t);
  TS_HIP(hipMemcpyAsync(qd, qidx, nb, hipMemcpyHostToDevice, st));
  TS_HIP(hipMemcpyAsync(qd + kMaxPairs, didx, nb, hipMemcpyHostToDevice, st));
  const int64_t threads = npairs > kCellBuf ? npairs : kCellBuf;
  hipLaunchKernelGGL(k_noalloc_items, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, qd,
                     (int32_t)npairs, items, bufs.cells);
  TS_HIP(hipGetLastError());
  TS_TRY(refine_fixed(e, *p, db, qs, items, npairs, vals, bufs));
  TS_HIP(hipMemcpyAsync(out, vals, (size_t)npairs * sizeof(float), hipMemcpyDeviceToHost, st));
  TS_HIP(hipStreamSynchronize(st));
  return TS_OK;
}
