// dbfile.hip -- trajectory database files: a set in the device layout, so a
// database is written once and loaded by copying sections into HBM.
//
// The reference keeps no database between calls: each call takes host arrays
// of fixed-length trajectories (main.cu:68-74, dtw.cu:125-138). A tsdb file
// holds a variable-length database, little-endian:
//
//   header  256 bytes: magic "TSDBSOA1", version, header size, count,
//           shortest and longest length, padding unit (8 points), points,
//           padded points, each section's byte offset and size, and an
//           FNV-1a 64 checksum of the header (taken with that field zero)
//   len     int32[count]             trajectory lengths
//   x, y    float32[padded points]   SoA points; trajectory i starts at the
//                                    sum of the earlier lengths rounded up to
//                                    8 and its last point fills its padding
//   mbr     float32x4[count]         (min x, max x, min y, max y)
//   ends    float32x4[count]         (first x, first y, last x, last y)
//   order   int32[count]             trajectory ids by (length, id)
//   class   int64[26]                length index: order[class[c],
//                                    class[c+1]) are the lengths in
//                                    [2^c, 2^(c+1)), c = 0..24
//
// Each section starts on a 256-byte boundary. The loader checks the header,
// the lengths against the sizes and the length index, copies len, x and y
// into the set's device arrays (the points through pinned staging buffers), recomputes every summary there (box kernels
// run in well under a millisecond per million trajectories), and rejects the
// file when the recomputed MBRs and endpoints differ from the stored ones.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "ts_internal.h"

namespace ts {
namespace {

constexpr char kMagic[8] = {'T', 'S', 'D', 'B', 'S', 'O', 'A', '1'};
constexpr uint32_t kVersion = 1;
constexpr int kSections = 7;  // len, x, y, mbr, ends, order, class
constexpr int kClasses = 25;  // lengths 1 .. 2^24
constexpr uint64_t kAlign = 256;

struct DbHeader {
  char magic[8];
  uint32_t version, header_bytes;
  int32_t count, min_len, max_len, pad;
  int64_t points, dev_points;
  uint64_t sec_off[kSections], sec_bytes[kSections];
  uint64_t checksum;
  uint8_t reserved[88];
};
static_assert(sizeof(DbHeader) == 256, "tsdb header is 256 bytes");

uint64_t fnv1a(const void* p, size_t n) {
  const uint8_t* b = (const uint8_t*)p;
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
  return h;
}

uint64_t header_checksum(DbHeader h) {
  h.checksum = 0;
  return fnv1a(&h, sizeof h);
}

// Everything a file holds, on the host.
struct HostDb {
  std::vector<int32_t> lens;
  std::vector<int64_t> doff;
  std::vector<float> x, y;
  std::vector<float4> mbr, ends;
  std::vector<int32_t> order;
  std::vector<int64_t> cls;
  int64_t points = 0;
};

int length_class(int32_t len) {
  int c = 0;
  while (c + 1 < kClasses && ((int64_t)1 << (c + 1)) <= len) ++c;
  return c;
}

// Ids by (length, id) and the start of each length class in that order.
void length_index(const std::vector<int32_t>& lens, std::vector<int32_t>& order, std::vector<int64_t>& cls) {
  const int32_t n = (int32_t)lens.size();
  order.resize(n);
  for (int32_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return lens[a] < lens[b]; });
  cls.assign(kClasses + 1, n);
  int64_t i = 0;
  for (int c = 0; c <= kClasses; ++c) {
    while (i < n && length_class(lens[order[i]]) < c) ++i;
    cls[c] = i;
  }
}

// MBR and endpoints of every trajectory, as k_summaries computes them (min
// and max from the first point over the trajectory's own points).
void host_summaries(HostDb& d) {
  const int32_t n = (int32_t)d.lens.size();
  d.mbr.resize(n);
  d.ends.resize(n);
#pragma omp parallel for schedule(static) if (n > 4096)
  for (int32_t t = 0; t < n; ++t) {
    const int64_t o = d.doff[t];
    const int32_t len = d.lens[t];
    float mnx = d.x[o], mxx = mnx, mny = d.y[o], mxy = mny;
    for (int32_t i = 1; i < len; ++i) {
      mnx = fminf(mnx, d.x[o + i]);
      mxx = fmaxf(mxx, d.x[o + i]);
      mny = fminf(mny, d.y[o + i]);
      mxy = fmaxf(mxy, d.y[o + i]);
    }
    d.mbr[t] = make_float4(mnx, mxx, mny, mxy);
    d.ends[t] = make_float4(d.x[o], d.y[o], d.x[o + len - 1], d.y[o + len - 1]);
  }
}

ts_status write_file(const char* path, const HostDb& d) {
  const int32_t n = (int32_t)d.lens.size();
  DbHeader h;
  std::memset(&h, 0, sizeof h);
  std::memcpy(h.magic, kMagic, 8);
  h.version = kVersion;
  h.header_bytes = sizeof(DbHeader);
  h.count = n;
  h.min_len = *std::min_element(d.lens.begin(), d.lens.end());
  h.max_len = *std::max_element(d.lens.begin(), d.lens.end());
  h.pad = kPad;
  h.points = d.points;
  h.dev_points = d.doff[n];
  const void* data[kSections] = {d.lens.data(), d.x.data(), d.y.data(), d.mbr.data(), d.ends.data(),
                                 d.order.data(), d.cls.data()};
  const uint64_t bytes[kSections] = {(uint64_t)n * 4, (uint64_t)h.dev_points * 4, (uint64_t)h.dev_points * 4,
                                     (uint64_t)n * 16, (uint64_t)n * 16, (uint64_t)n * 4,
                                     (uint64_t)(kClasses + 1) * 8};
  uint64_t at = kAlign;
  for (int i = 0; i < kSections; ++i) {
    h.sec_off[i] = at;
    h.sec_bytes[i] = bytes[i];
    at = (at + bytes[i] + kAlign - 1) / kAlign * kAlign;
  }
  h.checksum = header_checksum(h);
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(TS_ERR_INVALID, std::string("tsdb: cannot create ") + path);
  static const char zeros[kAlign] = {};
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
  uint64_t pos = sizeof h;
  for (int i = 0; i < kSections && ok; ++i) {
    ok = std::fwrite(zeros, 1, h.sec_off[i] - pos, f) == h.sec_off[i] - pos;
    if (ok && bytes[i]) ok = std::fwrite(data[i], 1, bytes[i], f) == bytes[i];
    pos = h.sec_off[i] + bytes[i];
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(TS_ERR_INVALID, std::string("tsdb: write failed: ") + path);
  return TS_OK;
}

// A read-only mapping of a whole file.
struct Mapped {
  void* p = MAP_FAILED;
  size_t n = 0;
  int fd = -1;
  ~Mapped() {
    if (p != MAP_FAILED) munmap(p, n);
    if (fd >= 0) close(fd);
  }
};

// Copies n bytes from pageable memory (the file's mapping) to the device
// through two pinned staging buffers: host threads fill one while the copy
// engine drains the other.
hipError_t upload_staged(void* dst, const void* src, size_t n) {
  constexpr size_t kChunk = (size_t)32 << 20;
  if (n <= kChunk) return hipMemcpy(dst, src, n, hipMemcpyHostToDevice);
  char* pin[2] = {nullptr, nullptr};
  hipEvent_t free_again[2] = {nullptr, nullptr};
  hipStream_t st = nullptr;
  hipError_t e = hipHostMalloc((void**)&pin[0], kChunk, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void**)&pin[1], kChunk, hipHostMallocDefault);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&free_again[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&free_again[1], hipEventDisableTiming);
  for (size_t off = 0, i = 0; e == hipSuccess && off < n; off += kChunk, ++i) {
    const size_t len = std::min(kChunk, n - off);
    char* b = pin[i & 1];
    if (i >= 2) e = hipEventSynchronize(free_again[i & 1]);
    if (e != hipSuccess) break;
    const char* from = (const char*)src + off;
#pragma omp parallel for schedule(static)
    for (int t = 0; t < 16; ++t) {
      const size_t a = len * t / 16, z = len * (t + 1) / 16;
      std::memcpy(b + a, from + a, z - a);
    }
    e = hipMemcpyAsync((char*)dst + off, b, len, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(free_again[i & 1], st);
  }
  if (st) {
    const hipError_t e2 = hipStreamSynchronize(st);
    if (e == hipSuccess) e = e2;
    (void)hipStreamDestroy(st);
  }
  for (hipEvent_t ev : free_again)
    if (ev) (void)hipEventDestroy(ev);
  for (char* b : pin)
    if (b) (void)hipHostFree(b);
  return e;
}

}  // namespace
}  // namespace ts

using namespace ts;

extern "C" {

ts_status ts_db_write(const char* path, const ts_point* xy, const int64_t* offsets, int32_t count) {
  if (!path) return fail(TS_ERR_INVALID, "ts_db_write: null path");
  HostDb d;
  TS_TRY(pad_layout(xy, offsets, count, 0, d.lens, d.doff, d.x, d.y));
  d.points = offsets[count];
  host_summaries(d);
  length_index(d.lens, d.order, d.cls);
  return write_file(path, d);
}

ts_status ts_set_save(const ts_set* s, const char* path) {
  if (!s || !path) return fail(TS_ERR_INVALID, "ts_set_save: null argument");
  HostDb d;
  d.lens = s->h_len;
  d.doff.resize(s->count + 1);
  d.doff[0] = 0;
  for (int32_t i = 0; i < s->count; ++i) d.doff[i + 1] = d.doff[i] + ((d.lens[i] + kPad - 1) / kPad) * kPad;
  d.points = s->points;
  d.x.resize(s->dev_points);
  d.y.resize(s->dev_points);
  d.mbr.resize(s->count);
  d.ends.resize(s->count);
  TS_HIP(hipDeviceSynchronize());
  TS_HIP(hipMemcpy(d.x.data(), s->x, s->dev_points * sizeof(float), hipMemcpyDeviceToHost));
  TS_HIP(hipMemcpy(d.y.data(), s->y, s->dev_points * sizeof(float), hipMemcpyDeviceToHost));
  TS_HIP(hipMemcpy(d.mbr.data(), s->mbr, s->count * sizeof(float4), hipMemcpyDeviceToHost));
  TS_HIP(hipMemcpy(d.ends.data(), s->ends, s->count * sizeof(float4), hipMemcpyDeviceToHost));
  length_index(d.lens, d.order, d.cls);
  return write_file(path, d);
}

ts_status ts_set_load(const char* path, ts_set** out) {
  if (!path || !out) return fail(TS_ERR_INVALID, "ts_set_load: null argument");
  *out = nullptr;
  const std::string where = std::string("tsdb ") + path + ": ";
  Mapped m;
  m.fd = open(path, O_RDONLY);
  if (m.fd < 0) return fail(TS_ERR_INVALID, where + "cannot open");
  struct stat st;
  if (fstat(m.fd, &st) != 0 || st.st_size < (off_t)sizeof(DbHeader)) return fail(TS_ERR_INVALID, where + "too short");
  m.n = (size_t)st.st_size;
  m.p = mmap(nullptr, m.n, PROT_READ, MAP_PRIVATE, m.fd, 0);
  if (m.p == MAP_FAILED) return fail(TS_ERR_INVALID, where + "cannot map");
  const char* base = (const char*)m.p;
  DbHeader h;
  std::memcpy(&h, base, sizeof h);
  if (std::memcmp(h.magic, kMagic, 8) != 0) return fail(TS_ERR_INVALID, where + "not a tsdb file");
  if (h.version != kVersion || h.header_bytes != sizeof(DbHeader))
    return fail(TS_ERR_INVALID, where + "unsupported version " + std::to_string(h.version));
  if (h.checksum != header_checksum(h)) return fail(TS_ERR_INVALID, where + "header checksum mismatch");
  if (h.count < 1 || h.pad != kPad || h.dev_points < h.count || h.dev_points % kPad != 0)
    return fail(TS_ERR_INVALID, where + "bad counts");
  const uint64_t n = (uint64_t)h.count, P = (uint64_t)h.dev_points;
  const uint64_t want[kSections] = {n * 4, P * 4, P * 4, n * 16, n * 16, n * 4, (uint64_t)(kClasses + 1) * 8};
  for (int i = 0; i < kSections; ++i) {
    if (h.sec_bytes[i] != want[i] || h.sec_off[i] % 4 != 0 || h.sec_off[i] < sizeof(DbHeader) ||
        h.sec_off[i] > m.n || h.sec_bytes[i] > m.n - h.sec_off[i])
      return fail(TS_ERR_INVALID, where + "section " + std::to_string(i) + " out of place");
  }
  const int32_t* lens = (const int32_t*)(base + h.sec_off[0]);
  const float* fx = (const float*)(base + h.sec_off[1]);
  const float* fy = (const float*)(base + h.sec_off[2]);
  const float4* fmbr = (const float4*)(base + h.sec_off[3]);
  const float4* fends = (const float4*)(base + h.sec_off[4]);
  const int32_t* order = (const int32_t*)(base + h.sec_off[5]);
  const int64_t* cls = (const int64_t*)(base + h.sec_off[6]);
  // lengths against the sizes, then the length index
  std::vector<int64_t> doff(h.count + 1), hoff(h.count + 1);
  doff[0] = hoff[0] = 0;
  int32_t mn = INT32_MAX, mx = 0;
  for (int32_t i = 0; i < h.count; ++i) {
    const int32_t l = lens[i];
    if (l < 1 || l > (1 << 24)) return fail(TS_ERR_INVALID, where + "trajectory " + std::to_string(i) + " length");
    doff[i + 1] = doff[i] + ((l + kPad - 1) / kPad) * kPad;
    hoff[i + 1] = hoff[i] + l;
    mn = std::min(mn, l);
    mx = std::max(mx, l);
  }
  if (doff[h.count] != h.dev_points || hoff[h.count] != h.points || mn != h.min_len || mx != h.max_len)
    return fail(TS_ERR_INVALID, where + "lengths disagree with the header");
  // the length index in one pass: a permutation in increasing (length, id)
  // whose class starts fall where the classes change
  {
    std::vector<uint8_t> seen(n, 0);
    bool ok = cls[0] == 0 && cls[kClasses] == h.count;
    int c = 0;
    for (int32_t i = 0; ok && i < h.count; ++i) {
      const int32_t id = order[i];
      ok = id >= 0 && id < h.count && !seen[id];
      if (!ok) break;
      seen[id] = 1;
      if (i > 0) {
        const int32_t prev = order[i - 1];
        ok = lens[prev] < lens[id] || (lens[prev] == lens[id] && prev < id);
      }
      const int ci = length_class(lens[id]);
      while (ok && c < ci) ok = cls[++c] == i;  // classes c+1..ci start here
      ok = ok && cls[ci] <= i && i < cls[ci + 1];
    }
    for (int cc = c + 1; ok && cc < kClasses; ++cc) ok = cls[cc] == h.count;
    if (!ok) return fail(TS_ERR_INVALID, where + "length index disagrees with the lengths");
  }
  // device arrays, points, summaries recomputed from the points
  auto s = new ts_set();
  auto bail = [&](ts_status e) {
    ts_set_destroy(s);
    return e;
  };
  (void)hipGetDevice(&s->device);
  ts_status r = set_reserve(s, h.dev_points, h.count, false);
  if (r != TS_OK) return bail(r);
  s->count = h.count;
  s->points = h.points;
  s->dev_points = h.dev_points;
  s->min_len = mn;
  s->max_len = mx;
  s->h_off = hoff;
  s->h_len.assign(lens, lens + h.count);
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = upload_staged(s->x, fx, P * sizeof(float));
  if (e == hipSuccess) e = upload_staged(s->y, fy, P * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(s->off, doff.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->len, s->h_len.data(), n * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_summaries(nullptr, s->view(), s->mbr, s->ends, s->tiny);
  if (e == hipSuccess) e = launch_blocks(nullptr, s->view(), h.dev_points, s->blk, s->sblk);
  std::vector<float4> gm(n), ge(n);
  if (e == hipSuccess) e = hipMemcpy(gm.data(), s->mbr, n * sizeof(float4), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(ge.data(), s->ends, n * sizeof(float4), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return bail(hip_fail(e, "tsdb load", __FILE__, __LINE__));
  if (std::memcmp(gm.data(), fmbr, n * 16) != 0 || std::memcmp(ge.data(), fends, n * 16) != 0)
    return bail(fail(TS_ERR_INVALID, where + "stored MBRs or endpoints disagree with the points"));
  *out = s;
  return TS_OK;
}

ts_status ts_set_lengths(const ts_set* s, int32_t* out_len) {
  if (!s || !out_len) return fail(TS_ERR_INVALID, "ts_set_lengths: null argument");
  std::copy(s->h_len.begin(), s->h_len.end(), out_len);
  return TS_OK;
}

}  // extern "C"
