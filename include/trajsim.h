/* trajsim.h -- C ABI of the MI355X trajectory k-NN engine (libtrajsim.so).
 *
 * One library replaces the reference's per-measure GPU entry points
 * (1li2yang3/CudaRuntime1: dtw.cuh:6, frechet.cuh:6, lcss.cuh:6-7,
 * hausdorff.cuh:11-19, euclidean.cuh:10-13). The reference functions each
 * upload two fixed-length trajectory batches, filter, pick candidates on the
 * host, refine on the GPU and return the best value per query. Here the
 * database and the query batch are uploaded once into device-resident sets
 * and a k-NN call runs filter -> candidate selection -> exact refine -> top-k
 * on the device. The 12 reference signatures themselves are re-exported with
 * C++ linkage by trajsim_compat.hpp.
 *
 * Conventions: every call returns ts_status (TS_OK == 0); a failing call
 * leaves a message for ts_last_error() (thread-local). Callers own all host
 * buffers; the library owns device memory behind opaque handles. One engine
 * per host thread. All calls are synchronous unless noted.
 */
#ifndef TRAJSIM_H
#define TRAJSIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TS_API_VERSION 1

/* 2-D point, layout-identical to the reference's Point (common.cuh:9-11). */
typedef struct ts_point { float x, y; } ts_point;

typedef enum ts_status {
  TS_OK = 0,
  TS_ERR_INVALID = 1,     /* bad argument (null pointer, size, unknown measure...) */
  TS_ERR_HIP = 2,         /* a HIP runtime call failed */
  TS_ERR_OOM = 3,         /* device memory exhausted */
  TS_ERR_UNSUPPORTED = 4, /* combination not implemented (e.g. band with Frechet) */
  TS_ERR_NO_DEVICE = 5,   /* no usable GPU */
  TS_ERR_BUSY = 6,        /* serving: admission refused, queue full */
  TS_ERR_TIMEOUT = 7,     /* serving: result not ready before the deadline */
  TS_ERR_CANCELLED = 8    /* serving: request cancelled or server shut down */
} ts_status;

/* Measures; the numbering is the library's own. */
typedef enum ts_measure {
  TS_DTW = 0,       /* dtw.cpp:64-105: sum of Euclidean point costs, min-path */
  TS_LCSS = 1,      /* lcss.cpp:87-105: per-axis strict eps matches / min(n,m) */
  TS_FRECHET = 2,   /* frechet.cpp:83-112: discrete Frechet (Eiter-Mannila) */
  TS_HAUSDORFF = 3, /* hausdorff.cpp:172-182: symmetric Hausdorff */
  TS_EUCLIDEAN = 4  /* euclidean.cpp:27-40: lock-step, equal lengths, f64 sum */
} ts_measure;

typedef enum ts_mode {
  /* The reference pipeline: kc candidates chosen by the reference filter
   * (per-point envelope score for DTW/Frechet/LCSS, nearest MBRs for
   * Hausdorff/Euclidean), refined exactly; the best k of them are returned.
   * out_val[q*k] equals the reference's h_results[q]. */
  TS_MODE_REF = 0,
  /* True top-k over the whole database, pruned by admissible bounding-box /
   * endpoint bounds. */
  TS_MODE_EXACT = 1
} ts_mode;

typedef struct ts_params {
  int32_t measure; /* ts_measure */
  int32_t mode;    /* ts_mode */
  int32_t k;       /* results per query (REF: k <= kc) */
  int32_t kc;      /* REF candidates per query (reference: 10; Euclidean 50) */
  int32_t r;       /* REF envelope half-width (reference: 5) */
  int32_t band;    /* DTW Sakoe-Chiba half-width: w = band when > 0, w =
                      ceil(-band * max(n,m) / 100) per pair when in
                      [-100, -1] (a percentage), 0 = unconstrained. Cell
                      (i,j) is inside iff |i*m - j*n| <= w * max(n,m). */
  int32_t delta;   /* LCSS index window, < 0 = none (the reference) */
  float epsilon;   /* LCSS per-axis match threshold */
} ts_params;

/* Fill p with the reference's defaults for a measure (k=kc=10, r=5, no band,
 * no window, eps=0.5; Euclidean kc=50). */
void ts_params_default(ts_params* p, int32_t measure, int32_t mode);

typedef struct ts_stats {
  double ms_total;        /* host wall time of the call */
  double ms_device;       /* GPU time between the call's first and last kernel */
  int64_t pairs;          /* query x database pairs covered (Q*N) */
  int64_t refined_pairs;  /* pairs that went through the exact refine */
  int64_t cells;          /* DP cells (DTW/Frechet/LCSS) or point pairs refined */
  int64_t launches;       /* kernel launches issued */
  int64_t overflow_retries;
} ts_stats;

const char* ts_status_string(ts_status s);
const char* ts_last_error(void);

/* Select the device (default 0) and create the runtime context. */
ts_status ts_init(int32_t device);
ts_status ts_device_name(char* buf, int32_t buflen);

/* ----------------------------------------------------------------- sets
 * A device-resident set of trajectories: the database or a query batch.
 * xy holds offsets[count] points, trajectory i is points
 * [offsets[i], offsets[i+1]); offsets[0] == 0; every length >= 1. The points
 * are stored SoA (x[], y[]) with per-trajectory MBR, endpoints and length. */
typedef struct ts_set ts_set;
ts_status ts_set_create(const ts_point* xy, const int64_t* offsets, int32_t count, ts_set** out);
ts_status ts_set_destroy(ts_set* s);
/* Replace a set's trajectories in place, keeping its device buffers when the
 * new data fits (they grow otherwise). No call that reads the set may be in
 * flight. */
ts_status ts_set_reload(ts_set* s, const ts_point* xy, const int64_t* offsets, int32_t count);
ts_status ts_set_info(const ts_set* s, int32_t* count, int64_t* points, int32_t* min_len, int32_t* max_len);
/* Device memory held by the set, in bytes. */
ts_status ts_set_bytes(const ts_set* s, int64_t* bytes);
/* Lengths of the set's trajectories into out_len[count]. */
ts_status ts_set_lengths(const ts_set* s, int32_t* out_len);
/* Properties found in the points when they were loaded (no counterpart in the
 * reference). Bit TS_SET_TINY_COORD: some trajectory has a coordinate c with
 * 0 < |c| < 2^-25. Bit TS_SET_HUGE_COORD: some trajectory has a coordinate
 * with |c| >= 2^62, or one that is not finite. Between two trajectories
 * without either, DTW's squared point distances are 0 or in [2^-96, 2^127),
 * and the DP kernels skip their test for other values; the kernels keep
 * both flags per trajectory. Waits for the set's pending uploads. */
#define TS_SET_TINY_COORD 1
#define TS_SET_HUGE_COORD 2
ts_status ts_set_flags(const ts_set* s, int32_t* flags);
/* Append trajectories (offsets[0] == 0, as for ts_set_create) behind the
 * set's current ones; they get the next indices. Device arrays grow when
 * needed, keeping their contents; summaries are computed for the new
 * trajectories only. No call that reads the set may be in flight. (No
 * counterpart in the reference, which keeps no database between calls.) */
ts_status ts_set_append(ts_set* s, const ts_point* xy, const int64_t* offsets, int32_t count);

/* -------------------------------------------------------- database files
 * A tsdb file holds a set in the device layout (lengths, padded SoA points,
 * MBRs, endpoints, a length index; layout in csrc/dbfile.hip), so loading is
 * a copy of sections into HBM. No counterpart in the reference, which keeps
 * no database between calls (main.cu:68-74).
 *   ts_set_save   writes a device-resident set;
 *   ts_db_write   writes the same file from host arrays, without a GPU;
 *   ts_set_load   maps a file, checks its header, lengths and length index,
 *                 copies the points to the current device, recomputes the
 *                 summaries there and rejects the file (TS_ERR_INVALID) when
 *                 its stored MBRs or endpoints disagree with its points. */
ts_status ts_set_save(const ts_set* s, const char* path);
ts_status ts_db_write(const char* path, const ts_point* xy, const int64_t* offsets, int32_t count);
ts_status ts_set_load(const char* path, ts_set** out);

/* --------------------------------------------------------------- engine */
typedef struct ts_engine ts_engine;
ts_status ts_engine_create(ts_engine** out);
ts_status ts_engine_destroy(ts_engine* e);
/* Record a HIP event pair around every kernel launch (per-kernel times for
 * ts_engine_kernel_stats); off by default. */
ts_status ts_engine_set_timing(ts_engine* e, int32_t on);
/* Replay batches of ts_knn from HIP graphs (off by default; not used while
 * timing is on, nor by ts_knn_plain). A batch's launch sequence is recorded
 * the second time the same shape is seen (same set handles, lengths range,
 * parameters, batch range and workspace) and replayed from then on; the data
 * behind the handles may change between calls (ts_set reload in place).
 * Recorded EXACT batches size their candidate lists on the device within a
 * capacity; a call whose candidates exceed it grows the capacity and runs
 * again (ts_stats.overflow_retries). Results are identical either way.
 * With on == 1, batches of more than 2^26 (query, DB) pairs launch eagerly
 * (there launch overhead is a negligible share); on == 2 records and
 * replays batches of every size. */
ts_status ts_engine_set_graphs(ts_engine* e, int32_t on);
/* Graphs recorded and replayed so far, and graphs currently cached. */
ts_status ts_engine_graph_info(ts_engine* e, int64_t* captures, int64_t* replays, int64_t* cached);

/* DP launch shapes. The wavefront DP runs a (query, DB) pair on a group of
 * `lanes` lanes (16, 32 or 64: a quarter, half or whole wave; 128 or 256:
 * the 2 or 4 waves of a block, the rows between their strips passed in LDS)
 * that updates `cells` DP cells per lane step; pairs are split into 6
 * buckets by the row width the DP visits (shorter length, or band width):
 * <= 31, 63, 127, 255, 1023, larger. Variants: 0 DTW, 1 DTW with a band,
 * 2 Frechet, 3 LCSS, 4 LCSS with a window. Each (variant, bucket) has a
 * shape and a persistent grid of waves_per_cu waves per CU, in four tables
 * chosen by how many pairs a refine can hold: table 4 up to 256 (at most a
 * pair per CU), table 1 up to 4096 (one pair's latency decides), table 3 up
 * to 65536 (a few rounds of the grid), table 0 above (throughput). Engines
 * start from built-in defaults, then read dp_tuning.txt beside the library
 * (or the file TS_DP_TUNING names), written by the autotuner
 * (cudaruntime1_amd/tune.py): lines "<variant> <bucket> <lanes> <cells>
 * <waves_per_cu> [lat|mid|few]" with variant one of dtw, dtw_band, frechet,
 * lcss, lcss_window ("lat" lines fill table 1, "mid" lines table 3, "few"
 * lines table 4, others table 0 and the tables no line of theirs filled
 * before). ts_engine_set_dp_shape sets table 0, 1, 3 or 4, or all of them
 * (2); ts_engine_get_dp_shape reads table 0, 1, 3 or 4. Results never
 * depend on the shape. */
/* DP cells of an n x m pair under p's measure, band and window, as the
 * engine's cell counters and the bench count them: n*m, the in-band cells
 * of a DTW band (w = band, or ceil(-band * max(n,m) / 100) for a negative
 * band: cells with |r*C - j*R| <= w*R, R = max(n,m), C = min(n,m)), or the
 * cells within an LCSS window; n for Euclidean. Host only (no GPU); -1 for
 * a null p or a length below 1. No counterpart in the reference (whose
 * cores visit every cell). */
int64_t ts_dp_cells(const ts_params* p, int32_t n, int32_t m);
int32_t ts_dp_shape_count(void);
ts_status ts_dp_shape(int32_t i, int32_t* lanes, int32_t* cells);
ts_status ts_engine_load_tuning(ts_engine* e, const char* path);
ts_status ts_engine_set_dp_shape(ts_engine* e, int32_t table, int32_t variant, int32_t bucket, int32_t lanes,
                                 int32_t cells, int32_t waves_per_cu);
ts_status ts_engine_get_dp_shape(ts_engine* e, int32_t table, int32_t variant, int32_t bucket, int32_t* lanes,
                                 int32_t* cells, int32_t* waves_per_cu);
/* DP launches so far whose shape's LDS halo did not fit a block (a long band
 * on a 16- or 32-lane group) and that ran the bucket's default shape. */
ts_status ts_engine_dp_fallbacks(ts_engine* e, int64_t* n);
/* Per-kernel totals since the last reset: launches, summed kernel ms and
 * summed algorithmic bytes / work units that the launcher attributes to it.
 * idx enumerates kernels; returns TS_ERR_INVALID past the last one. */
ts_status ts_engine_kernel_stats(ts_engine* e, int32_t idx, char* name, int32_t namelen,
                                 int64_t* launches, double* ms, double* bytes, double* units);
ts_status ts_engine_reset_stats(ts_engine* e);

/* Batched k-NN: every query of the batch in one launch sequence (filter,
 * candidate selection, persistent refine, top-k). out_idx/out_val are host
 * arrays of n_queries*k; slots without a result hold idx -1. */
ts_status ts_knn(ts_engine* e, const ts_set* db, const ts_set* queries, const ts_params* p,
                 int32_t* out_idx, float* out_val, ts_stats* stats);

/* Batched k-NN where query i asks for its own k: k_per_query[i] in [1, p->k].
 * out_idx/out_val are n_queries * p->k; query i's first k_per_query[i]
 * entries equal what ts_knn with k = k_per_query[i] returns, its later slots
 * hold idx -1 (EXACT: each query's pruning threshold comes from its own k;
 * REF: the best of its p->kc candidates). No counterpart in the reference;
 * the serving loop batches requests that differ only in k through it. */
ts_status ts_knn_multi_k(ts_engine* e, const ts_set* db, const ts_set* queries, const ts_params* p,
                         const int32_t* k_per_query, int32_t* out_idx, float* out_val, ts_stats* stats);

/* Plain k-NN: the same result, one launch sequence per query (the drop-in
 * execution model of the reference entry points, with the database resident). */
ts_status ts_knn_plain(ts_engine* e, const ts_set* db, const ts_set* queries, const ts_params* p,
                       int32_t* out_idx, float* out_val, ts_stats* stats);

/* Building blocks, exposed for parity tests and the compat layer. */
/* REF filter score of every (query, db) pair into host scores[Q*N]
 * (DTW: LB sum dtw.cpp:45-58; Frechet: LB max frechet.cpp:46-62; LCSS: count
 * ratio lcss.cpp:46-63; Hausdorff/Euclidean: squared MBR gap). */
ts_status ts_filter_ref(ts_engine* e, const ts_set* db, const ts_set* queries, const ts_params* p,
                        float* scores);
/* Exact value of npairs given (query, db) index pairs into host out[npairs]. */
ts_status ts_pair_values(ts_engine* e, const ts_set* db, const ts_set* queries, const ts_params* p,
                         const int32_t* qidx, const int32_t* didx, int64_t npairs, float* out);
/* ts_pair_values without any device allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (serving loops
 * that keep a fixed device footprint, harnesses that forbid hipMalloc after
 * start-up). Same arguments and results as ts_pair_values, bit for bit;
 * at most TS_NOALLOC_MAX_PAIRS pairs per call (more: TS_ERR_INVALID). Every
 * device buffer of the call lies in a static scratch array of about 2 MB
 * that the library reserves on each GPU at ts_init (or at the first call on
 * a device that is current then) and holds until the process ends. Calls on
 * one GPU share that array; overlapping calls would race on it, so the
 * library runs them one at a time per GPU. Synchronous. */
#define TS_NOALLOC_MAX_PAIRS 16384
ts_status ts_pair_values_noalloc(ts_engine* e, const ts_set* db, const ts_set* queries, const ts_params* p,
                                 const int32_t* qidx, const int32_t* didx, int64_t npairs, float* out);
/* Per-row top-k of a host score matrix [rows][cols] by (value, column)
 * ascending, or value descending for descending != 0; used for parity tests
 * of the device selection. */
ts_status ts_topk_rows(ts_engine* e, const float* scores, int32_t rows, int32_t cols, int32_t k,
                       int32_t descending, int32_t* out_idx, float* out_val);

/* --------------------------------------------------------------- serving
 * Continuous batching over a resident database: requests of any measure,
 * length and k are admitted while device memory allows, batched as they
 * arrive, and run concurrently on several streams. */
typedef struct ts_server ts_server;
typedef struct ts_server_config {
  int32_t streams;          /* concurrent in-flight batches (default 4) */
  int32_t max_batch;        /* queries per batch (default 256) */
  int32_t max_queue;        /* pending requests before TS_ERR_BUSY (default 65536) */
  int64_t mem_reserve;      /* bytes of device memory never handed out */
  int64_t fake_free_mem;    /* >0: pretend this much device memory is free (tests) */
  int32_t batch_wait_us;    /* how long the scheduler waits to grow a batch */
} ts_server_config;

void ts_server_config_default(ts_server_config* c);
ts_status ts_server_create(const ts_set* db, const ts_server_config* c, ts_server** out);
/* Queue one query; the server copies q. ticket identifies the request. */
ts_status ts_server_submit(ts_server* s, const ts_point* q, int32_t len, const ts_params* p, uint64_t* ticket);
/* Wait up to timeout_ms (<0: forever) for a result; out arrays hold p->k. */
ts_status ts_server_wait(ts_server* s, uint64_t ticket, int32_t timeout_ms, int32_t* out_idx,
                         float* out_val, double* latency_ms);
ts_status ts_server_cancel(ts_server* s, uint64_t ticket);
/* Counters: [0] admitted, [1] refused, [2] completed, [3] batches, [4] cancelled. */
ts_status ts_server_counters(ts_server* s, int64_t* out5);
ts_status ts_server_destroy(ts_server* s);

/* --------------------------------------------------------------- data
 * Synthetic random walks (CPU): start U[0,1)^2, Gaussian steps of sigma,
 * accumulated in double and stored as float; a counter-based generator keyed
 * by (seed, trajectory, point) makes the bytes identical on any host. */
ts_status ts_gen_lengths(uint64_t seed, int32_t count, int32_t lo, int32_t hi, int32_t log_uniform,
                         int32_t* out_len);
ts_status ts_gen_walks(uint64_t seed, int32_t count, const int64_t* offsets, double sigma, ts_point* out_xy);

#ifdef __cplusplus
}
#endif
#endif /* TRAJSIM_H */
