// oracle.cpp -- CPU restatement of the reference trajectory k-NN path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP engine in
// cudaruntime1_amd/; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it. The product library never links or calls it.
//
// It restates the reference CPU functions of 1li2yang3/CudaRuntime1 from
// their arithmetic, generalised to per-trajectory lengths, Q != N, any k, a
// Sakoe-Chiba band for DTW and an index window delta for LCSS. With the
// reference's shapes (fixed lengths, Q == N) and band/delta off it reduces to
// the reference bit for bit (checked against oracle/_ref, the reference's own
// dtw.cpp/frechet.cpp/lcss.cpp compiled here, and the fixtures in
// tests/golden/).
//
// Build: g++ -O3 -fopenmp -ffp-contract=off (no -march=native): contraction
// would change f32 rounding against the reference's non-contracting build.
//
// Pinning of each function (what checks it besides the reference build):
//   envelope              brute-force windowed min/max (tests/test_oracle.py)
//   lb_dtw/lb_frechet/ub  per-pair closed forms on tiny inputs, _ref values
//   dtw                   brute force over all warping paths (n,m <= 6),
//                         DTW(a,a)=0, n=1 closed form, symmetry, _ref
//   frechet               brute force over couplings, n=1 closed form, _ref
//   lcss                  textbook LCS known answer ("ABCBDAB","BDCABA")=4
//                         on a lattice, symmetry, bounds, _ref
//   hausdorff             O(nm) definition vs numpy, symmetric, H(a,a)=0
//   euclidean             numpy float64 closed form
//   knn EXACT             brute force over the whole DB == pruned search
//   mbr / box distance    numpy closed form; tie order vs Boost's R-tree is
//                         parity unpinned (Boost is not installed here)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <utility>
#include <vector>
#include <omp.h>

namespace {

struct Pt { float x, y; };
struct Env { float minx, maxx, miny, maxy; };

enum Measure { DTW = 0, LCSS = 1, FRECHET = 2, HAUSDORFF = 3, EUCLIDEAN = 4 };
enum Mode { REF = 0, EXACT = 1 };

constexpr float kInf = 1e20f;  // DP sentinel, dtw.cpp:68-69,81

// Envelope of q over the index window [i-r, i+r] clipped to [0, n-1].
// dtw.cpp:20-34 (frechet.cpp:19-33 and lcss.cpp:20-35 are the same; lcss.cpp
// uses fmin/fmax, identical without NaNs).
void envelope(const Pt* q, int n, int r, Env* out) {
  for (int i = 0; i < n; ++i) {
    float mnx = 1e20f, mxx = -1e20f, mny = 1e20f, mxy = -1e20f;
    int a = std::max(0, i - r), b = std::min(n - 1, i + r);
    for (int k = a; k <= b; ++k) {
      mnx = std::min(mnx, q[k].x); mxx = std::max(mxx, q[k].x);
      mny = std::min(mny, q[k].y); mxy = std::max(mxy, q[k].y);
    }
    out[i] = {mnx, mxx, mny, mxy};
  }
}

// Point-to-box distance used by the DTW and Frechet filters: dtw.cpp:54-56.
inline float box_dist(const Env& e, Pt p) {
  float dx = std::max(0.0f, std::max(e.minx - p.x, p.x - e.maxx));
  float dy = std::max(0.0f, std::max(e.miny - p.y, p.y - e.maxy));
  return std::sqrt(dx * dx + dy * dy);
}

// LB_DTW: left-fold f32 sum from 0.0f over the DB points, index map v*n/m.
// dtw.cpp:45-58.
float lb_dtw(const Env* env, int n, const Pt* c, int m) {
  float s = 0.0f;
  for (int v = 0; v < m; ++v) s += box_dist(env[(int64_t)v * n / m], c[v]);
  return s;
}

// LB_Frechet: running max from 0.0f. frechet.cpp:46-62.
float lb_frechet(const Env* env, int n, const Pt* c, int m) {
  float s = 0.0f;
  for (int v = 0; v < m; ++v) s = std::max(s, box_dist(env[(int64_t)v * n / m], c[v]));
  return s;
}

// UB_LCSS count: DB points strictly inside the eps-inflated envelope box.
// lcss.cpp:50-61 (floor((double)v*n/m) equals the integer division for v>=0).
int ub_lcss_count(const Env* env, int n, const Pt* c, int m, float eps) {
  int cnt = 0;
  for (int v = 0; v < m; ++v) {
    const Env& e = env[(int64_t)v * n / m];
    Pt p = c[v];
    if (p.x > e.minx - eps && p.x < e.maxx + eps && p.y > e.miny - eps && p.y < e.maxy + eps) ++cnt;
  }
  return cnt;
}

// Point cost d(p,r) in precision T. dtw.cpp:88-90, frechet.cpp:72-76.
template <class T> inline T pdist(Pt a, Pt b) {
  T dx = (T)a.x - (T)b.x, dy = (T)a.y - (T)b.y;
  return std::sqrt(dx * dx + dy * dy);
}

// Sakoe-Chiba band predicate (build extension, no reference counterpart):
// cell (i,j), 0-based, is inside iff |i*m - j*n| <= w*max(n,m). For n == m it
// is |i-j| <= w; w >= max(n,m) admits every cell; w >= 1 keeps (0,0),
// (n-1,m-1) and a monotone path between them inside. band > 0 is w itself;
// band < 0 asks for w = ceil(-band * max(n,m) / 100), a per-pair percentage.
inline int band_width(int band, int n, int m) {
  return band > 0 ? band : (-band * std::max(n, m) + 99) / 100;
}
inline bool in_band(int i, int j, int n, int m, int w) {
  if (w <= 0) return true;
  int64_t d = (int64_t)i * m - (int64_t)j * n;
  if (d < 0) d = -d;
  return d <= (int64_t)w * std::max(n, m);
}

// DTW with rolling rows, D[0][0]=0, borders 1e20f, Euclidean point cost.
// dtw.cpp:64-100. Cells outside the band are 1e20f.
template <class T> T dtw(const Pt* q, int n, const Pt* c, int m, int band) {
  const T inf = (T)kInf;
  const int w = band == 0 ? 0 : band_width(band, n, m);
  std::vector<T> prev(m + 1, inf), curr(m + 1, inf);
  prev[0] = 0;
  // Row u (1-based) visits columns [vlo, vhi]: the cells in_band admits, a
  // contiguous run whose ends never move left from one row to the next.
  // Every other entry of a row's buffer is inf: before a buffer takes row u
  // its entries from row u-2 left of vlo are reset (those right of vhi are
  // inf already, as vhi never decreases).
  int lo_curr = 1, lo_prev = 1;  // first visited column of the rows the buffers hold
  int vlo = 1, vhi = 0;
  for (int u = 1; u <= n; ++u) {
    if (w > 0) {
      while (vlo <= m && !in_band(u - 1, vlo - 1, n, m, w)) ++vlo;
      if (vhi < vlo - 1) vhi = vlo - 1;
      while (vhi < m && in_band(u - 1, vhi, n, m, w)) ++vhi;
      for (int v = lo_curr; v < vlo && v <= m; ++v) curr[v] = inf;
    } else {
      vhi = m;
    }
    curr[0] = inf;
    for (int v = vlo; v <= vhi; ++v) {
      T cost = pdist<T>(q[u - 1], c[v - 1]);
      curr[v] = cost + std::min({prev[v], curr[v - 1], prev[v - 1]});
    }
    lo_curr = vlo;
    std::swap(prev, curr);
    std::swap(lo_prev, lo_curr);
  }
  return prev[m];
}

// Discrete Frechet (Eiter-Mannila), full table. frechet.cpp:83-112.
template <class T> T frechet(const Pt* q, int n, const Pt* c, int m) {
  std::vector<T> dp((size_t)(n + 1) * (m + 1), (T)0);
  auto at = [&](int u, int v) -> T& { return dp[(size_t)u * (m + 1) + v]; };
  at(1, 1) = pdist<T>(q[0], c[0]);
  for (int u = 2; u <= n; ++u) at(u, 1) = std::max(at(u - 1, 1), pdist<T>(q[u - 1], c[0]));
  for (int v = 2; v <= m; ++v) at(1, v) = std::max(at(1, v - 1), pdist<T>(q[0], c[v - 1]));
  for (int u = 2; u <= n; ++u)
    for (int v = 2; v <= m; ++v) {
      T d = pdist<T>(q[u - 1], c[v - 1]);
      T pm = std::min({at(u - 1, v), at(u, v - 1), at(u - 1, v - 1)});
      at(u, v) = std::max(d, pm);
    }
  return at(n, m);
}

// LCSS match count: per-axis strict epsilon, lcss.cpp:87-103. The index window
// delta (build extension) only restricts which cells may match; delta < 0 is
// no window (the reference).
int lcss(const Pt* q, int n, const Pt* c, int m, float eps, int delta) {
  std::vector<int> prev(m + 1, 0), curr(m + 1, 0);
  for (int u = 1; u <= n; ++u) {
    curr[0] = 0;
    Pt p1 = q[u - 1];
    for (int v = 1; v <= m; ++v) {
      Pt p2 = c[v - 1];
      float dx = std::abs(p1.x - p2.x), dy = std::abs(p1.y - p2.y);
      bool win = delta < 0 || std::abs(u - v) <= delta;
      if (win && dx < eps && dy < eps) curr[v] = prev[v - 1] + 1;
      else curr[v] = std::max(prev[v], curr[v - 1]);
    }
    std::swap(prev, curr);
  }
  return prev[m];
}

// Similarity as the reference reports it: (float)count / min(n,m), lcss.cpp:105.
inline float lcss_sim(int cnt, int n, int m) { return (float)cnt / std::min(n, m); }

// Directed Hausdorff, squared: max over a of min over b of dx*dx+dy*dy in f32.
// hausdorff.cpp:11-109 computes the same value with a grid accelerator; its
// early prune (:83) only skips points that cannot raise the max.
template <class T> T directed_h2(const Pt* a, int na, const Pt* b, int nb) {
  T best = 0;
  for (int i = 0; i < na; ++i) {
    T mn = (T)kInf;
    for (int j = 0; j < nb; ++j) {
      T dx = (T)a[i].x - (T)b[j].x, dy = (T)a[i].y - (T)b[j].y;
      T d2 = dx * dx + dy * dy;
      if (d2 < mn) mn = d2;
    }
    if (mn > best) best = mn;
  }
  return best;
}

// Symmetric Hausdorff distance: sqrt(max(h(q->c), h(c->q))), hausdorff.cpp:172-182.
template <class T> T hausdorff(const Pt* q, int n, const Pt* c, int m) {
  T d1 = directed_h2<T>(q, n, c, m);
  T d2 = directed_h2<T>(c, m, q, n);
  return std::sqrt(std::max(d1, d2));
}

// Lock-step Euclidean in double, sequential sum, euclidean.cpp:27-40 without
// the early abandon (which only returns early once the sum cannot win).
double euclid2(const Pt* q, const Pt* c, int n) {
  double s = 0.0;
  for (int j = 0; j < n; ++j) {
    double dx = (double)q[j].x - (double)c[j].x, dy = (double)q[j].y - (double)c[j].y;
    s += std::pow(dx, 2) + std::pow(dy, 2);
  }
  return s;
}

// Minimum bounding rectangle, hausdorff.cpp:124-134.
Env mbr(const Pt* t, int len) {
  Env e{t[0].x, t[0].x, t[0].y, t[0].y};
  for (int i = 1; i < len; ++i) {
    e.minx = std::min(e.minx, t[i].x); e.maxx = std::max(e.maxx, t[i].x);
    e.miny = std::min(e.miny, t[i].y); e.maxy = std::max(e.maxy, t[i].y);
  }
  return e;
}

// Squared box-box distance (the R-tree's comparable distance between MBRs,
// hausdorff.cpp:162, euclidean.cpp:96), in f32.
inline float box_box_d2(const Env& a, const Env& b) {
  float gx = std::max(0.0f, std::max(b.minx - a.maxx, a.minx - b.maxx));
  float gy = std::max(0.0f, std::max(b.miny - a.maxy, a.miny - b.maxy));
  return gx * gx + gy * gy;
}

struct Params {
  int32_t measure, mode, k, kc, r, band, delta;
  float eps;
};

struct Set {
  const Pt* xy; const int64_t* off; int n;
  const Pt* t(int i) const { return xy + off[i]; }
  int len(int i) const { return (int)(off[i + 1] - off[i]); }
};

// Exact value of one pair in the measure's reported unit (f32).
float exact_value(const Params& p, const Pt* q, int n, const Pt* c, int m) {
  switch (p.measure) {
    case DTW: return dtw<float>(q, n, c, m, p.band);
    case FRECHET: return frechet<float>(q, n, c, m);
    case LCSS: return lcss_sim(lcss(q, n, c, m, p.eps, p.delta), n, m);
    case HAUSDORFF: return hausdorff<float>(q, n, c, m);
    case EUCLIDEAN: return (float)std::sqrt(euclid2(q, c, n));
  }
  return 0.0f;
}

double exact_value_f64(const Params& p, const Pt* q, int n, const Pt* c, int m) {
  switch (p.measure) {
    case DTW: return dtw<double>(q, n, c, m, p.band);
    case FRECHET: return frechet<double>(q, n, c, m);
    case LCSS: return (double)lcss(q, n, c, m, p.eps, p.delta) / std::min(n, m);
    case HAUSDORFF: return hausdorff<double>(q, n, c, m);
    case EUCLIDEAN: return std::sqrt(euclid2(q, c, n));
  }
  return 0.0;
}

inline bool is_max_measure(int measure) { return measure == LCSS; }

// Candidate order used everywhere: (value, idx) ascending, or for LCSS
// (value descending, idx ascending). dtw.cpp:62 partial_sort on
// pair<float,int>; lcss.cpp:67-70 stable_sort descending.
struct Cand { float v; int32_t i; };
inline bool better(int measure, const Cand& a, const Cand& b) {
  if (a.v != b.v) return is_max_measure(measure) ? a.v > b.v : a.v < b.v;
  return a.i < b.i;
}

// REF-mode filter score of one pair (per-point envelope filters a3/a4/a5).
float ref_score(const Params& p, const Env* env, int n, const Pt* c, int m) {
  switch (p.measure) {
    case DTW: return lb_dtw(env, n, c, m);
    case FRECHET: return lb_frechet(env, n, c, m);
    case LCSS: return (float)ub_lcss_count(env, n, c, m, p.eps) / std::min(n, m);
  }
  return 0.0f;
}

// Admissible summary lower bound of a pair's f32 distance (EXACT mode), and
// for LCSS an upper bound of the similarity. Derivations:
//  DTW: every warping path has >= max(n,m) cells, the first is (0,0) and the
//       last (n-1,m-1); every cell cost >= the MBR gap b (f32 rounding is
//       monotone, so the computed gap never exceeds a computed cost). The
//       f32 running sum loses at most a relative (L+8)*2^-23, removed here.
//  Frechet: max over couplings >= both endpoint costs and >= b; all exact in
//       f32 because min/max do not round.
//  Hausdorff, Euclidean: >= b (Euclidean: sqrt(n)*b up to rounding).
//  LCSS: matches need a point of q inside c's eps-inflated MBR.
float summary_bound(const Params& p, const Pt* q, int n, const Env& qb, const Pt* c, int m, const Env& cb) {
  float g2 = box_box_d2(qb, cb);
  float b = std::sqrt(g2);
  switch (p.measure) {
    case DTW: {
      float e = pdist<float>(q[0], c[0]);
      int L = std::max(n, m);
      if (L >= 2) e = e + pdist<float>(q[n - 1], c[m - 1]);
      float s = e + (float)std::max(0, L - 2) * b;
      return s * (1.0f - (float)(L + 8) * 1.1920929e-7f);
    }
    case FRECHET:
      return std::max(b, std::max(pdist<float>(q[0], c[0]), pdist<float>(q[n - 1], c[m - 1])));
    case HAUSDORFF: return b;
    case EUCLIDEAN: return (float)(std::sqrt((double)n * (double)g2) * (1.0 - 1e-6));
    case LCSS: {
      int cnt = 0;
      for (int i = 0; i < n; ++i) {
        Pt a = q[i];
        if (a.x - cb.maxx < p.eps && cb.minx - a.x < p.eps && a.y - cb.maxy < p.eps && cb.miny - a.y < p.eps) ++cnt;
      }
      return lcss_sim(std::min(cnt, std::min(n, m)), n, m);
    }
  }
  return 0.0f;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- per pair
void orc_envelope(const float* q, int n, int r, float* env) { envelope((const Pt*)q, n, r, (Env*)env); }
float orc_lb_dtw(const float* env, int n, const float* c, int m) { return lb_dtw((const Env*)env, n, (const Pt*)c, m); }
float orc_lb_frechet(const float* env, int n, const float* c, int m) { return lb_frechet((const Env*)env, n, (const Pt*)c, m); }
int orc_ub_lcss_count(const float* env, int n, const float* c, int m, float eps) {
  return ub_lcss_count((const Env*)env, n, (const Pt*)c, m, eps);
}
float orc_dtw(const float* q, int n, const float* c, int m, int band) { return dtw<float>((const Pt*)q, n, (const Pt*)c, m, band); }
double orc_dtw_f64(const float* q, int n, const float* c, int m, int band) { return dtw<double>((const Pt*)q, n, (const Pt*)c, m, band); }
float orc_frechet(const float* q, int n, const float* c, int m) { return frechet<float>((const Pt*)q, n, (const Pt*)c, m); }
double orc_frechet_f64(const float* q, int n, const float* c, int m) { return frechet<double>((const Pt*)q, n, (const Pt*)c, m); }
int orc_lcss(const float* q, int n, const float* c, int m, float eps, int delta) {
  return lcss((const Pt*)q, n, (const Pt*)c, m, eps, delta);
}
float orc_hausdorff(const float* q, int n, const float* c, int m) { return hausdorff<float>((const Pt*)q, n, (const Pt*)c, m); }
double orc_hausdorff_f64(const float* q, int n, const float* c, int m) { return hausdorff<double>((const Pt*)q, n, (const Pt*)c, m); }
float orc_euclidean(const float* q, const float* c, int n) { return (float)std::sqrt(euclid2((const Pt*)q, (const Pt*)c, n)); }
void orc_mbr(const float* t, int len, float* out4) { *(Env*)out4 = mbr((const Pt*)t, len); }
float orc_box_box_d2(const float* a4, const float* b4) { return box_box_d2(*(const Env*)a4, *(const Env*)b4); }

// Summary bound of one pair (EXACT-mode filter value).
float orc_summary_bound(const int32_t* params, const float* q, int n, const float* c, int m) {
  const Params& p = *(const Params*)params;
  Env qb = mbr((const Pt*)q, n), cb = mbr((const Pt*)c, m);
  return summary_bound(p, (const Pt*)q, n, qb, (const Pt*)c, m, cb);
}

// ------------------------------------------------------------ whole k-NN
// REF-mode filter scores for every (query, DB trajectory) pair, row-major
// [Q][N]. DTW/Frechet/LCSS only (the per-point envelope filters).
void orc_filter_ref(const int32_t* params, const float* db_xy, const int64_t* db_off, int N,
                    const float* q_xy, const int64_t* q_off, int Q, float* scores) {
  const Params& p = *(const Params*)params;
  Set db{(const Pt*)db_xy, db_off, N}, qs{(const Pt*)q_xy, q_off, Q};
#pragma omp parallel for schedule(dynamic, 1)
  for (int qi = 0; qi < Q; ++qi) {
    int n = qs.len(qi);
    std::vector<Env> env(n);
    envelope(qs.t(qi), n, p.r, env.data());
    for (int j = 0; j < N; ++j) scores[(int64_t)qi * N + j] = ref_score(p, env.data(), n, db.t(j), db.len(j));
  }
}

// k-NN of every query against the DB.
//  REF mode (the reference pipeline, dtw.cpp:36-109 et al.): choose kc
//    candidates by the reference filter (DTW/Frechet/LCSS: per-point envelope
//    score; Hausdorff/Euclidean: nearest MBRs, hausdorff.cpp:158-162,
//    euclidean.cpp:90-93), refine them exactly, and report the best k of them
//    in candidate order. out_val[q*k] is the reference's h_results[q].
//  EXACT mode: true top-k over the whole DB (brute != 0: refine every pair;
//    otherwise prune with the summary bound, which must give the same set).
// Euclidean compares only DB trajectories of the query's length.
// out_idx/out_val are [Q][k]; unused slots get idx -1. stats[0] receives the
// number of refined pairs.
int orc_knn(const int32_t* params, int brute, const float* db_xy, const int64_t* db_off, int N,
            const float* q_xy, const int64_t* q_off, int Q, int32_t* out_idx, float* out_val, int64_t* stats) {
  const Params& p = *(const Params*)params;
  Set db{(const Pt*)db_xy, db_off, N}, qs{(const Pt*)q_xy, q_off, Q};
  const int k = p.k;
  std::vector<Env> dbb(N);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < N; ++j) dbb[j] = mbr(db.t(j), db.len(j));
  int64_t refined = 0;
  // Fewer queries than threads: queries run one after another and each
  // query's bounds and refines run on every thread instead (a long query
  // would otherwise hold one core while the others idle).
  const bool par_queries = Q >= omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : refined) if (par_queries)
  for (int qi = 0; qi < Q; ++qi) {
    const Pt* q = qs.t(qi);
    int n = qs.len(qi);
    std::vector<Cand> result;
    auto eligible = [&](int j) { return p.measure != EUCLIDEAN || db.len(j) == n; };
    if (p.mode == REF) {
      std::vector<Cand> cand;
      if (p.measure == DTW || p.measure == FRECHET || p.measure == LCSS) {
        std::vector<Env> env(n);
        envelope(q, n, p.r, env.data());
        cand.resize(N);
#pragma omp parallel for schedule(static) if (!par_queries)
        for (int j = 0; j < N; ++j) cand[j] = {ref_score(p, env.data(), n, db.t(j), db.len(j)), j};
      } else {
        Env qb = mbr(q, n);
        for (int j = 0; j < N; ++j)
          if (eligible(j)) cand.push_back({box_box_d2(qb, dbb[j]), j});
      }
      int kc = std::min<int>(p.kc, (int)cand.size());
      int fm = (p.measure == LCSS) ? LCSS : DTW;  // filter order: LCSS descending, others ascending
      std::partial_sort(cand.begin(), cand.begin() + kc, cand.end(),
                        [&](const Cand& a, const Cand& b) { return better(fm, a, b); });
      for (int t = 0; t < kc; ++t) {
        int j = cand[t].i;
        result.push_back({exact_value(p, q, n, db.t(j), db.len(j)), j});
        ++refined;
      }
    } else if (brute) {
      for (int j = 0; j < N; ++j)
        if (eligible(j)) { result.push_back({exact_value(p, q, n, db.t(j), db.len(j)), j}); ++refined; }
    } else {
      // Best-first by bound; stop once the bound cannot beat the k-th best.
      Env qb = mbr(q, n);
      std::vector<Cand> order(N);
#pragma omp parallel for schedule(static) if (!par_queries)
      for (int j = 0; j < N; ++j)
        order[j] = {eligible(j) ? summary_bound(p, q, n, qb, db.t(j), db.len(j), dbb[j]) : NAN, j};
      order.erase(std::remove_if(order.begin(), order.end(), [](const Cand& c) { return std::isnan(c.v); }),
                  order.end());
      bool mx = is_max_measure(p.measure);
      std::sort(order.begin(), order.end(), [&](const Cand& a, const Cand& b) { return better(p.measure, a, b); });
      std::vector<Cand> top;  // current best k, sorted
      // Candidates in bound order, a chunk at a time: the chunk's exact
      // values are computed in parallel (those whose bound already cannot
      // beat the k-th best at the chunk's start are skipped), then merged
      // one by one with the sequential rule: stop at the first bound that
      // cannot beat the current k-th best. Every merged value is the exact
      // value, so the result equals the one-at-a-time search.
      const size_t chunk = par_queries ? 1 : (size_t)4 * omp_get_max_threads();
      std::vector<float> ex;
      bool stop = false;
      for (size_t c0 = 0; c0 < order.size() && !stop; c0 += chunk) {
        const size_t c1 = std::min(order.size(), c0 + chunk);
        const bool full = (int)top.size() == k;
        const float kv = full ? top.back().v : 0.0f;
        ex.assign(c1 - c0, 0.0f);
#pragma omp parallel for schedule(dynamic, 1) if (!par_queries)
        for (size_t t = c0; t < c1; ++t) {
          const Cand& o = order[t];
          const bool beyond = full && (mx ? (o.v < kv) : (o.v > kv));
          if (!beyond && !(mx && o.v == 0.0f)) ex[t - c0] = exact_value(p, q, n, db.t(o.i), db.len(o.i));
        }
        for (size_t t = c0; t < c1; ++t) {
          const Cand& o = order[t];
          if ((int)top.size() == k) {
            const Cand& kth = top.back();
            if (mx ? (o.v < kth.v) : (o.v > kth.v)) {
              stop = true;
              break;
            }
          }
          Cand c{(mx && o.v == 0.0f) ? 0.0f : ex[t - c0], o.i};
          if (!(mx && o.v == 0.0f)) ++refined;
          auto it = std::lower_bound(top.begin(), top.end(), c, [&](const Cand& a, const Cand& b) { return better(p.measure, a, b); });
          top.insert(it, c);
          if ((int)top.size() > k) top.pop_back();
        }
      }
      result = top;
    }
    int kk = std::min<int>(k, (int)result.size());
    std::partial_sort(result.begin(), result.begin() + kk, result.end(),
                      [&](const Cand& a, const Cand& b) { return better(p.measure, a, b); });
    for (int t = 0; t < k; ++t) {
      out_idx[(int64_t)qi * k + t] = t < kk ? result[t].i : -1;
      out_val[(int64_t)qi * k + t] = t < kk ? result[t].v : 0.0f;
    }
  }
  if (stats) stats[0] = refined;
  return 0;
}

// FP64 value of given (query, DB) pairs, for the stated f32 error per measure.
void orc_pair_values_f64(const int32_t* params, const float* db_xy, const int64_t* db_off,
                         const float* q_xy, const int64_t* q_off, const int32_t* qidx,
                         const int32_t* didx, int npairs, double* out) {
  const Params& p = *(const Params*)params;
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < npairs; ++t) {
    const Pt* q = (const Pt*)q_xy + q_off[qidx[t]];
    int n = (int)(q_off[qidx[t] + 1] - q_off[qidx[t]]);
    const Pt* c = (const Pt*)db_xy + db_off[didx[t]];
    int m = (int)(db_off[didx[t] + 1] - db_off[didx[t]]);
    out[t] = exact_value_f64(p, q, n, c, m);
  }
}

void orc_pair_values(const int32_t* params, const float* db_xy, const int64_t* db_off,
                     const float* q_xy, const int64_t* q_off, const int32_t* qidx,
                     const int32_t* didx, int npairs, float* out) {
  const Params& p = *(const Params*)params;
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < npairs; ++t) {
    const Pt* q = (const Pt*)q_xy + q_off[qidx[t]];
    int n = (int)(q_off[qidx[t] + 1] - q_off[qidx[t]]);
    const Pt* c = (const Pt*)db_xy + db_off[didx[t]];
    int m = (int)(db_off[didx[t] + 1] - db_off[didx[t]]);
    out[t] = exact_value(p, q, n, c, m);
  }
}

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

}  // extern "C"
