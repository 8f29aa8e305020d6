// CPU model of k_dp_mw's schedule (cudaruntime1_amd/csrc/dp_mw.hip): a pair
// spans NW waves of a block, strip s (64 rows) on wave s mod NW, strips
// started at the table's steps T_s, a block barrier every kMwB steps, the
// rows between strips handed through two LDS halo buffers.
//
// The waves of a block run independently between barriers, so the model
// runs each superstep (kMwB steps) wave by wave in a given order -- forward
// or reverse -- and tags every halo slot with the pair, strip, chunk and
// superstep that wrote it. Lane 0's halo reads must find the right strip
// and chunk, written before the last barrier that precedes the read: a read
// of a value written in the same superstep by another wave is a race (the
// model flags it whichever order it runs the waves in), a wrong tag a stale
// or overwritten slot. Every halo index is bounds-checked against the
// buffer size the launcher gives (dp_mw_hcap). Results are compared bit for
// bit with the oracle (DTW with and without a band, discrete Frechet, LCSS
// with and without a window).
// Early abandoning of DTW and Frechet pairs as the kernel does it, against
// thresholds below, at, just above or far above each pair's value (or
// none): the flag a strip's last row raises is seen at the next barrier,
// the pair's result is never written before or in the superstep of a drop,
// a pair at or below its threshold is never dropped.
// Usage: dp_mw_schedule [pairs per case] [lag]; exit 0 when all match (a
// lag below 64 + kMwB shows the hazard the lag prevents).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

extern "C" {
float orc_dtw(const float* q, int n, const float* c, int m, int band);
float orc_frechet(const float* q, int n, const float* c, int m);
int orc_lcss(const float* q, int n, const float* c, int m, float eps, int delta);
}
static constexpr float kInf = 1e20f;
static constexpr int kMwB = 8;
enum { DTW = 0, LCSS = 1, FRECHET = 2 };
static int g_lag = 64 + kMwB;
static long g_fail = 0, g_drops = 0;
#define CHK(c)                                                                  \
  do {                                                                          \
    if (!(c)) {                                                                 \
      if (g_fail++ < 5) fprintf(stderr, "check fail line %d: %s\n", __LINE__, #c); \
      return false;                                                             \
    }                                                                           \
  } while (0)

struct Strip { int t0, cs, ce; };
struct Tag { float v; int pair, strip, chunk, epoch; };

static int band_half_width(int band, int R, int C) {
  const int w = band > 0 ? band : (-band * R + 99) / 100;
  return std::min(w, C);
}

// dp_mw_hcap: halo floats per buffer
static int mw_hcap(int u, bool ranged, int cmax, int wmax) {
  long cols = cmax;
  if (ranged) cols = std::min<long>(cmax, (long)wmax + 64 + 2 * u);
  return (int)(((cols + u - 1) / u + 1) * u);
}

template <int MEAS, int U, int NW, bool BAND, bool DELTA>
static bool run_pair(int pid, const std::vector<float>& A, const std::vector<float>& Bpad, int R, int C, int band,
                     int delta, float eps, int hcap, std::vector<Tag> (&halo)[2], bool reverse, float& result, float tau,
                     bool& dropped) {
  constexpr bool kAbandon = MEAS != LCSS;
  const float BND = MEAS == LCSS ? 0.0f : kInf;
  const int lastch = (C - 1) / U;
  const int wb = BAND ? band_half_width(band, R, C) : 0;
  const int bw = wb * R;
  const int res_row = DELTA ? std::min(R - 1, C - 1 + delta) : R - 1;
  const int S = res_row / 64 + 1;
  std::vector<Strip> tbl(S);
  for (int s = 0; s < S; ++s) {
    int lo = 0, hi = C - 1;
    if (BAND) {
      const int r0 = s * 64, r1 = std::min(s * 64 + 63, R - 1);
      const uint32_t a0 = (uint32_t)r0 * (uint32_t)C, a1 = (uint32_t)r1 * (uint32_t)C;
      lo = std::max(0, (int)((a0 + (uint32_t)R - 1u) / (uint32_t)R) - wb);
      hi = std::min(C - 1, (int)(a1 / (uint32_t)R) + wb);
    } else if (DELTA) {
      const int r0 = s * 64, r1 = std::min(s * 64 + 63, R - 1);
      lo = std::max(0, r0 - delta);
      hi = std::min(C - 1, r1 + delta);
    }
    tbl[s].cs = lo / U;
    tbl[s].ce = hi / U + 1;
  }
  int total = 0;
  for (int s = 0; s < S; ++s) {
    int t0 = 0;
    if (s > 0) t0 = tbl[s - 1].t0 + g_lag + (tbl[s].cs - tbl[s - 1].cs);
    if (s >= NW) t0 = std::max(t0, tbl[s - NW].t0 + (tbl[s - NW].ce - tbl[s - NW].cs));
    tbl[s].t0 = t0;
    total = std::max(total, t0 + 63 + (tbl[s].ce - tbl[s].cs));
  }
  const Strip none{INT_MAX / 2, 0, 0};
  auto rlast_of = [&](int row) {
    const int r = std::min(row, R - 1);
    return BAND ? std::min(C - 1, (int)(((uint32_t)r * (uint32_t)C) / (uint32_t)R) + wb) / U : lastch;
  };
  struct Lane {
    int s, row;
    Strip st;
    float ax, ay, o[U], hv[U], up_last, hdiag;
    int hv_tag_ok;
    int rlast;
    float rowmin;
  };
  bool flag = false;
  int flag_epoch = -1;
  dropped = false;
  std::vector<std::vector<Lane>> W(NW, std::vector<Lane>(64));
  for (int w = 0; w < NW; ++w)
    for (int l = 0; l < 64; ++l) {
      Lane& L = W[w][l];
      L.s = w;
      L.st = w < S ? tbl[w] : none;
      L.row = w * 64 + l;
      L.ax = L.ay = 0.0f;
      if (w < S) {
        L.ax = A[2 * std::min(L.row, R - 1)];
        L.ay = A[2 * std::min(L.row, R - 1) + 1];
      }
      for (int j = 0; j < U; ++j) L.o[j] = L.hv[j] = BND;
      L.up_last = L.hdiag = BND;
      L.rlast = rlast_of(L.row);
      L.rowmin = INFINITY;
    }
  bool got = false;
  // lane 0's read for step t1 of wave w, seeing writes of epochs <= vis
  auto lane0_read = [&](int w, int t1, int vis) -> bool {
    Lane& L = W[w][0];
    int s1 = L.s, k1 = t1 - L.st.t0 + L.st.cs;
    Strip st1 = L.st;
    if (k1 >= L.st.ce) {
      s1 = L.s + NW;
      st1 = s1 < S ? tbl[s1] : none;
      k1 = t1 - st1.t0 + st1.cs;
    }
    if (!(s1 < S && k1 >= st1.cs && k1 < st1.ce)) return true;
    if (s1 == 0) {
      for (int j = 0; j < U; ++j) L.hv[j] = BND;
      L.hdiag = BND;
      return true;
    }
    const Strip pv = tbl[s1 - 1];
    const std::vector<Tag>& hb = halo[(s1 - 1) & 1];
    if (k1 < pv.ce) {
      for (int j = 0; j < U; ++j) {
        const int i = (k1 - pv.cs) * U + j;
        CHK(i >= 0 && i < hcap);
        const Tag& g = hb[i];
        CHK(g.pair == pid && g.strip == s1 - 1 && g.chunk == k1 && g.epoch <= vis);
        L.hv[j] = g.v;
      }
    } else {
      for (int j = 0; j < U; ++j) L.hv[j] = BND;
    }
    if (k1 == st1.cs) {
      const int dc = st1.cs * U - 1;
      if (dc >= 0 && dc >= pv.cs * U && dc < pv.ce * U && dc < C) {
        const int i = dc - pv.cs * U;
        CHK(i >= 0 && i < hcap);
        const Tag& g = hb[i];
        CHK(g.pair == pid && g.strip == s1 - 1 && g.chunk == dc / U && g.epoch <= vis);
        L.hdiag = g.v;
      } else {
        L.hdiag = BND;
      }
    }
    return true;
  };
  // one step of wave w
  auto step = [&](int w, int t, int epoch) -> bool {
    std::vector<Lane>& Ls = W[w];
    float oprev[64][U];
    int ps[64], pk[64];  // each lane's strip and chunk at step t-1
    for (int l = 0; l < 64; ++l) {
      for (int j = 0; j < U; ++j) oprev[l][j] = Ls[l].o[j];
      ps[l] = Ls[l].s;
      pk[l] = (t - 1) - Ls[l].st.t0 - l + Ls[l].st.cs;
    }
    for (int l = 0; l < 64; ++l) {
      Lane& L = Ls[l];
      int k = t - L.st.t0 - l + L.st.cs;
      if (k >= L.st.ce) {
        L.s += NW;
        L.st = L.s < S ? tbl[L.s] : none;
        L.row = L.s * 64 + l;
        if (L.s < S) {
          L.ax = A[2 * std::min(L.row, R - 1)];
          L.ay = A[2 * std::min(L.row, R - 1) + 1];
          L.rlast = rlast_of(L.row);
        }
        k = t - L.st.t0 - l + L.st.cs;
      }
      const bool active = L.s < S && k >= L.st.cs && k < L.st.ce;
      const int kc = std::min(std::max(k, 0), lastch);
      float cx[U], cy[U];
      for (int j = 0; j < U; ++j) {
        const int c = kc * U + j;
        CHK(c < (int)Bpad.size() / 2);
        cx[j] = Bpad[2 * c];
        cy[j] = Bpad[2 * c + 1];
      }
      float up[U], on[U];
      for (int j = 0; j < U; ++j) up[j] = l == 0 ? L.hv[j] : oprev[l - 1][j];
      if (active && l > 0) {
        // lane l-1 ran the same strip's same chunk one step before
        CHK(ps[l - 1] == L.s && pk[l - 1] == k);
      }
      const bool first = k == L.st.cs;
      float d = first ? ((L.row == 0) ? (L.st.cs == 0 ? 0.0f : BND) : (l == 0 ? L.hdiag : BND)) : L.up_last;
      if (MEAS == LCSS && first && L.st.cs == 0) d = 0.0f;
      float lf = first ? BND : L.o[U - 1];
      const int bd = BAND ? L.row * C - kc * U * R : 0;
      for (int j = 0; j < U; ++j) {
        const float dx = L.ax - cx[j], dy = L.ay - cy[j];
        float v;
        if (MEAS == DTW) {
          v = std::sqrt(dx * dx + dy * dy) + std::min(std::min(up[j], lf), d);
          if (BAND) {
            const int bj = bd - j * R;
            if ((bj < 0 ? -bj : bj) > bw) v = kInf;
          }
        } else if (MEAS == FRECHET) {
          v = std::max(dx * dx + dy * dy, std::min(std::min(up[j], lf), d));
        } else {
          bool match = std::fabs(dx) < eps && std::fabs(dy) < eps;
          float uu = up[j], ll = lf;
          if (DELTA) {
            const int dd = L.row - (kc * U + j);
            match = match && (dd < 0 ? -dd : dd) <= delta;
            if (dd - 1 > delta || 1 - dd > delta) uu = d;
            if (dd + 1 > delta || -1 - dd > delta) ll = d;
          }
          v = match ? d + 1.0f : std::max(uu, ll);
        }
        on[j] = v;
        d = up[j];
        lf = v;
      }
      L.up_last = up[U - 1];
      for (int j = 0; j < U; ++j) L.o[j] = on[j];
      if (l == 63 && active) {
        for (int j = 0; j < U; ++j) {
          const int i = (k - L.st.cs) * U + j;
          CHK(i >= 0 && i < hcap);
          halo[L.s & 1][i] = Tag{on[j], pid, L.s, k, epoch};
        }
      }
      if (kAbandon) {
        float mn = on[0];
        for (int j = 1; j < U; ++j) mn = fminf(mn, on[j]);
        L.rowmin = fminf(first ? INFINITY : L.rowmin, mn);
        if (l == 63 && active && k == L.rlast && L.row + 1 < R &&
            (MEAS == FRECHET ? std::sqrt(L.rowmin) : L.rowmin) > tau && !flag) {
          CHK(!got);  // never after the result
          flag = true;
          flag_epoch = epoch;
        }
      }
      if (active && L.row == res_row && k == lastch) {
        // a drop raised in this superstep would be a race with this write
        CHK(!flag || flag_epoch < epoch);
        if (!flag) {
          const int rc = (C - 1) - lastch * U;
          float v = on[rc];
          if (MEAS == FRECHET) v = std::sqrt(v);
          if (MEAS == LCSS) v = v / (float)C;
          CHK(!got);
          result = v;
          got = true;
        }
      }
    }
    return true;
  };
  const int nsup = (total + kMwB - 1) / kMwB;
  for (int j = 0; j < nsup; ++j) {
    for (int wi = 0; wi < NW; ++wi) {
      const int w = reverse ? NW - 1 - wi : wi;
      for (int t = j * kMwB; t < std::min(total, (j + 1) * kMwB); ++t) {
        if (!step(w, t, j)) return false;
        // reads before the superstep's barrier see earlier supersteps only
        if ((t + 1) % kMwB != 0 && !lane0_read(w, t + 1, j - 1)) return false;
      }
    }
    // after the barrier: every thread sees a drop raised in this superstep
    if (flag) {
      CHK(flag_epoch == j);
      dropped = true;
      return true;
    }
    if ((j + 1) * kMwB <= total)
      for (int w = 0; w < NW; ++w)
        if (!lane0_read(w, (j + 1) * kMwB, j)) return false;
  }
  CHK(got);
  return true;
}

template <int MEAS, int U, int NW, bool BAND, bool DELTA>
static bool run_case(std::mt19937& rng, int npairs, int band, int delta, float eps, int maxlen) {
  std::uniform_int_distribution<int> len(1, maxlen);
  std::normal_distribution<double> stepd(0.0, 1e-2);
  bool ok = true;
  for (int p = 0; p < npairs; ++p) {
    const int n = len(rng), m = len(rng);
    auto walk = [&](int L) {
      std::vector<float> v(2 * L);
      double x = 0.3, y = 0.6;
      for (int i = 0; i < L; ++i) {
        x += stepd(rng);
        y += stepd(rng);
        v[2 * i] = (float)x;
        v[2 * i + 1] = (float)y;
      }
      return v;
    };
    std::vector<float> q = walk(n), c = walk(m);
    const bool qrow = n >= m;
    const std::vector<float>& A = qrow ? q : c;
    const std::vector<float>& B = qrow ? c : q;
    const int R = std::max(n, m), C = std::min(n, m);
    // the device's padding: to 8 points with the last one
    std::vector<float> Bpad(B);
    while ((Bpad.size() / 2) % 8) {
      Bpad.push_back(B[2 * (C - 1)]);
      Bpad.push_back(B[2 * (C - 1) + 1]);
    }
    int width = C;
    if (BAND) width = std::min(C, 2 * band_half_width(band, R, C) + 1);
    if (DELTA) width = std::min(C, 2 * delta + 1);
    const int hcap = mw_hcap(U, BAND || DELTA, C, width);
    std::vector<Tag> halo[2] = {std::vector<Tag>(hcap, Tag{0, -1, -1, -1, -1}),
                                std::vector<Tag>(hcap, Tag{0, -1, -1, -1, -1})};
    float want;
    if (MEAS == DTW) want = orc_dtw(q.data(), n, c.data(), m, BAND ? band : 0);
    else if (MEAS == FRECHET) want = orc_frechet(q.data(), n, c.data(), m);
    else want = (float)orc_lcss(q.data(), n, c.data(), m, eps, DELTA ? delta : -1) / (float)C;
    static const float kFactor[7] = {0.25f, 0.8f, 1.0f, 1.0f, 1.0001f, 3.0f, INFINITY};
    const int f = std::uniform_int_distribution<int>(0, 6)(rng);
    const float tau = f == 3 ? nextafterf(want, 0.0f) : want * kFactor[f];
    for (int rev = 0; rev < 2; ++rev) {
      float got = NAN;
      bool dropped = false;
      if (!run_pair<MEAS, U, NW, BAND, DELTA>(p, A, Bpad, R, C, band, delta, eps, hcap, halo, rev != 0, got, tau,
                                              dropped)) {
        ok = false;
        continue;
      }
      if (dropped) {
        ++g_drops;
        if (!(want > tau)) {
          if (g_fail++ < 5) fprintf(stderr, "dropped at value %.9g, tau %.9g (meas %d U %d NW %d)\n", want, tau, MEAS, U, NW);
          ok = false;
        }
        continue;
      }
      if (memcmp(&got, &want, 4) != 0) {
        if (g_fail++ < 5)
          fprintf(stderr, "mismatch meas %d U %d NW %d n %d m %d: %.9g vs %.9g\n", MEAS, U, NW, n, m, got, want);
        ok = false;
      }
    }
  }
  return ok;
}

int main(int argc, char** argv) {
  const int np = argc > 1 ? atoi(argv[1]) : 6;
  if (argc > 2) g_lag = atoi(argv[2]);
  std::mt19937 rng(12345);
  bool ok = true;
#define CASES(U, NW)                                                                   \
  ok &= run_case<DTW, U, NW, false, false>(rng, np, 0, -1, 0, 700);                   \
  ok &= run_case<DTW, U, NW, true, false>(rng, np, -10, -1, 0, 900);                  \
  ok &= run_case<DTW, U, NW, true, false>(rng, np, 7, -1, 0, 600);                    \
  ok &= run_case<FRECHET, U, NW, false, false>(rng, np, 0, -1, 0, 500);               \
  ok &= run_case<LCSS, U, NW, false, false>(rng, np, 0, -1, 2e-2f, 500);              \
  ok &= run_case<LCSS, U, NW, false, true>(rng, np, 0, 16, 2e-2f, 900);               \
  ok &= run_case<LCSS, U, NW, false, true>(rng, np, 0, 200, 2e-2f, 700);
  CASES(4, 2)
  CASES(8, 2)
  CASES(4, 4)
  CASES(8, 4)
  printf("%s (%ld failures, %ld pairs dropped early)\n", ok ? "ALL OK" : "FAILED", g_fail, g_drops);
  return ok ? 0 : 1;
}
