// CPU lockstep model of k_dp_lean_r's schedule (cudaruntime1_amd/csrc/dp.hip):
// banded DTW and windowed LCSS. Waves of 64 lanes in groups of W share the
// device work counter and step in the kernel's order -- the previous step's
// result, the group's fetch (descriptor, staged trajectories, the strip
// table of (cstart, cend), the pair's step total and result step), the pair
// start, the U cells, the advance with the next step's LDS reads (strip
// table entry, halo chunk clamped to the row's last chunk, lane 0's diagonal
// at a strip start), lane W-1's halo store into the ring. The halo ring has
// the size dp_halo gives the kernel. Every read that a real cell uses is
// checked to come from the right pair, row and chunk; results are compared
// bit for bit with the oracle's banded DTW and windowed LCSS.
// Early abandoning of banded DTW pairs as the kernel does it (thresholds
// below, at, just above or far above each pair's oracle value, or none): a
// pair at or below its threshold is never dropped, a dropped one exceeds it,
// and the pairs fetched after a drop pass every check.
// Usage: dp_lean_ranged_schedule [trials per shape] [ring divisor] [late];
// exit 0 when all match (a ring divisor above 1 shrinks the ring to show a
// hazard; "late" tests a row's minimum at its strip's padded end, where the
// result may already be out: the model must then fail).
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
int orc_lcss(const float* q, int n, const float* c, int m, float eps, int delta);
}
static constexpr float kInf = 1e20f;
enum { DTW = 0, LCSS = 1 };
static const int kCaps[6] = {31, 63, 127, 255, 1023, 1 << 24};
#define CHK(c)                                                   \
  do {                                                           \
    if (!(c)) {                                                  \
      fprintf(stderr, "check fail line %d: %s\n", __LINE__, #c); \
      abort();                                                   \
    }                                                            \
  } while (0)

// The kernels' next event (group fetch at rem == lg, pair start at rem == 0,
// result at rem + 1 == res_rem): the largest trigger below rem, INT_MIN if none.
static int next_event(int rem, int lg, int res_rem) {
  int ev = INT_MIN;
  if (0 < rem) ev = 0;
  if (lg < rem) ev = std::max(ev, lg);
  if (res_rem != INT_MIN && res_rem - 1 < rem) ev = std::max(ev, res_rem - 1);
  return ev;
}

struct Pair { std::vector<float> a, b; bool dead; float tau; };
static bool g_late = false;
struct Tag { float v; int pair, row, chunk; };

static int band_half_width(int band, int R, int C) {
  const int w = band > 0 ? band : (-band * R + 99) / 100;
  return std::min(w, C);
}

template <bool BAND, bool DELTA, int W, int U>
static void strip_range(int s, int R, int C, int wb, int delta, int lastch, int& clo, int& chi) {
  const int r0 = s * W, r1 = std::min(s * W + W - 1, R - 1);
  int lo, hi;
  if (BAND) {
    const uint32_t a0 = (uint32_t)r0 * (uint32_t)C, a1 = (uint32_t)r1 * (uint32_t)C;
    lo = std::max(0, (int)((a0 + (uint32_t)R - 1u) / (uint32_t)R) - wb);
    hi = std::min(C - 1, (int)(a1 / (uint32_t)R) + wb);
  } else {
    lo = std::max(0, r0 - delta);
    hi = std::min(C - 1, r1 + delta);
  }
  clo = lo / U;
  chi = hi / U;
  (void)lastch;
}

template <int MEAS, int W, int U>
static std::vector<float> emulate(const std::vector<Pair>& pairs, int band, int delta, float eps, int hlen,
                                  int hmask, int ccap, int rmax, int smax, int nwaves, std::vector<int>* dropped) {
  constexpr bool BAND = MEAS == DTW, DELTA = MEAS == LCSS;
  constexpr int G = 64 / W;
  constexpr float BND = MEAS == LCSS ? 0.0f : kInf;
  const int n_items = pairs.size();
  std::vector<float> out(n_items, NAN);
  std::vector<int> written(n_items, 0);
  dropped->assign(n_items, 0);
  const float kNoTau = INFINITY;
  // last chunk holding an in-band cell of row r (the kernel's band_last_chunk)
  auto last_chunk = [](int r, int R, int C, int wb) {
    return std::min(C - 1, (int)(((uint32_t)r * (uint32_t)C) / (uint32_t)R) + wb) / U;
  };
  int next = 0;
  struct Lane {
    int R = 1, C = 1, lastch = 0, S = 1, rem, res_rem = INT_MIN, slot = 0, pair = -1;
    int ev = 0;
    int strip = 0, k = 0, cs = 0, ce = W + 1, pcs = 0, pce = 0, row;
    int bw = 0, bd = 0;
    int nxs = 0, nxe = W + 1, nx_pair = -1;
    int gpar = 0, par = 0;
    bool done = false, wrapped = false, live = false;
    int fR = 1, fC = 1, f_slot = 0, f_total = 0, f_ires = 0, f_lgres = -1, f_pair = -1;
    bool f_ok = false, f_live = false;
    float o[U], hv[U];
    Tag cx[U], cy[U], hvt[U];
    Tag ax{0, -1, -1, -1}, ay{0, -1, -1, -1}, rx{0, -1, -1, -1}, ry{0, -1, -1, -1};
    float up_last = 0;
    Tag hdiag{0, -2, -1, -1};
    int last_row = -1, last_k = -1;
    float f_tau = INFINITY, tau = INFINITY, rowmin = INFINITY;
    int wbl = 0, rlast = 0;
  };
  struct Group {
    std::vector<Tag> bnd, colx, coly, rowx, rowy;
    std::vector<int> ts, te, tpair;
    int slot_item = 0;
  };
  struct Wave { Lane L[64]; Group Gr[G]; bool fin = false; };
  std::vector<Wave> waves(nwaves);
  for (auto& w : waves) {
    for (int l = 0; l < 64; ++l) {
      w.L[l].row = l & (W - 1);
      w.L[l].rem = l & (W - 1);
      w.L[l].ev = w.L[l].rem;
      w.L[l].hdiag.v = BND;
      for (int j = 0; j < U; ++j) { w.L[l].o[j] = 0; w.L[l].hv[j] = BND; w.L[l].cx[j] = w.L[l].cy[j] = {0, -1, -1, -1}; }
    }
    for (auto& g : w.Gr) {
      g.bnd.assign(hlen, {0, -1, -1, -1});
      g.colx.assign(2 * ccap, {0, -1, -1, -1});
      g.coly.assign(2 * ccap, {0, -1, -1, -1});
      g.rowx.assign(rmax, {0, -1, -1, -1});
      g.rowy.assign(rmax, {0, -1, -1, -1});
      g.ts.assign(smax, 0);
      g.te.assign(smax, 0);
      g.tpair.assign(smax, -1);
    }
  }
  auto lgof = [](int l) { return l & (W - 1); };
  int live_waves = nwaves;
  while (live_waves > 0) {
    for (auto& w : waves) {
      if (w.fin) continue;
      Lane* L = w.L;
      // the kernel tests rem == ev; every lane's test must agree with the
      // three conditions it stands for
      bool any_ev = false;
      for (int l = 0; l < 64; ++l) {
        const bool three = L[l].rem == lgof(l) || L[l].rem == 0 || L[l].rem + 1 == L[l].res_rem;
        CHK(three == (L[l].rem == L[l].ev));
        any_ev |= L[l].rem == L[l].ev;
      }
      if (any_ev) {
        for (int l = 0; l < 64; ++l) {
          Lane& ln = L[l];
          if (ln.rem + 1 != ln.res_rem) continue;
          const int res_row = DELTA ? std::min(ln.R - 1, ln.C - 1 + delta) : ln.R - 1;
          CHK(ln.live && ln.last_row == res_row && ln.last_k == ln.lastch);
          const int rc = (ln.C - 1) - ln.lastch * U;
          float v = ln.o[rc];
          if (MEAS == LCSS) v = v / (float)ln.C;
          CHK(!written[ln.slot]);
          written[ln.slot] = 1;
          out[ln.slot] = v;
        }
        for (int g = 0; g < G; ++g) {
          const int l0 = g * W;
          if (L[l0].rem != 0) {
            for (int lg = 0; lg < W; ++lg) CHK(L[l0 + lg].rem != lg);
            continue;
          }
          for (int lg = 0; lg < W; ++lg) CHK(L[l0 + lg].rem == lg);
          Group& gr = w.Gr[g];
          gr.slot_item = next++;
          const int it = gr.slot_item;
          for (int lg = 0; lg < W; ++lg) {
            Lane& ln = L[l0 + lg];
            ln.f_ok = it < n_items;
            ln.f_live = false;
            if (ln.f_ok) {
              const Pair& p = pairs[it];
              ln.f_live = !p.dead;
              if (ln.f_live) {
                const int n = p.a.size() / 2, m = p.b.size() / 2;
                ln.fR = std::max(n, m);
                ln.fC = std::min(n, m);
                ln.f_slot = it;
                ln.f_pair = it;
                ln.f_tau = MEAS == DTW ? p.tau : kNoTau;
              }
            }
          }
          // staging and table (whole group)
          Lane& l0n = L[l0];
          if (l0n.f_ok && l0n.f_live) {
            const Pair& p = pairs[it];
            const int n = p.a.size() / 2, m = p.b.size() / 2;
            const std::vector<float>& A = n >= m ? p.a : p.b;
            const std::vector<float>& B = n >= m ? p.b : p.a;
            for (int i = 0; i < ((l0n.fC + 3) / 4) * 4; ++i) {
              CHK(i < ccap);
              const int src = std::min(i, l0n.fC - 1);
              gr.colx[l0n.gpar * ccap + i] = {B[2 * src], it, -1, i};
              gr.coly[l0n.gpar * ccap + i] = {B[2 * src + 1], it, -1, i};
            }
            for (int i = 0; i < ((l0n.fR + 3) / 4) * 4; ++i) {
              CHK(i < rmax);
              const int src = std::min(i, l0n.fR - 1);
              gr.rowx[i] = {A[2 * src], it, i, -1};
              gr.rowy[i] = {A[2 * src + 1], it, i, -1};
            }
          }
          const int R2 = l0n.f_live ? l0n.fR : 1, C2 = l0n.f_live ? l0n.fC : 1;
          const int lastch2 = (C2 - 1) / U;
          const int wb2 = BAND ? band_half_width(band, R2, C2) : 0;
          const int res_row2 = DELTA ? std::min(R2 - 1, C2 - 1 + delta) : R2 - 1;
          const int S2 = DELTA ? res_row2 / W + 1 : (R2 + W - 1) / W;
          const int sres = res_row2 / W;
          int total = 0, before = 0, at_res = 0;
          CHK(S2 <= smax);
          for (int s2 = 0; s2 < S2; ++s2) {
            int clo, chi, nlo, nhi;
            strip_range<BAND, DELTA, W, U>(s2, R2, C2, wb2, delta, lastch2, clo, chi);
            nlo = clo;
            if (s2 + 1 < S2) strip_range<BAND, DELTA, W, U>(s2 + 1, R2, C2, wb2, delta, lastch2, nlo, nhi);
            const int dur = std::max(chi - clo + 1, W + (nlo - clo) + 1);
            gr.ts[s2] = clo;
            gr.te[s2] = clo + dur;
            gr.tpair[s2] = l0n.f_live ? it : -3;
            total += dur;
            if (s2 < sres) before += dur;
            if (s2 == sres) at_res = lastch2 - clo;
          }
          for (int lg = 0; lg < W; ++lg) {
            Lane& ln = L[l0 + lg];
            ln.f_total = total;
            ln.f_ires = before + at_res;
            ln.f_lgres = ln.f_live ? res_row2 - sres * W : -1;
            ln.gpar ^= 1;
          }
        }
        for (int l = 0; l < 64; ++l) {
          Lane& ln = L[l];
          const int lg = lgof(l);
          Group& gr = w.Gr[l / W];
          if (ln.rem != 0) continue;
          if (!ln.f_ok) {
            ln.done = true;
            ln.rem = -2;
            ln.res_rem = INT_MIN;
            ln.tau = kNoTau;
            continue;
          }
          ln.live = ln.f_live;
          ln.pair = ln.f_live ? ln.f_pair : -3;
          ln.R = ln.f_live ? ln.fR : 1;
          ln.C = ln.f_live ? ln.fC : 1;
          ln.slot = ln.f_slot;
          ln.lastch = (ln.C - 1) / U;
          if (BAND) {
            ln.wbl = band_half_width(band, ln.R, ln.C);
            ln.bw = ln.wbl * ln.R;
          }
          ln.tau = ln.f_live ? ln.f_tau : kNoTau;
          if (BAND) ln.rlast = last_chunk(ln.row, ln.R, ln.C, ln.wbl);
          ln.S = DELTA ? std::min(ln.R - 1, ln.C - 1 + delta) / W + 1 : (ln.R + W - 1) / W;
          ln.rem = ln.f_total;
          ln.res_rem = lg == ln.f_lgres ? ln.f_total - ln.f_ires : INT_MIN;
          ln.strip = 0;
          ln.row = lg;
          ln.wrapped = false;
          CHK(gr.tpair[0] == ln.pair);
          ln.cs = gr.ts[0];
          ln.ce = gr.te[0];
          ln.pcs = ln.pce = 0;
          ln.k = ln.cs;
          const int ni = std::min(1, ln.S - 1);
          ln.nxs = gr.ts[ni];
          ln.nxe = gr.te[ni];
          ln.nx_pair = gr.tpair[ni];
          if (BAND) ln.bd = ln.row * ln.C - ln.cs * U * ln.R;
          ln.par = ln.gpar ^ 1;
          const int lc = U * std::min(ln.cs, ln.lastch);
          for (int j = 0; j < U; ++j) {
            CHK(lc + j < ccap);
            ln.cx[j] = gr.colx[ln.par * ccap + lc + j];
            ln.cy[j] = gr.coly[ln.par * ccap + lc + j];
          }
          ln.ax = gr.rowx[std::min(lg, ln.R - 1)];
          ln.ay = gr.rowy[std::min(lg, ln.R - 1)];
          for (int j = 0; j < U; ++j) { ln.hv[j] = BND; ln.hvt[j] = {BND, -2, -1, -1}; }
        }
        for (int l = 0; l < 64; ++l) L[l].ev = next_event(L[l].rem, lgof(l), L[l].res_rem);
        bool all = true;
        for (int l = 0; l < 64; ++l) all &= L[l].done;
        if (all) {
          w.fin = true;
          --live_waves;
          continue;
        }
      }
      for (int l = 0; l < 64; ++l)
        if (L[l].wrapped) { L[l].ax = L[l].rx; L[l].ay = L[l].ry; }
      float oldo[64][U], on[64][U];
      for (int l = 0; l < 64; ++l) for (int j = 0; j < U; ++j) oldo[l][j] = L[l].o[j];
      for (int l = 0; l < 64; ++l) {
        Lane& ln = L[l];
        const int lg = lgof(l);
        float up[U];
        for (int j = 0; j < U; ++j) up[j] = lg == 0 ? ln.hv[j] : oldo[l - 1][j];
        const bool real = !ln.done && ln.live && ln.row < ln.R && ln.k <= ln.lastch;
        const bool first = ln.k == ln.cs;
        if (real) {
          CHK(ln.ax.pair == ln.pair && ln.ax.row == ln.row);
          for (int j = 0; j < U && ln.k * U + j < ln.C; ++j) CHK(ln.cx[j].pair == ln.pair && ln.cx[j].chunk == ln.k * U + j);
          if (lg == 0 && ln.row > 0 && ln.k < ln.pce) {
            for (int j = 0; j < U; ++j)
              CHK(ln.hvt[j].pair == ln.pair && ln.hvt[j].row == ln.row - 1 && ln.hvt[j].chunk == ln.k);
          }
          if (lg == 0 && ln.row > 0 && first && ln.hdiag.pair != -2) {
            CHK(ln.hdiag.pair == ln.pair && ln.hdiag.row == ln.row - 1 && ln.hdiag.chunk == (ln.cs * U - 1) / U);
          }
        }
        float d = first ? ((ln.row == 0) ? (ln.cs == 0 ? 0.0f : BND) : (lg == 0 ? ln.hdiag.v : BND)) : ln.up_last;
        if (MEAS == LCSS && first && ln.cs == 0) d = 0.0f;
        float lft = first ? BND : ln.o[U - 1];
        for (int j = 0; j < U; ++j) {
          const float dx = ln.ax.v - ln.cx[j].v, dy = ln.ay.v - ln.cy[j].v;
          float v;
          if (MEAS == DTW) {
            v = sqrtf(dx * dx + dy * dy) + std::min(std::min(up[j], lft), d);
            const int bj = ln.bd - j * ln.R;
            if ((bj < 0 ? -bj : bj) > ln.bw) v = kInf;
          } else {
            bool match = fabsf(dx) < eps && fabsf(dy) < eps;
            float uu = up[j], ll = lft;
            const int dd = ln.row - (ln.k * U + j);
            match = match && (dd < 0 ? -dd : dd) <= delta;
            if (dd - 1 > delta || 1 - dd > delta) uu = d;
            if (dd + 1 > delta || -1 - dd > delta) ll = d;
            v = match ? d + 1.0f : std::max(uu, ll);
          }
          on[l][j] = v;
          d = up[j];
          lft = v;
        }
        ln.up_last = up[U - 1];
        ln.last_row = ln.row;
        ln.last_k = ln.k;
      }
      for (int l = 0; l < 64; ++l) for (int j = 0; j < U; ++j) L[l].o[j] = on[l][j];
      // early abandoning (banded DTW): lane W-1's row, not the pair's last,
      // at its last in-band chunk, with a minimum above the threshold
      if (MEAS == DTW) {
        bool stop_g[G] = {};
        for (int l = 0; l < 64; ++l) {
          Lane& ln = L[l];
          float mn = on[l][0];
          for (int j = 1; j < U; ++j) mn = fminf(mn, on[l][j]);
          ln.rowmin = fminf(ln.k == ln.cs ? kNoTau : ln.rowmin, mn);
          const int at = g_late ? ln.ce - 1 : ln.rlast;
          if (lgof(l) == W - 1 && ln.k == at && ln.row + 1 < ln.R && ln.rowmin > ln.tau) stop_g[l / W] = true;
        }
        for (int g = 0; g < G; ++g) {
          if (!stop_g[g]) continue;
          Lane& lw = L[g * W + W - 1];
          CHK(lw.live && !lw.done && lw.slot >= 0 && lw.slot < n_items && !written[lw.slot]);
          written[lw.slot] = 1;
          (*dropped)[lw.slot] = 1;
          out[lw.slot] = NAN;
          for (int lg = 0; lg < W; ++lg) {
            Lane& ln = L[g * W + lg];
            ln.rem = lg + 1;
            ln.ev = lg;
            ln.res_rem = INT_MIN;
            ln.tau = kNoTau;
            ln.live = false;  // the rest of the dropped pair is garbage nothing reads
            ln.pair = -1;
          }
        }
      }
      // advance and the next step's reads (before the halo store)
      int kh[64];
      for (int l = 0; l < 64; ++l) {
        Lane& ln = L[l];
        Group& gr = w.Gr[l / W];
        kh[l] = ln.k;
        ln.wrapped = ln.k + 1 == ln.ce;
        if (ln.wrapped) {
          ++ln.strip;
          ln.row += W;
          if (BAND) ln.rlast = last_chunk(std::min(ln.row, ln.R - 1), ln.R, ln.C, ln.wbl);
          ln.pcs = ln.cs;
          ln.pce = ln.ce;
          if (!ln.done && ln.live && ln.strip < ln.S) CHK(ln.nx_pair == ln.pair);
          ln.cs = ln.nxs;
          ln.ce = ln.nxe;
          ln.k = ln.cs;
          if (BAND) ln.bd = ln.row * ln.C - ln.cs * U * ln.R;
          const int dc = ln.cs * U - 1;
          if (dc >= 0 && dc >= ln.pcs * U && dc < ln.pce * U && dc < ln.C) {
            CHK((dc & hmask) >= 0 && (dc & hmask) < hlen);
            ln.hdiag = gr.bnd[dc & hmask];
            ln.hdiag.chunk = ln.hdiag.chunk;  // tagged with the writer's chunk
          } else {
            ln.hdiag = {BND, -2, -1, -1};
          }
        } else {
          ++ln.k;
          if (BAND) ln.bd -= U * ln.R;
        }
        const int ni = std::min(ln.strip + 1, ln.S - 1);
        CHK(ni >= 0 && ni < smax);
        ln.nxs = gr.ts[ni];
        ln.nxe = gr.te[ni];
        ln.nx_pair = gr.tpair[ni];
        const int lc = U * std::min(ln.k, ln.lastch);
        ln.rx = gr.rowx[std::min(ln.row, ln.R - 1)];
        ln.ry = gr.rowy[std::min(ln.row, ln.R - 1)];
        for (int j = 0; j < U; ++j) {
          CHK(lc + j < ccap);
          ln.cx[j] = gr.colx[ln.par * ccap + lc + j];
          ln.cy[j] = gr.coly[ln.par * ccap + lc + j];
        }
        if (ln.row == 0 || ln.k >= ln.pce) {
          for (int j = 0; j < U; ++j) { ln.hv[j] = BND; ln.hvt[j] = {BND, -2, -1, -1}; }
        } else {
          for (int j = 0; j < U; ++j) {
            const int idx = (lc & hmask) + j;
            CHK(idx >= 0 && idx < hlen);
            ln.hv[j] = gr.bnd[idx].v;
            ln.hvt[j] = gr.bnd[idx];
          }
        }
      }
      for (int g = 0; g < G; ++g) {
        const int l = g * W + W - 1;
        Lane& ln = L[l];
        Group& gr = w.Gr[g];
        if (kh[l] <= ln.lastch) {
          for (int j = 0; j < U; ++j) {
            const int idx = ((kh[l] * U) & hmask) + j;
            CHK(idx >= 0 && idx < hlen);
            gr.bnd[idx] = {on[l][j], ln.pair, ln.last_row, kh[l]};
          }
        }
      }
      for (int l = 0; l < 64; ++l) --L[l].rem;
    }
  }
  for (int i = 0; i < n_items; ++i) CHK(pairs[i].dead || written[i]);
  return out;
}

// dp_halo for a ranged launch (bucket b of the widest pair)
static void halo_of(int W, int U, int cmax_c, int wmax_c, int& hlen, int& hmask, int div) {
  int b = 0;
  while (wmax_c > kCaps[b]) ++b;
  const int chunks = std::max((cmax_c + U - 1) / U, W);
  const int full = ((chunks * U + 3) / 4) * 4;
  const int wd = std::min(wmax_c, kCaps[b]);
  const int need = std::max(W + wd + 4 * U, W * U + W + 4 * U);
  int h = 4;
  while (h < need) h *= 2;
  if (h >= full && div == 1) { hlen = full; hmask = -1; return; }
  h /= div;
  if (h < U) h = U;
  hlen = h;
  hmask = h - 1;
}

template <int MEAS, int W, int U>
static int run(std::mt19937& rng, int trials, int div) {
  int bad = 0, drops = 0;
  for (int t = 0; t < trials; ++t) {
    const int maxlen = std::uniform_int_distribution<int>(1, 400)(rng);
    const int npairs = std::uniform_int_distribution<int>(1, 30)(rng);
    const int band = std::uniform_int_distribution<int>(0, 1)(rng) ? std::uniform_int_distribution<int>(1, 40)(rng)
                                                                     : -std::uniform_int_distribution<int>(1, 60)(rng);
    const int delta = std::uniform_int_distribution<int>(0, 70)(rng);
    std::vector<Pair> pairs(npairs);
    int cmax_c = 1, rmax_c = 1, wmax_c = 1;
    for (auto& p : pairs) {
      const int n = std::uniform_int_distribution<int>(1, maxlen)(rng), m = std::uniform_int_distribution<int>(1, maxlen)(rng);
      p.dead = std::uniform_int_distribution<int>(0, 9)(rng) == 0;
      std::normal_distribution<float> nd(0.0f, 0.03f);
      float x = 0, y = 0;
      for (int i = 0; i < n; ++i) { x += nd(rng); y += nd(rng); p.a.push_back(x); p.a.push_back(y); }
      x = 0.01f; y = 0;
      for (int i = 0; i < m; ++i) { x += nd(rng); y += nd(rng); p.b.push_back(x); p.b.push_back(y); }
      const int c = std::min(n, m), r = std::max(n, m);
      cmax_c = std::max(cmax_c, c);
      rmax_c = std::max(rmax_c, r);
      int e = c;
      if (MEAS == DTW) { const int w = band > 0 ? band : (-band * r + 99) / 100; e = std::min(c, 2 * w + 1); }
      else e = std::min(c, 2 * delta + 1);
      wmax_c = std::max(wmax_c, e);
    }
    int hlen, hmask;
    halo_of(W, U, cmax_c, wmax_c, hlen, hmask, div);
    const int rmax = ((rmax_c + 3) / 4) * 4;
    const int smax = (rmax + W - 1) / W;
    const float eps = 0.03f;
    int ccap = ((cmax_c + 3) / 4) * 4;
    ccap = ((ccap + U - 1) / U) * U;  // launch_dp_s: whole chunks
    std::vector<float> wants(npairs, 0.0f);
    for (int i = 0; i < npairs; ++i) {
      Pair& p = pairs[i];
      const int n = p.a.size() / 2, m = p.b.size() / 2;
      if (MEAS == DTW) wants[i] = orc_dtw(p.a.data(), n, p.b.data(), m, band);
      else wants[i] = (float)orc_lcss(p.a.data(), n, p.b.data(), m, eps, delta) / (float)std::min(n, m);
      static const float kFactor[7] = {0.25f, 0.8f, 1.0f, 1.0f, 1.0001f, 3.0f, INFINITY};
      const int f = std::uniform_int_distribution<int>(0, 6)(rng);
      p.tau = f == 3 ? nextafterf(wants[i], 0.0f) : wants[i] * kFactor[f];
    }
    std::vector<int> dropped;
    std::vector<float> got = emulate<MEAS, W, U>(pairs, band, delta, eps, hlen, hmask, ccap, rmax, smax, 1 + t % 3,
                                                 &dropped);
    for (int i = 0; i < npairs; ++i) {
      if (pairs[i].dead) continue;
      const Pair& p = pairs[i];
      const int n = p.a.size() / 2, m = p.b.size() / 2;
      const float want = wants[i];
      if (dropped[i]) {
        ++drops;
        if (!(want > p.tau)) {
          if (bad < 5) fprintf(stderr, "MEAS %d W %d U %d trial %d pair %d: dropped at value %.9g, tau %.9g\n", MEAS, W, U, t, i, want, p.tau);
          ++bad;
        }
        continue;
      }
      if (memcmp(&want, &got[i], 4) != 0) {
        if (bad < 5) fprintf(stderr, "MEAS %d W %d U %d trial %d pair %d n %d m %d band %d delta %d: got %.9g want %.9g\n",
                             MEAS, W, U, t, i, n, m, band, delta, got[i], want);
        ++bad;
      }
    }
  }
  printf("MEAS %d W %2d U %d: %d trials, %d pairs dropped early, %d mismatches\n", MEAS, W, U, trials, drops, bad);
  return bad;
}

int main(int argc, char** argv) {
  const int T = argc > 1 ? std::atoi(argv[1]) : 40;
  const int div = argc > 2 ? std::atoi(argv[2]) : 1;
  g_late = argc > 3 && std::strcmp(argv[3], "late") == 0;
  std::mt19937 rng(4242);
  int bad = 0;
  bad += run<DTW, 16, 1>(rng, T, div); bad += run<DTW, 16, 4>(rng, T, div); bad += run<DTW, 32, 2>(rng, T, div);
  bad += run<DTW, 32, 8>(rng, T, div); bad += run<DTW, 64, 4>(rng, T, div); bad += run<DTW, 64, 8>(rng, T, div);
  bad += run<LCSS, 16, 2>(rng, T, div); bad += run<LCSS, 16, 4>(rng, T, div); bad += run<LCSS, 32, 4>(rng, T, div);
  bad += run<LCSS, 64, 2>(rng, T, div); bad += run<LCSS, 64, 8>(rng, T, div);
  printf(bad ? "FAIL\n" : "ALL OK\n");
  return bad != 0;
}
