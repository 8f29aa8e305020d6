// CPU lockstep model of k_dp_lean's schedule (cudaruntime1_amd/csrc/dp.hip):
// one or more waves of 64 lanes in groups of W share the device work counter
// and step in the kernel's order -- the previous step's result, the group's
// pair fetch (descriptor and both trajectories staged), the pair start, the
// next step's LDS reads, the U cells with the previous lane's outputs as
// "up", lane W-1's halo store, the advance. Every LDS index is bounds-checked
// against the sizes the host gives the kernel (halo of max(chunks, W+1)
// chunks, rows rounded up to 4 points), every staged operand carries the
// pair and index that wrote it and is checked where a real cell uses it, and
// results are compared bit for bit with the oracle's DTW, Frechet and LCSS.
// Early abandoning as the kernel does it: every DTW / Frechet pair has a
// threshold (below, at, just above or far above its oracle value, or none);
// lane W-1's finished row minimum above it drops the pair for the group,
// which fetches the next one at once. A pair whose value is at most its
// threshold must never be dropped, a dropped one must exceed it, and the
// pairs fetched after a drop pass every staging and halo check.
// DTW pairs also test the row minimum plus the suffix table the group builds
// at the fetch (the sum of each later row's bound to the nearest 64-point box
// of the column trajectory), scaled by the kernel's discount: the boxes are
// staged over the group's halo row, which the model then fills with tags no
// real cell may read, and the bounds take v_sqrt_f32's largest result.
// Usage: dp_lean_schedule [trials per shape] [late|nodisc]; exit status 0
// when all match. "late" tests the row minimum at the row's last chunk
// including its idle ones, a step too late, and "nodisc" drops the suffix
// test's discount: the model must then fail.
#include <algorithm>
#include <array>
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
enum { DTW = 0, LCSS = 1, FRECHET = 2 };
#define CHK(c)                                                   \
  do {                                                           \
    if (!(c)) {                                                  \
      fprintf(stderr, "check fail line %d: %s\n", __LINE__, #c); \
      abort();                                                   \
    }                                                            \
  } while (0)

// The kernels' next event (group fetch at rem == lg, pair start at rem == 0,
// result at rem + 1 == res_rem): the largest trigger below rem, INT_MIN if none.
static bool g_late = false;    // test rows at their padded end (a planted mistake)
static bool g_nodisc = false;  // suffix test without its discount (a planted mistake)
static bool g_sfx1 = false;    // suffix table one strip short: it counts the tested strip's rows (a planted mistake)
static long g_sfx_drops = 0;   // drops the suffix test made (with or without the row test)

// sqrt_lo with v_sqrt_f32's largest result (1 ulp above the root)
static float sqrt_lo_max(float x) {
  return x < 0x1p-96f ? 0.0f : nextafterf(sqrtf(x), INFINITY) * (1.0f - 0x1p-21f);
}
static float pt_box(float px, float py, const float b[4]) {
  const float gx = std::max(0.0f, std::max(b[0] - px, px - b[1]));
  const float gy = std::max(0.0f, std::max(b[2] - py, py - b[3]));
  return gx * gx + gy * gy;
}

static int next_event(int rem, int lg, int res_rem) {
  int ev = INT_MIN;
  if (0 < rem) ev = 0;
  if (lg < rem) ev = std::max(ev, lg);
  if (res_rem != INT_MIN && res_rem - 1 < rem) ev = std::max(ev, res_rem - 1);
  return ev;
}

struct Pair { std::vector<float> a, b; bool dead; float tau; };  // a: query (xy), b: db (xy)

struct Tagged { float v; int pair; int idx; };

template <int MEAS, int W, int U>
std::vector<float> emulate(const std::vector<Pair>& pairs, float eps, int cmax, int rmax, int nwaves, long* steps,
                           std::vector<int>* dropped, int sfx) {
  constexpr int G = 64 / W;
  constexpr float BND = MEAS == LCSS ? 0.0f : kInf;
  const int n_items = pairs.size();
  std::vector<float> out(n_items, NAN);
  std::vector<int> written(n_items, 0);
  dropped->assign(n_items, 0);
  const float kNoTau = INFINITY;
  int next = 0;
  struct Lane {
    int R = 1, C = 1, lastch = 0, D = W + 1, k = 0, row, rem, res_rem = INT_MIN, slot = 0, par = 0, gpar = 0;
    int ev = 0;
    bool done = false, wrapped = false;
    int fR = 1, fC = 1, f_slot = 0, f_pair = -1, pair = -1;
    bool f_ok = false, f_live = false, live = false;
    float o[U], cx[U], cy[U], hv[U];
    Tagged ax{0, -1, -1}, ay{0, -1, -1}, rx{0, -1, -1}, ry{0, -1, -1};
    int cxpair = -1, cxidx = -1;
    int hvpair = -2, hvidx = -1, hvrow = -1;
    float up_last = 0, dfirst = 0;
    int lastres_row = -1, lastres_k = -1;
    float f_tau = INFINITY, tau = INFINITY, rowmin = INFINITY;
    float f_disc = 0, disc = 0;
  };
  struct Group {
    std::vector<Tagged> bnd, sbx, sby, sax, say;
    int slot_item = 0;
    std::vector<float> sfxt;  // suffix table
    std::vector<int> bnd_row;  // row whose values each halo slot holds
  };
  struct Wave { Lane L[64]; Group Gr[G]; bool fin = false; };
  std::vector<Wave> waves(nwaves);
  for (auto& w : waves) {
    for (int l = 0; l < 64; ++l) {
      Lane& ln = w.L[l];
      ln.row = l & (W - 1);
      ln.rem = l & (W - 1);
      ln.ev = ln.rem;
      for (int j = 0; j < U; ++j) ln.o[j] = ln.cx[j] = ln.cy[j] = ln.hv[j] = 0;
      ln.dfirst = (MEAS == LCSS || (l & (W - 1)) == 0) ? 0.0f : BND;
    }
    for (auto& g : w.Gr) {
      g.bnd.assign(cmax, {0, -1, -1});
      g.bnd_row.assign(cmax, -1);
      g.sbx.assign(2 * cmax, {0, -1, -1});
      g.sby.assign(2 * cmax, {0, -1, -1});
      g.sax.assign(rmax, {0, -1, -1});
      g.say.assign(rmax, {0, -1, -1});
      g.sfxt.assign(sfx, NAN);
    }
  }
  long st = 0;
  int live_waves = nwaves;
  while (live_waves > 0) {
    for (auto& w : waves) {
      if (w.fin) continue;
      ++st;
      Lane* L = w.L;
      auto lgof = [](int l) { return l & (W - 1); };
      // the kernel tests rem == ev; every lane's test must agree with the
      // three conditions it stands for
      bool any_ev = false;
      for (int l = 0; l < 64; ++l) {
        const bool three = L[l].rem == lgof(l) || L[l].rem == 0 || L[l].rem + 1 == L[l].res_rem;
        CHK(three == (L[l].rem == L[l].ev));
        any_ev |= L[l].rem == L[l].ev;
      }
      if (any_ev) {
        // result from the outputs of the step before
        for (int l = 0; l < 64; ++l) {
          Lane& ln = L[l];
          if (ln.rem + 1 != ln.res_rem) continue;
          CHK(ln.live && !ln.done && ln.lastres_row == ln.R - 1 && ln.lastres_k == ln.lastch);
          const int rc = (ln.C - 1) - ln.lastch * U;
          CHK(rc >= 0 && rc < U);
          float v = ln.o[rc];
          if (MEAS == FRECHET) v = sqrtf(v);
          if (MEAS == LCSS) v = v / (float)ln.C;
          CHK(ln.slot >= 0 && ln.slot < n_items && !written[ln.slot]);
          written[ln.slot] = 1;
          out[ln.slot] = v;
        }
        // fetch: lane 0 of each fetching group takes the counter
        for (int g = 0; g < G; ++g) {
          const int l0 = g * W;
          if (L[l0].rem == 0) {
            for (int lg = 0; lg < W; ++lg) CHK(L[l0 + lg].rem == lg);
            w.Gr[g].slot_item = next++;
          } else {
            for (int lg = 0; lg < W; ++lg) CHK(L[l0 + lg].rem != lg);
          }
        }
        for (int g = 0; g < G; ++g) {
          const int l0 = g * W;
          if (L[l0].rem != 0) continue;
          Group& gr = w.Gr[g];
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
                const std::vector<float>& A = n >= m ? p.a : p.b;
                const std::vector<float>& B = n >= m ? p.b : p.a;
                ln.fR = std::max(n, m);
                ln.fC = std::min(n, m);
                ln.f_slot = it;
                ln.f_pair = it;
                ln.f_tau = MEAS == LCSS ? kNoTau : p.tau;
                for (int c = lg; c < ((ln.fC + 3) >> 2); c += W)
                  for (int e = 0; e < 4; ++e) {
                    const int i = 4 * c + e, src = std::min(i, ln.fC - 1);
                    CHK(ln.gpar * cmax + i < 2 * cmax);
                    gr.sbx[ln.gpar * cmax + i] = {B[2 * src], it, i};
                    gr.sby[ln.gpar * cmax + i] = {B[2 * src + 1], it, i};
                  }
                for (int c = lg; c < ((ln.fR + 3) >> 2); c += W)
                  for (int e = 0; e < 4; ++e) {
                    const int i = 4 * c + e, src = std::min(i, ln.fR - 1);
                    CHK(i < rmax);
                    gr.sax[i] = {A[2 * src], it, i};
                    gr.say[i] = {A[2 * src + 1], it, i};
                  }
              }
            }
            ln.gpar ^= 1;
          }
          // suffix table (lean_suffix_table), group-uniform
          Lane& lz = L[l0];
          for (int lg = 0; lg < W; ++lg) L[l0 + lg].f_disc = 0.0f;
          const int S = (lz.fR + W - 1) / W;
          const int nbox = (((lz.fC + 7) >> 3) + 7) >> 3;
          if (MEAS == DTW && sfx > 0 && lz.f_ok && lz.f_live && lz.f_tau < kNoTau && S >= 2 && S - 1 <= sfx &&
              4 * nbox <= cmax) {
            const Pair& p = pairs[it];
            const int n = p.a.size() / 2, m = p.b.size() / 2;
            const std::vector<float>& A = n >= m ? p.a : p.b;
            const std::vector<float>& B = n >= m ? p.b : p.a;
            const int R = lz.fR, C = lz.fC;
            // 64-point boxes of B (the set's superblocks)
            std::vector<std::array<float, 4>> box(nbox);
            for (int t = 0; t < nbox; ++t) {
              box[t] = {B[2 * 64 * t], B[2 * 64 * t], B[2 * 64 * t + 1], B[2 * 64 * t + 1]};
              for (int i = 64 * t; i < std::min(C, 64 * t + 64); ++i) {
                box[t][0] = std::min(box[t][0], B[2 * i]); box[t][1] = std::max(box[t][1], B[2 * i]);
                box[t][2] = std::min(box[t][2], B[2 * i + 1]); box[t][3] = std::max(box[t][3], B[2 * i + 1]);
              }
            }
            // the boxes overwrite the halo row: tag it so no real cell reads it
            for (int i = 0; i < 4 * nbox; ++i) { gr.bnd[i] = {NAN, -3, -3}; gr.bnd_row[i] = -3; }
            // strips from the last, each summed by the DPP reduction into lane W-1
            float acc = 0.0f;
            for (int s = S - 1; s >= 1; --s) {
              float v[64];
              for (int lg = 0; lg < W; ++lg) {
                const int i = s * W + lg;
                float mb = INFINITY;
                if (i < R)
                  for (int t = 0; t < nbox; ++t) mb = std::min(mb, pt_box(A[2 * i], A[2 * i + 1], box[t].data()));
                v[lg] = i < R ? sqrt_lo_max(mb) : 0.0f;
              }
              for (int d = 1; d < 16; d <<= 1) {
                float nv[64];
                for (int lg = 0; lg < W; ++lg) nv[lg] = v[lg] + ((lg & 15) >= d ? v[lg - d] : 0.0f);
                for (int lg = 0; lg < W; ++lg) v[lg] = nv[lg];
              }
              if (W >= 32) for (int lg = 16; lg < 32; ++lg) v[lg] += v[15];
              if (W == 64) {
                for (int lg = 48; lg < 64; ++lg) v[lg] += v[47];
                for (int lg = 32; lg < 64; ++lg) v[lg] += v[31];
              }
              acc += v[W - 1];
              gr.sfxt[s - 1] = acc;
            }
            if (g_sfx1)  // entry s also counts strip s
              for (int s = S - 2; s >= 1; --s) gr.sfxt[s] = gr.sfxt[s - 1];
            const float disc = g_nodisc ? 1.0f : 1.0f - (float)(2 * R + C + 64) * 0x1p-23f;
            for (int lg = 0; lg < W; ++lg) L[l0 + lg].f_disc = disc;
          }
        }
        // start
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
          ln.tau = ln.f_live ? ln.f_tau : kNoTau;
          ln.disc = ln.f_live ? ln.f_disc : 0.0f;
          ln.pair = ln.f_live ? ln.f_pair : -1;
          ln.R = ln.f_live ? ln.fR : 1;
          ln.C = ln.f_live ? ln.fC : 1;
          ln.slot = ln.f_slot;
          ln.lastch = (ln.C - 1) / U;
          ln.D = std::max(ln.lastch + 1, W + 1);
          const int S = (ln.R + W - 1) / W;
          ln.rem = S * ln.D;
          ln.res_rem = (ln.f_live && lg == ln.R - 1 - (S - 1) * W) ? ln.D - ln.lastch : INT_MIN;
          ln.k = 0;
          ln.row = lg;
          ln.par = ln.gpar ^ 1;
          const int r0 = std::min(lg, ln.R - 1);
          CHK(r0 < rmax);
          ln.ax = gr.sax[r0];
          ln.ay = gr.say[r0];
          for (int j = 0; j < U; ++j) {
            CHK(ln.par * cmax + j < 2 * cmax);
            ln.cx[j] = gr.sbx[ln.par * cmax + j].v;
            ln.cy[j] = gr.sby[ln.par * cmax + j].v;
          }
          ln.cxpair = gr.sbx[ln.par * cmax].pair;
          ln.cxidx = gr.sbx[ln.par * cmax].idx;
          for (int j = 0; j < U; ++j) ln.hv[j] = BND;
          ln.hvrow = -1;
          ln.hvpair = -2;
          ln.dfirst = (MEAS == LCSS || lg == 0) ? 0.0f : BND;
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
      // next step's operands, read before this step's halo store
      struct Nx { bool wrap; int k1, row1; Tagged rx, ry; float cx[U], cy[U], hv[U]; int cxpair, cxidx, hvpair, hvidx, hvrow; float dfirst; };
      static Nx nx[64];
      for (int l = 0; l < 64; ++l) {
        Lane& ln = L[l];
        Group& gr = w.Gr[l / W];
        Nx& n = nx[l];
        n.wrap = ln.k + 1 == ln.D;
        n.k1 = n.wrap ? 0 : ln.k + 1;
        n.row1 = n.wrap ? ln.row + W : ln.row;
        const int rr = std::min(n.row1, ln.R - 1);
        CHK(rr >= 0 && rr < rmax);
        n.rx = gr.sax[rr];
        n.ry = gr.say[rr];
        for (int j = 0; j < U; ++j) {
          CHK(ln.par * cmax + n.k1 * U + j < 2 * cmax && n.k1 * U + j < cmax);
          n.cx[j] = gr.sbx[ln.par * cmax + n.k1 * U + j].v;
          n.cy[j] = gr.sby[ln.par * cmax + n.k1 * U + j].v;
        }
        n.cxpair = gr.sbx[ln.par * cmax + n.k1 * U].pair;
        n.cxidx = gr.sbx[ln.par * cmax + n.k1 * U].idx;
        if (n.row1 == 0) {
          for (int j = 0; j < U; ++j) n.hv[j] = BND;
          n.hvpair = -2;
        } else {
          for (int j = 0; j < U; ++j) {
            CHK(n.k1 * U + j < cmax);
            n.hv[j] = gr.bnd[n.k1 * U + j].v;
          }
          n.hvpair = gr.bnd[n.k1 * U].pair;
          n.hvidx = gr.bnd[n.k1 * U].idx;
          n.hvrow = gr.bnd_row[n.k1 * U];
        }
        n.dfirst = (MEAS == LCSS || n.row1 == 0) ? 0.0f : BND;
      }
      // cells: up from the previous lane's outputs of the last step
      float oldo[64][U];
      for (int l = 0; l < 64; ++l) for (int j = 0; j < U; ++j) oldo[l][j] = L[l].o[j];
      float on[64][U];
      for (int l = 0; l < 64; ++l) {
        Lane& ln = L[l];
        const int lg = lgof(l);
        float up[U];
        for (int j = 0; j < U; ++j) up[j] = lg == 0 ? ln.hv[j] : oldo[l - 1][j];
        const bool real = !ln.done && ln.live && ln.row < ln.R && ln.k <= ln.lastch && ln.pair >= 0;
        if (real) {
          CHK(ln.ax.pair == ln.pair && ln.ax.idx == ln.row);
          CHK(ln.ay.pair == ln.pair && ln.ay.idx == ln.row);
          CHK(ln.cxpair == ln.pair && ln.cxidx == ln.k * U);
          if (lg == 0 && ln.row > 0) CHK(ln.hvpair == ln.pair && ln.hvrow == ln.row - 1 && ln.hvidx == ln.k * U);
        }
        const bool first = ln.k == 0;
        float d = first ? ln.dfirst : ln.up_last;
        float lft = first ? BND : ln.o[U - 1];
        for (int j = 0; j < U; ++j) {
          const float dx = ln.ax.v - ln.cx[j], dy = ln.ay.v - ln.cy[j];
          float v;
          if (MEAS == DTW) v = sqrtf(dx * dx + dy * dy) + std::min(std::min(up[j], lft), d);
          else if (MEAS == FRECHET) v = std::max(dx * dx + dy * dy, std::min(std::min(up[j], lft), d));
          else {
            const bool match = fabsf(dx) < eps && fabsf(dy) < eps;
            v = match ? d + 1.0f : std::max(up[j], lft);
          }
          on[l][j] = v;
          d = up[j];
          lft = v;
        }
        ln.up_last = up[U - 1];
        ln.lastres_row = ln.row;
        ln.lastres_k = ln.k;
      }
      for (int l = 0; l < 64; ++l) for (int j = 0; j < U; ++j) L[l].o[j] = on[l][j];
      // halo store by lane W-1 of each group
      for (int g = 0; g < G; ++g) {
        Lane& ln = L[g * W + W - 1];
        Group& gr = w.Gr[g];
        for (int j = 0; j < U; ++j) {
          CHK(ln.k * U + j < cmax);
          gr.bnd[ln.k * U + j] = {on[g * W + W - 1][j], ln.pair, ln.k * U};
          gr.bnd_row[ln.k * U + j] = ln.row;
        }
      }
      // early abandoning (DTW, Frechet): lane W-1's row, not the pair's last,
      // at its last real chunk (later, in the row's idle chunks, a lane of the
      // next strip may already hold the pair's result) with a minimum above
      // the threshold drops the group's pair
      bool stop_g[G] = {};
      for (int l = 0; l < 64; ++l) {
        Lane& ln = L[l];
        float mn = on[l][0];
        for (int j = 1; j < U; ++j) mn = fminf(mn, on[l][j]);
        ln.rowmin = fminf(ln.k == 0 ? kNoTau : ln.rowmin, mn);
        if (MEAS != LCSS && lgof(l) == W - 1 && ln.k == (g_late ? ln.D - 1 : ln.lastch) && ln.row + 1 < ln.R) {
          const float v = MEAS == DTW ? ln.rowmin : sqrtf(ln.rowmin);
          if (v > ln.tau) stop_g[l / W] = true;
          if (MEAS == DTW && ln.disc > 0.0f) {
            const int s = ln.row / W;
            CHK(s >= 0 && s < sfx && s < (ln.R + W - 1) / W - 1);
            const float tv = w.Gr[l / W].sfxt[s];
            CHK(tv == tv);  // written for this pair
            if ((ln.rowmin + tv) * ln.disc > ln.tau) stop_g[l / W] = true, ++g_sfx_drops;
          }
        }
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
          ln.tau = kNoTau;  // no second drop before the next pair starts
          // the dropped pair's remaining cells are garbage nothing reads (the
          // next pair's rows may already be staged under them)
          ln.live = false;
          ln.pair = -1;
        }
      }
      // advance
      for (int l = 0; l < 64; ++l) {
        Lane& ln = L[l];
        Nx& n = nx[l];
        --ln.rem;
        ln.k = n.k1;
        ln.row = n.row1;
        if (n.wrap) { ln.ax = n.rx; ln.ay = n.ry; }
        for (int j = 0; j < U; ++j) { ln.cx[j] = n.cx[j]; ln.cy[j] = n.cy[j]; ln.hv[j] = n.hv[j]; }
        ln.cxpair = n.cxpair; ln.cxidx = n.cxidx;
        ln.hvpair = n.hvpair; ln.hvidx = n.hvidx; ln.hvrow = n.hvrow;
        ln.dfirst = n.dfirst;
      }
    }
  }
  for (int i = 0; i < n_items; ++i) CHK(pairs[i].dead || written[i]);
  *steps = st;
  return out;
}

template <int MEAS, int W, int U>
int run(std::mt19937& rng, int trials) {
  int bad = 0, drops = 0;
  for (int t = 0; t < trials; ++t) {
    const int maxlen = std::uniform_int_distribution<int>(1, 300)(rng);
    const int npairs = std::uniform_int_distribution<int>(1, 40)(rng);
    std::vector<Pair> pairs(npairs);
    int cmaxlen = 1, rmaxlen = 1;
    for (auto& p : pairs) {
      int n = std::uniform_int_distribution<int>(1, maxlen)(rng), m = std::uniform_int_distribution<int>(1, maxlen)(rng);
      // one-point trajectories: DTW is the sum of the other's distances to
      // the point, and the suffix bound is tight up to its rounding
      if (std::uniform_int_distribution<int>(0, 7)(rng) == 0) (n < m ? n : m) = 1;
      p.dead = std::uniform_int_distribution<int>(0, 9)(rng) == 0;
      std::normal_distribution<float> nd(0.0f, 0.03f);
      float x = 0, y = 0;
      for (int i = 0; i < n; ++i) { x += nd(rng); y += nd(rng); p.a.push_back(x); p.a.push_back(y); }
      x = 0.01f; y = 0;
      for (int i = 0; i < m; ++i) { x += nd(rng); y += nd(rng); p.b.push_back(x); p.b.push_back(y); }
      cmaxlen = std::max(cmaxlen, std::min(n, m));
      rmaxlen = std::max(rmaxlen, std::max(n, m));
    }
    // host sizes as in dp_halo / launch_dp: halo row of max(chunks, W) chunks, rows rounded to 4
    const int chunks = std::max((cmaxlen + U - 1) / U, W + 1);
    const int cmax = ((chunks * U + 3) / 4) * 4;
    const int rmax = ((rmaxlen + 3) / 4) * 4;
    long steps = 0;
    const float eps = 0.03f;
    const int nwaves = 1 + t % 3;
    std::vector<float> wants(npairs, 0.0f);
    for (int i = 0; i < npairs; ++i) {
      Pair& p = pairs[i];
      const int n = p.a.size() / 2, m = p.b.size() / 2;
      if (MEAS == DTW) wants[i] = orc_dtw(p.a.data(), n, p.b.data(), m, 0);
      else if (MEAS == FRECHET) wants[i] = orc_frechet(p.a.data(), n, p.b.data(), m);
      else wants[i] = (float)orc_lcss(p.a.data(), n, p.b.data(), m, eps, -1) / (float)std::min(n, m);
      static const float kFactor[7] = {0.25f, 0.8f, 1.0f, 1.0f, 1.0001f, 3.0f, INFINITY};
      const int f = std::uniform_int_distribution<int>(0, 6)(rng);
      p.tau = f == 3 ? nextafterf(wants[i], 0.0f) : wants[i] * kFactor[f];
    }
    std::vector<int> dropped;
    // launch_dp_s: one entry per strip of the longest row trajectory, up to 256
    const int sfx = MEAS == DTW ? std::min(((rmax + W - 1) / W + 3) / 4 * 4, 256) : 0;
    std::vector<float> got = emulate<MEAS, W, U>(pairs, eps, cmax, rmax, nwaves, &steps, &dropped, sfx);
    for (int i = 0; i < npairs; ++i) {
      if (pairs[i].dead) continue;
      const Pair& p = pairs[i];
      const int n = p.a.size() / 2, m = p.b.size() / 2;
      const float want = wants[i];
      if (dropped[i]) {
        ++drops;
        if (!(want > p.tau)) {
          if (bad < 5) fprintf(stderr, "MEAS %d W %d U %d trial %d pair %d n %d m %d: dropped at value %.9g, tau %.9g\n", MEAS, W, U, t, i, n, m, want, p.tau);
          ++bad;
        }
        continue;
      }
      if (memcmp(&want, &got[i], 4) != 0) {
        if (bad < 5) fprintf(stderr, "MEAS %d W %d U %d trial %d pair %d n %d m %d: got %.9g want %.9g\n", MEAS, W, U, t, i, n, m, got[i], want);
        ++bad;
      }
    }
  }
  printf("MEAS %d W %2d U %d: %d trials, %d pairs dropped early, %d mismatches\n", MEAS, W, U, trials, drops, bad);
  return bad;
}

int main(int argc, char** argv) {
  std::mt19937 rng(12345);
  int bad = 0;
  const int T = argc > 1 ? std::atoi(argv[1]) : 60;
  g_late = argc > 2 && std::strcmp(argv[2], "late") == 0;
  g_nodisc = argc > 2 && std::strcmp(argv[2], "nodisc") == 0;
  g_sfx1 = argc > 2 && std::strcmp(argv[2], "sfx1") == 0;
  bad += run<DTW, 16, 1>(rng, T); bad += run<DTW, 16, 2>(rng, T); bad += run<DTW, 16, 4>(rng, T);
  bad += run<DTW, 32, 2>(rng, T); bad += run<DTW, 32, 4>(rng, T); bad += run<DTW, 64, 2>(rng, T);
  bad += run<DTW, 64, 4>(rng, T); bad += run<DTW, 32, 8>(rng, T); bad += run<DTW, 64, 8>(rng, T);
  bad += run<FRECHET, 16, 4>(rng, T); bad += run<FRECHET, 64, 8>(rng, T); bad += run<FRECHET, 32, 2>(rng, T);
  bad += run<LCSS, 16, 1>(rng, T); bad += run<LCSS, 32, 4>(rng, T); bad += run<LCSS, 64, 8>(rng, T);
  printf("suffix-bound drops: %ld\n", g_sfx_drops);
  printf(bad ? "FAIL\n" : "ALL OK\n");
  return bad != 0;
}
