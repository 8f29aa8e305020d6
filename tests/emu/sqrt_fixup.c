/* The DTW refine computes sqrt(x) as v_sqrt_f32(x) (within 1 ulp) followed by
 * a neighbour decision with fma residuals (cudaruntime1_amd/csrc/ts_internal.h,
 * sqrt_cr): with s the start, sm = s - 1 ulp (0 for s = 0) and sp = s + 1 ulp,
 * the result is sm + clamp01(bits(x - sm*s)) + clamp01(bits(x - sp*s)) in
 * ulps, clamp01 being v_med3_i32(v, 0, 1) on the residual's bit pattern. For
 * x = 0 and x in [2^-96, 2^127) every start within 1 ulp of the root must
 * give the correctly rounded sqrtf. This checks that on the CPU (fmaf and
 * sqrtf correctly rounded): every float in [2^-96, 2^-93), [1, 4) and
 * [2^125, 2^127), a stride over all others, and 0; with the argument "all",
 * every float in [2^-96, 2^127) (build with -fopenmp to use every core).
 * Exit status 0 when all match. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static uint32_t clamp01(float f) {
  int32_t v;
  memcpy(&v, &f, 4);
  return v < 0 ? 0u : (v > 1 ? 1u : (uint32_t)v);
}

static float fixup(float s, float x) {
  const uint32_t sb = f2u(s), smb = sb == 0 ? 0 : sb - 1u;
  const float fm = fmaf(-u2f(smb), s, x), fp = fmaf(-u2f(sb + 1u), s, x);
  return u2f(smb + clamp01(fm) + clamp01(fp));
}

static unsigned long long check(uint32_t u) {
  const float x = u2f(u), want = sqrtf(x);
  unsigned long long bad = 0;
  for (int e = -1; e <= 1; ++e) {
    if (f2u(fixup(u2f(f2u(want) + (uint32_t)e), x)) != f2u(want)) {
      if (bad == 0) printf("mismatch x=%a start %+d ulp\n", x, e);
      ++bad;
    }
  }
  return bad;
}

int main(int argc, char** argv) {
  unsigned long long bad = 0, n = 0;
  if (argc > 1 && strcmp(argv[1], "all") == 0) {
#pragma omp parallel for reduction(+ : bad, n) schedule(dynamic, 1)
    for (uint32_t hi = 0x0f8; hi < 0x7f0; ++hi)
      for (uint32_t lo = 0; lo < 0x100000; ++lo) {
        bad += check((hi << 20) | lo);
        n += 3;
      }
  } else {
    const uint32_t ranges[3][2] = {{0x0f800000u, 0x11000000u}, {0x3f800000u, 0x40800000u}, {0x7e000000u, 0x7f000000u}};
    for (int r = 0; r < 3; ++r)
      for (uint32_t u = ranges[r][0]; u < ranges[r][1]; ++u) {
        bad += check(u);
        n += 3;
      }
    for (uint32_t u = 0x0f800000u; u < 0x7f000000u; u += 97) {
      bad += check(u);
      n += 3;
    }
  }
  if (f2u(fixup(0.0f, 0.0f)) != 0u) {
    printf("mismatch at 0\n");
    ++bad;
  }
  printf("%llu starts checked, %llu mismatches\n", n, bad);
  return bad != 0;
}
