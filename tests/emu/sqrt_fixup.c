/* The DTW refine computes sqrt(x) as v_sqrt_f32(x) (within 1 ulp) followed by
 * a neighbour test with fma residuals (cudaruntime1_amd/csrc/dp.hip,
 * sqrt_cr): for x >= 2^-96, x = 0 and x = +inf the test must turn every
 * start within 1 ulp of the root into the correctly rounded sqrtf. This
 * checks that on the CPU (fmaf and sqrtf correctly rounded): every float in
 * [2^-96, 2^-93), [1, 4) and [2^126, 2^128), a stride over all others, and
 * the two specials. Exit status 0 when all match. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float fixup(float s, float x) {
  const float sm = u2f(f2u(s) - 1u), sp = u2f(f2u(s) + 1u);
  const float t = fmaf(-sm, s, x) <= 0.0f ? sm : s;
  return fmaf(-sp, s, x) > 0.0f ? sp : t;
}

static unsigned long long bad, n;
static void check(uint32_t u) {
  const float x = u2f(u), want = sqrtf(x);
  for (int e = -1; e <= 1; ++e) {
    if (f2u(fixup(u2f(f2u(want) + (uint32_t)e), x)) != f2u(want)) {
      if (bad < 5) printf("mismatch x=%a start %+d ulp\n", x, e);
      ++bad;
    }
    ++n;
  }
}

int main(void) {
  const uint32_t ranges[3][2] = {{0x0f800000u, 0x11000000u}, {0x3f800000u, 0x40800000u}, {0x7e800000u, 0x7f800000u}};
  for (int r = 0; r < 3; ++r)
    for (uint32_t u = ranges[r][0]; u < ranges[r][1]; ++u) check(u);
  for (uint32_t u = 0x0f800000u; u < 0x7f800000u; u += 97) check(u);
  if (f2u(fixup(0.0f, 0.0f)) != 0u || fixup(INFINITY, INFINITY) != INFINITY) {
    printf("mismatch at 0 or inf\n");
    ++bad;
  }
  printf("%llu starts checked, %llu mismatches\n", n, bad);
  return bad != 0;
}
