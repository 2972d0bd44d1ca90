/* Check of the integer-count division of the multigrid kernels (k_mg.hip,
 * div_count): for a count c in 1..6 with r = RN(1/c),
 *   q0 = x * r,  e = fma(-q0, c, x),  q = fma(e, r, q0),
 * and q is replaced by the IEEE quotient x / c when q is subnormal.  The
 * claim: the result equals x / c bit for bit for every finite float x except
 * -0 (the kernels never divide -0: their x is a neighbour sum started at +0
 * minus h^2 b).  Also counted: how many x the unguarded q gets wrong (all of
 * them must have a subnormal q).
 *
 *   div_count_check          sampled (seconds): every 3rd x with
 *                            |x| < 2^-124 for c = 6 (every 41st otherwise),
 *                            every 127th mantissa of 12 exponents spread over
 *                            the range, 2^20 random patterns per c
 *   div_count_check --all    all 2^32 bit patterns (about a minute, OpenMP)
 *
 * Exit status 0 if no x is wrong.  The host is IEEE binary32 with
 * round-to-nearest and subnormals (x86-64 SSE), as the device is here. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float from_bits(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

/* returns 1 if the guarded quotient differs from x / c; *raw_bad is set if
 * the unguarded quotient differs while not subnormal (a hole in the guard) */
static int check_one(uint32_t u, float c, float r, int *raw_bad_unflagged) {
  const float x = from_bits(u);
  *raw_bad_unflagged = 0;
  if (!isfinite(x) || u == 0x80000000u) return 0;
  const float ieee = x / c;
  const float q0 = x * r;
  const float e = fmaf(-q0, c, x);
  float q = fmaf(e, r, q0);
  const int sub = fpclassify(q) == FP_SUBNORMAL;
  if (memcmp(&q, &ieee, 4) != 0 && !sub) *raw_bad_unflagged = 1;
  if (sub) q = x / c;
  return memcmp(&q, &ieee, 4) != 0;
}

int main(int argc, char **argv) {
  const int all = argc > 1 && strcmp(argv[1], "--all") == 0;
  uint64_t bad = 0, hole = 0, tested = 0;
  for (int k = 1; k <= 6; k++) {
    const volatile float cv = (float)k;
    const float c = cv, r = 1.0f / cv;
    if (all) {
      uint64_t b = 0, h = 0;
#pragma omp parallel for reduction(+ : b, h) schedule(static)
      for (int64_t hi = 0; hi < 65536; hi++)
        for (uint32_t lo = 0; lo < 65536; lo++) {
          int hb;
          b += check_one(((uint32_t)hi << 16) | lo, c, r, &hb);
          h += hb;
        }
      bad += b; hole += h; tested += 1ull << 32;
      continue;
    }
    /* |x| < 2^-124 (subnormals and the smallest normals, where c = 6 needs
     * the guard), both signs: every pattern for c = 6, every 5th otherwise */
    const uint32_t band_step = k == 6 ? 3u : 41u;
    for (uint32_t m = 0; m < (3u << 23); m += band_step)
      for (uint32_t s = 0; s < 2; s++) {
        int hb;
        bad += check_one(m | (s << 31), c, r, &hb);
        hole += hb;
        tested++;
      }
    /* every 16th mantissa of 12 exponents across the normal range */
    for (int ei = 0; ei < 12; ei++) {
      const uint32_t ex = 3u + (uint32_t)ei * 21u;
      for (uint32_t m = (uint32_t)ei; m < (1u << 23); m += 127u)
        for (uint32_t s = 0; s < 2; s++) {
          int hb;
          bad += check_one((s << 31) | (ex << 23) | m, c, r, &hb);
          hole += hb;
          tested++;
        }
    }
    /* random bit patterns (xorshift) */
    uint32_t st = 0x9e3779b9u * (uint32_t)k;
    for (int i = 0; i < (1 << 20); i++) {
      st ^= st << 13; st ^= st >> 17; st ^= st << 5;
      int hb;
      bad += check_one(st, c, r, &hb);
      hole += hb;
      tested++;
    }
  }
  printf("div_count: %llu inputs, %llu wrong, %llu unguarded errors with a normal quotient\n",
         (unsigned long long)tested, (unsigned long long)bad, (unsigned long long)hole);
  return bad != 0 || hole != 0;
}
