/* Checks the division-by-constant sequence used in the device PPM code
 * against IEEE division: q = x * RN(1/c); rho = fma(-q, c, x);
 * result = copysign(fma(rho, RN(1/c), q), x), for c = 3, 6, 12.
 * Prints "<c> <samples> <mismatches>" per constant; mismatches must be 0
 * for every finite x whose quotient is normal or zero. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t next(void) {  /* splitmix64 */
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double from_bits(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

static double div_c(double x, double c, double r) {
  const double q = x * r;
  const double rho = fma(-q, c, x);
  return copysign(fma(rho, r, q), x);
}

int main(int argc, char** argv) {
  long n = argc > 1 ? atol(argv[1]) : 10000000;
  const double cs[3] = {3.0, 6.0, 12.0};
  int bad_total = 0;
  for (int k = 0; k < 3; ++k) {
    const double c = cs[k], r = 1.0 / c;
    long bad = 0, tried = 0;
    for (long i = 0; i < n; ++i) {
      uint64_t b = next();
      /* exponents across the whole normal range, both signs */
      b = (b & 0x800FFFFFFFFFFFFFull) | ((uint64_t)(1 + next() % 2044) << 52);
      double x = from_bits(b);
      double want = x / c;
      if (fabs(want) < 0x1p-1022 && want != 0.0) continue;  /* subnormal quotient */
      ++tried;
      if (bits(div_c(x, c, r)) != bits(want)) ++bad;
    }
    /* significands with every residue mod 3 near the top and bottom of a binade */
    for (long m = 0; m < 3000000; ++m) {
      uint64_t sig = (m % 2 ? (1ull << 52) + m : (1ull << 53) - 1 - m) & ((1ull << 52) - 1);
      for (int e = 1; e < 2047; e += 97) {
        double x = from_bits(((uint64_t)e << 52) | sig);
        double want = x / c;
        if (fabs(want) < 0x1p-1022 && want != 0.0) continue;
        ++tried;
        if (bits(div_c(x, c, r)) != bits(want)) ++bad;
      }
    }
    const double zeros[2] = {0.0, -0.0};
    for (int z = 0; z < 2; ++z) {
      ++tried;
      if (bits(div_c(zeros[z], c, r)) != bits(zeros[z] / c)) ++bad;
    }
    printf("%g %ld %ld\n", c, tried, bad);
    bad_total += bad != 0;
  }
  return bad_total;
}
