/* Sweep of the division used by the HIP kernels for shared divisors (exdiv in
 * shark_amd/csrc/physics.h): q = fma(fma(-b, a*y, a), y, a*y) with y = RN(1/b), carrying the
 * sign of a*y (so a zero dividend gives the zero of a/b's sign).  Compared with IEEE a/b on random operands over a wide exponent range, on divisors
 * with all-ones significands, and on zeros of both signs.  Prints the mismatch count. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t s = 0x9E3779B97F4A7C15ULL;
static uint64_t xr(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double rnd(int emin, int emax) {
  uint64_t m = xr() & ((1ULL << 52) - 1);
  int e = emin + (int)(xr() % (uint64_t)(emax - emin + 1));
  uint64_t bits = ((uint64_t)(e + 1023) << 52) | m;
  double d;
  memcpy(&d, &bits, 8);
  return (xr() & 1) ? -d : d;
}
static double exdiv(double a, double b, double y) {
  const double q0 = a * y;
  const double r = fma(-b, q0, a);
  const double q = fma(r, y, q0);
  return copysign(q, q0);
}
static int same(double x, double y) { return memcmp(&x, &y, 8) == 0; }

int main(int argc, char** argv) {
  long n = argc > 1 ? atol(argv[1]) : 10000000;
  long bad = 0;
  for (long t = 0; t < n; ++t) {
    double a = rnd(-300, 300), b = rnd(-300, 300);
    if ((t & 255) == 0) { uint64_t u; memcpy(&u, &b, 8); u |= (1ULL << 52) - 1; memcpy(&b, &u, 8); }
    if ((t & 255) == 1) { uint64_t u; memcpy(&u, &a, 8); u |= (1ULL << 52) - 1; memcpy(&a, &u, 8); }
    if ((t & 1023) == 2) a = (xr() & 1) ? 0.0 : -0.0;
    const double q = a / b;
    if (!isnormal(q) && q != 0.0) continue;   /* the method covers quotients in the normal range */
    if (!isnormal(1.0 / b)) continue;
    if (!same(exdiv(a, b, 1.0 / b), q)) {
      if (bad < 5) printf("mismatch a=%a b=%a\n", a, b);
      ++bad;
    }
  }
  printf("%ld\n", bad);
  return 0;
}
