/* Host check of cos_2pi (rmagine_amd/csrc/rmh_noise.hip), the cosine of the Box-Muller
 * step: the same operations in C, evaluated for every u = k 2^-24 the noise kernel can draw
 * (k = 0 .. 2^24 - 1), against libm's cos of the FP64 product 2 pi u (the oracle's formula)
 * and against cosl of 2 pi u in long double.  Exit status 1 when an error reaches 1e-15.
 * -DCOS_2PI_SWAP_QUADRANTS plants a mistake (quadrants 1 and 3 exchanged) that the check
 * must find. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>

static double cos_2pi(double u) {
  const double t = 4.0 * u;
  const double k = rint(t);
  const double x = (t - k) * 1.5707963267948966;
  const double x2 = x * x;
  double c = 1.0 / 20922789888000.0;
  c = fma(c, x2, -1.0 / 87178291200.0);
  c = fma(c, x2, 1.0 / 479001600.0);
  c = fma(c, x2, -1.0 / 3628800.0);
  c = fma(c, x2, 1.0 / 40320.0);
  c = fma(c, x2, -1.0 / 720.0);
  c = fma(c, x2, 1.0 / 24.0);
  c = fma(c, x2, -0.5);
  c = fma(c, x2, 1.0);
  double sn = 1.0 / 355687428096000.0;
  sn = fma(sn, x2, -1.0 / 1307674368000.0);
  sn = fma(sn, x2, 1.0 / 6227020800.0);
  sn = fma(sn, x2, -1.0 / 39916800.0);
  sn = fma(sn, x2, 1.0 / 362880.0);
  sn = fma(sn, x2, -1.0 / 5040.0);
  sn = fma(sn, x2, 1.0 / 120.0);
  sn = fma(sn, x2, -1.0 / 6.0);
  sn = fma(sn, x2, 1.0) * x;
  const int q = (int)k & 3;
  const double r = (q & 1) ? sn : c;
#ifdef COS_2PI_SWAP_QUADRANTS
  return (q == 3 || q == 2) ? -r : r;
#else
  return (q == 1 || q == 2) ? -r : r;
#endif
}

int main(void) {
  double max_libm = 0.0, max_exact = 0.0;
  uint32_t n = 0;
  for (uint32_t k = 0; k < (1u << 24); k++, n++) {
    const double u = (double)k * 0x1p-24;
    const double a = cos_2pi(u);
    const double e1 = fabs(a - cos(6.283185307179586 * u));
    const long double ref = cosl(2.0L * 3.14159265358979323846264338327950288L * (long double)u);
    const double e2 = (double)fabsl((long double)a - ref);
    if (!(e1 <= max_libm)) max_libm = e1;
    if (!(e2 <= max_exact)) max_exact = e2;
  }
  printf("%u %.3e %.3e\n", n, max_libm, max_exact);
  return (max_libm < 1e-15 && max_exact < 1e-15) ? 0 : 1;
}
