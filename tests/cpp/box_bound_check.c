/* Host check of the closest-point walk's box bound (rmagine_amd/csrc/rmh_closest.hip:
 * box_d2): for random queries and boxes, the FP32 value computed as the kernel computes it
 * -- per axis max(lo - q, q - hi, 0), the sum of squares, shrunk by 2^-20, infinity clamped
 * to FLT_MAX -- must never exceed the squared distance from the query to the box (taken
 * in FP64 and lowered by 2^-50 for its own rounding), or the walk could skip the box that
 * holds the nearest point.  Queries reach 1e4 (and 1e19 for the overflow clamp), boxes are
 * placed at gaps from 0 to 1e3, often with one float between plane and query.
 *
 * usage: box_bound_check [n]  -- prints the numbers of checked and violating cases; exit
 * status 0 when there is none.  Built with -DBOX_BOUND_WITHOUT_SHRINK the 2^-20 shrink is
 * left out: that variant must violate the bound in some cases, which shows the check can
 * fail.  Built without -ffast-math and with -ffp-contract=off, as the kernels are. */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static uint64_t rng_state = 0x2545F4914F6CDD1Dull;
static uint64_t next_u64(void) {
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double uni(void) { return (double)(next_u64() >> 11) * 0x1p-53; }

static float bound(float q[3], float lo[3], float hi[3]) {
  float d[3];
  for (int k = 0; k < 3; k++) d[k] = fmaxf(fmaxf(lo[k] - q[k], q[k] - hi[k]), 0.0f);
  float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
#ifndef BOX_BOUND_WITHOUT_SHRINK
  d2 = d2 * (1.0f - 0x1p-20f);
#endif
  return fminf(d2, FLT_MAX);
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? atol(argv[1]) : 4000000;
  long bad = 0;
  for (long it = 0; it < n; it++) {
    const double scale = (double[]){1.0, 100.0, 1e4, 1e19}[next_u64() % 4];
    float q[3], lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
      q[k] = (float)((uni() * 2.0 - 1.0) * scale);
      const uint64_t mode = next_u64() % 4;
      double gap = mode == 0 ? 0.0 : (mode == 1 ? uni() * 1e-3 : (mode == 2 ? uni() * 10.0 : uni() * 1e3));
      const double ext = uni() * 5.0;
      float a, b;
      if (next_u64() & 1) { a = (float)((double)q[k] + gap); b = (float)((double)a + ext); }
      else { b = (float)((double)q[k] - gap); a = (float)((double)b - ext); }
      if (mode == 1 && (next_u64() & 1)) a = nextafterf(q[k], INFINITY), b = nextafterf(a, INFINITY) + (float)ext;
      lo[k] = fminf(a, b);
      hi[k] = fmaxf(a, b);
    }
    double d2 = 0.0;
    for (int k = 0; k < 3; k++) {
      const double d = fmax(fmax((double)lo[k] - (double)q[k], (double)q[k] - (double)hi[k]), 0.0);
      d2 += d * d;
    }
    const double lowered = d2 * (1.0 - 0x1p-50);
    const float b = bound(q, lo, hi);
    if ((double)b > lowered && (double)b > 0.0) {
      if (bad < 5) printf("VIOLATION: bound %a > squared distance %a\n", (double)b, d2);
      bad++;
    }
  }
  printf("%ld %ld\n", n, bad);
  return bad == 0 ? 0 : 1;
}
