/* Host check of the conservativeness of the ray kernels' FP32 slab test
 * (rmagine_amd/csrc/rmh_trace.hpp: SlabRay set-up in trace(), slab2): for random rays and
 * boxes that the exact ray meets at some t* in [0, tbf], the test computed in FP32 exactly as
 * the kernels do (reciprocal 1.0f / d with components below 2^-64 replaced by +-2^-64, the
 * per-ray margin P = 2^-21 (max|o| + range_max) + 1e-30, fmaf(plane, 1/d, -(o +- P) / d),
 * tmin = max(near values, 0), tmax = min(far values, tbf), hit iff tmin <= tmax) must never
 * cull the box, and the entry distance it computes must not exceed t* (the stack's pop test
 * relies on that).
 *
 * Boxes are built around a point of the exact ray: p = o + t* d in FP64, then each bound
 * rounded outward to FP32 past the FP64 error of p, with random extents that are often zero
 * (degenerate boxes, rays that touch a face, an edge or a corner).  Origins reach 1e4,
 * directions include tiny, zero and axis-aligned components, range_max reaches 1e4, and tbf
 * is either range_max or t* rounded up to FP32 (the bound after a hit at t*: a box holding
 * an equal-distance triangle must survive; the kernels' bound, the nearest float plus one
 * ulp, is at least that, so it culls no more).
 *
 * usage: slab_check [n]  -- prints the numbers of checked and culled boxes; exit status 0
 * when none is culled.  Built with the C compiler without -ffast-math: fmaf is the
 * single-rounding fma.  Built with -DSLAB_WITHOUT_RANGE_TERM the margin lacks its
 * 2^-21 range_max term (which stands in for a relative slack on tmax): that variant must
 * cull some of the boxes, which shows the check can fail. */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64(void) {
  /* splitmix64 */
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double uni(void) { return (double)(next_u64() >> 11) * 0x1p-53; }

static float down_f(double v) { /* largest float <= v */
  float f = (float)v;
  if ((double)f > v) f = nextafterf(f, -INFINITY);
  return f;
}
static float up_f(double v) { /* smallest float >= v */
  float f = (float)v;
  if ((double)f < v) f = nextafterf(f, INFINITY);
  return f;
}

/* a random float coordinate of magnitude up to `scale`, sometimes exactly representable
 * small integers or zero */
static float coord(double scale) {
  const uint64_t k = next_u64() % 8;
  if (k == 0) return 0.0f;
  if (k == 1) return (float)((int)(next_u64() % 41) - 20);
  return (float)((uni() * 2.0 - 1.0) * scale);
}

static float dir_comp(void) {
  const uint64_t k = next_u64() % 10;
  if (k == 0) return 0.0f;
  if (k == 1) return (next_u64() & 1 ? 1.0f : -1.0f) * 1e-30f;
  if (k == 2) return (next_u64() & 1 ? 1.0f : -1.0f) * 0x1p-70f;
  if (k == 3) return (next_u64() & 1 ? 1.0f : -1.0f) * 1e-7f;
  return (float)(uni() * 2.0 - 1.0);
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? atol(argv[1]) : 4000000;
  long checked = 0, failed = 0;
  for (long it = 0; it < n; it++) {
    const double oscale = (double[]){1.0, 100.0, 1000.0, 10000.0}[next_u64() % 4];
    const float o[3] = {coord(oscale), coord(oscale), coord(oscale)};
    float d[3];
    do {
      d[0] = dir_comp(); d[1] = dir_comp(); d[2] = dir_comp();
    } while (fabsf(d[0]) + fabsf(d[1]) + fabsf(d[2]) < 1e-3f);
    {
      /* unit length as the kernels' rays (rotated sensor directions) roughly are */
      const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      for (int k = 0; k < 3; k++) d[k] = d[k] / len;
    }
    const float range_max = (float[]){1.0f, 100.0f, 130.0f, 10000.0f}[next_u64() % 4];
    /* t* in [0, range_max], sometimes exactly 0 or range_max */
    double ts;
    const uint64_t tk = next_u64() % 8;
    ts = tk == 0 ? 0.0 : (tk == 1 ? (double)range_max : uni() * (double)range_max);
    const int after_hit = (next_u64() & 1) != 0;
    const float tbf = after_hit ? up_f(ts) : range_max;

    /* the box around p = o + t* d (exact reals lie within err of the FP64 values) */
    float lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
      const double p = (double)o[k] + ts * (double)d[k];
      const double err = fabs(p) * 0x1p-50 + 0x1p-1000;
      const uint64_t ek = next_u64() % 4;
      const double a = ek == 0 ? 0.0 : uni() * (ek == 1 ? 1e-4 : (ek == 2 ? 1.0 : 10.0));
      const uint64_t fk = next_u64() % 4;
      const double b = fk == 0 ? 0.0 : uni() * (fk == 1 ? 1e-4 : (fk == 2 ? 1.0 : 10.0));
      lo[k] = down_f(p - err - a);
      hi[k] = up_f(p + err + b);
    }

    /* the kernels' slab test (trace(): SlabRay set-up; slab2: fmas, near/far, clamps) */
    const float kMinDir = 0x1p-64f;
    float id[3], lx[3], hx[3];
    const float omax = fmaxf(fmaxf(fabsf(o[0]), fabsf(o[1])), fabsf(o[2]));
#ifdef SLAB_WITHOUT_RANGE_TERM
    const float P = omax * 0x1p-21f + 1e-30f;
#else
    const float P = (omax + range_max) * 0x1p-21f + 1e-30f;
#endif
    for (int k = 0; k < 3; k++) {
      id[k] = 1.0f / (fabsf(d[k]) > kMinDir ? d[k] : copysignf(kMinDir, d[k]));
      lx[k] = -((o[k] + P) * id[k]);
      hx[k] = -((o[k] - P) * id[k]);
    }
    float tmin = 0.0f, tmax = tbf;
    for (int k = 0; k < 3; k++) {
      const float v0 = fmaf(lo[k], id[k], lx[k]);
      const float v1 = fmaf(hi[k], id[k], hx[k]);
      tmin = fmaxf(tmin, fminf(v0, v1));
      tmax = fminf(tmax, fmaxf(v0, v1));
    }
    checked++;
    if (!(tmin <= tmax) || !((double)tmin <= ts)) {
      if (failed < 10)
        printf("FAILED: o=(%a,%a,%a) d=(%a,%a,%a) range_max=%a t*=%a tbf=%a box=[%a,%a]x[%a,%a]x[%a,%a] "
               "tmin=%a tmax=%a\n",
               o[0], o[1], o[2], d[0], d[1], d[2], range_max, ts, tbf, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2],
               tmin, tmax);
      failed++;
    }
  }
  printf("%ld %ld\n", checked, failed);
  return failed == 0 ? 0 : 1;
}
