/* Test infrastructure: an independent closest-hit check for the oracle.  Brute force over
 * every triangle with the Moller-Trumbore ray/triangle test ("Fast, minimum storage
 * ray-triangle intersection", JGT 1997) in FP64: two-sided, 0 <= t <= tmax, equal distances
 * to the lower face index.  A different algorithm (no watertight shear, no BVH) than the
 * oracle's exact predicate over its SAH BVH.  OpenMP over rays. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

void mt_closest(const float* V, const uint32_t* F, int64_t nf, const double* O, const double* D, int64_t n,
                double tmax, double* t_out, int64_t* f_out) {
  double* T = (double*)malloc(sizeof(double) * 9 * (size_t)(nf > 0 ? nf : 1));
  for (int64_t f = 0; f < nf; f++) {
    const float* a = V + 3 * (size_t)F[3 * f];
    const float* b = V + 3 * (size_t)F[3 * f + 1];
    const float* c = V + 3 * (size_t)F[3 * f + 2];
    for (int k = 0; k < 3; k++) {
      T[9 * f + k] = a[k];
      T[9 * f + 3 + k] = (double)b[k] - (double)a[k];
      T[9 * f + 6 + k] = (double)c[k] - (double)a[k];
    }
  }
#pragma omp parallel for schedule(dynamic, 32)
  for (int64_t i = 0; i < n; i++) {
    const double* o = O + 3 * i;
    const double* d = D + 3 * i;
    double best = INFINITY;
    int64_t bf = -1;
    for (int64_t f = 0; f < nf; f++) {
      const double* v0 = T + 9 * f;
      const double* e1 = v0 + 3;
      const double* e2 = v0 + 6;
      const double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
      const double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
      if (det == 0.0) continue;
      const double inv = 1.0 / det;
      const double s[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
      const double u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
      if (u < 0.0 || u > 1.0) continue;
      const double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
      const double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
      if (v < 0.0 || u + v > 1.0) continue;
      const double t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
      if (t < 0.0 || t > tmax) continue;
      if (t < best) {
        best = t;
        bf = f;
      }
    }
    t_out[i] = best;
    f_out[i] = bf;
  }
  free(T);
}
