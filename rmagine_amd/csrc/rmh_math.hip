// Math layer of librmh: the MI355X counterpart of rmagine_cuda's math
// (src/rmagine_cuda/src/math/memory_math.cu, math_batched.cu, statistics.cu, linalg.cu).
//
//   rmh_transform_mul / _inv / _points / rmh_quat_normalize
//       element-wise pose algebra with N x N, N x 1 and 1 x N broadcasting
//       (memory_math.cu:16-160 multNxN/multNx1/mult1xN/invert, :402-413 normalizeInplace)
//   rmh_vec3_moments
//       segmented sum / mean / cross second moment of vector pairs with an optional
//       mask (memory_math.cu:415-469 cov_kernel, :1403-1549 sum/mean/cov;
//       math_batched.cu:30-546 sumBatched/covBatched), FP64 accumulation
//   rmh_statistics
//       point-to-point / point-to-plane correspondence statistics, segmented (one
//       segment per pose hypothesis), FP64 (statistics.cu:58-160 p2p, :190-276 p2l;
//       CPU statistics.cpp:20-126)
//   rmh_statistics_objectwise_p2l
//       point-to-plane statistics per object over a pixel window of image-shaped scans,
//       with per-object pre-transforms and constraints (statistics.cu:278-371,422-470)
//   rmh_linear_system_p2l
//       segmented FP64 point-to-plane Gauss-Newton system J^T w J, J^T w r
//       (optimization.h:113-137,170-200)
//   rmh_transform_pack_matrix / _quat
//       (R, t) -> Transform (memory_math.cu:313-339,1258-1279)
//   rmh_umeyama
//       per-segment FP64 3x3 SVD (one-sided Jacobi) and the Umeyama rotation with the
//       reflection fix (optimization.cpp:12-55, linalg.cu:1150-1190)
//
// Reductions are wave64 shuffles, then one LDS step across the waves of a block, then
// (for segments split over several blocks) a second pass over the per-block partials in a
// fixed order, so results are deterministic run to run.
#include "rmh_internal.hpp"

#include <cmath>
#include <mutex>
#include <vector>

namespace {

using namespace rmh;

constexpr int kBlk = 256;
constexpr int kWaves = kBlk / 64;

inline uint32_t grid1(uint64_t n) { return static_cast<uint32_t>((n + kBlk - 1) / kBlk); }

// ---- pose algebra -----------------------------------------------------------------------

__device__ __forceinline__ rmh_transform store_tf(const Tf& T, uint32_t stamp) {
  rmh_transform r;
  r.R.x = T.R.x; r.R.y = T.R.y; r.R.z = T.R.z; r.R.w = T.R.w;
  r.t.x = T.t.x; r.t.y = T.t.y; r.t.z = T.t.z;
  r.stamp = stamp;
  return r;
}

__global__ void k_tf_mul(const rmh_transform* A, uint32_t nA, const rmh_transform* B, uint32_t nB, rmh_transform* C,
                         uint32_t n) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  const rmh_transform a = A[nA == 1 ? 0 : i];
  const rmh_transform b = B[nB == 1 ? 0 : i];
  C[i] = store_tf(tmul(load_tf(a), load_tf(b)), a.stamp);
}

__global__ void k_tf_inv(const rmh_transform* A, rmh_transform* C, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  const rmh_transform a = A[i];
  const Tf T = load_tf(a);
  Tf R;
  R.R = qconj(T.R);                 // Transform.tcc:24-31
  const V3 t = qrot(R.R, T.t);
  R.t = v3(-t.x, -t.y, -t.z);
  C[i] = store_tf(R, a.stamp);
}

__global__ void k_tf_points(const rmh_transform* Tr, uint32_t nT, const rmh_vec3* p, rmh_vec3* out, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  const Tf T = load_tf(Tr[nT == 1 ? 0 : i]);
  const rmh_vec3 a = p[i];
  const V3 r = tapply(T, v3(a.x, a.y, a.z));
  rmh_vec3 o; o.x = r.x; o.y = r.y; o.z = r.z;
  out[i] = o;
}

__global__ void k_set_identity(rmh_transform* T, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  rmh_transform r{};
  r.R.w = 1.0f;
  T[i] = r;
}

__global__ void k_quat_normalize(rmh_transform* T, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  rmh_quat q = T[i].R;
  const float l = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);  // Quaternion::l2norm
  q.x /= l; q.y /= l; q.z /= l; q.w /= l;
  T[i].R = q;
}

// Matrix3x3 -> Quaternion in FP32 with the reference's operation order
// (src/rmagine_core/include/rmagine/math/types/MatrixOps.tcc:1492-1529).  M is column-major,
// at(r, c) = M[c * 3 + r].  The reference writes `sqrtf(tr + 1.0)` and
// `sqrtf(1.0 + a - b - c)` with a double literal: the sum is formed in double and rounded to
// float for sqrtf, which is what the casts below do.  No normalisation.
__device__ __forceinline__ rmh_quat mat_to_quat_f32(const float* M) {
  const float a00 = M[0], a10 = M[1], a20 = M[2];
  const float a01 = M[3], a11 = M[4], a21 = M[5];
  const float a02 = M[6], a12 = M[7], a22 = M[8];
  float tr = 0.0f;  // MatrixOps::trace: res += at(i, i)
  tr += a00;
  tr += a11;
  tr += a22;
  rmh_quat q;
  if (tr > 0.0f) {
    const float S = sqrtf(static_cast<float>(static_cast<double>(tr) + 1.0)) * 2.0f;
    q.w = 0.25f * S;
    q.x = (a21 - a12) / S;
    q.y = (a02 - a20) / S;
    q.z = (a10 - a01) / S;
  } else if (a00 > a11 && a00 > a22) {
    const float S = sqrtf(static_cast<float>(1.0 + static_cast<double>(a00) - a11 - a22)) * 2.0f;
    q.w = (a21 - a12) / S;
    q.x = 0.25f * S;
    q.y = (a01 + a10) / S;
    q.z = (a02 + a20) / S;
  } else if (a11 > a22) {
    const float S = sqrtf(static_cast<float>(1.0 + static_cast<double>(a11) - a00 - a22)) * 2.0f;
    q.w = (a02 - a20) / S;
    q.x = (a01 + a10) / S;
    q.y = 0.25f * S;
    q.z = (a12 + a21) / S;
  } else {
    const float S = sqrtf(static_cast<float>(1.0 + static_cast<double>(a22) - a00 - a11)) * 2.0f;
    q.w = (a10 - a01) / S;
    q.x = (a02 + a20) / S;
    q.y = (a12 + a21) / S;
    q.z = 0.25f * S;
  }
  return q;
}

// pack (src/rmagine_cuda/src/math/memory_math.cu:313-339): T.R from a rotation matrix or a
// quaternion, T.t = t; the stamp of T is left as it is.
__global__ void k_tf_pack_mat(const float* Rm, const rmh_vec3* t, rmh_transform* T, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  float M[9];
  for (int k = 0; k < 9; k++) M[k] = Rm[9 * static_cast<size_t>(i) + k];
  T[i].R = mat_to_quat_f32(M);
  T[i].t = t[i];
}

__global__ void k_tf_pack_quat(const rmh_quat* q, const rmh_vec3* t, rmh_transform* T, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  T[i].R = q[i];
  T[i].t = t[i];
}

// ---- reductions -------------------------------------------------------------------------

template <int N>
struct Acc {
  double v[N];
};

template <int N>
__device__ __forceinline__ void wave_sum(Acc<N>& a) {
#pragma unroll
  for (int k = 0; k < N; k++) {
    double x = a.v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    a.v[k] = x;
  }
}

// block-wide sum, result valid in thread 0
template <int N>
__device__ __forceinline__ void block_sum(Acc<N>& a, double (*lds)[N]) {
  wave_sum(a);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < N; k++) lds[w][k] = a.v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < N; k++) {
      double s = lds[0][k];
      for (int j = 1; j < kWaves; j++) s += lds[j][k];
      a.v[k] = s;
    }
  }
}

// moments: [0] count, [1..3] sum a, [4..6] sum b, [7..15] sum b_r a_c (column-major, c*3+r)
constexpr int kMom = 16;

__device__ __forceinline__ void acc_pair(Acc<kMom>& m, double ax, double ay, double az, double bx, double by, double bz) {
  m.v[0] += 1.0;
  m.v[1] += ax; m.v[2] += ay; m.v[3] += az;
  m.v[4] += bx; m.v[5] += by; m.v[6] += bz;
  m.v[7] += bx * ax;  m.v[8] += by * ax;  m.v[9] += bz * ax;
  m.v[10] += bx * ay; m.v[11] += by * ay; m.v[12] += bz * ay;
  m.v[13] += bx * az; m.v[14] += by * az; m.v[15] += bz * az;
}

__device__ __forceinline__ void zero(Acc<kMom>& m) {
#pragma unroll
  for (int k = 0; k < kMom; k++) m.v[k] = 0.0;
}

// ---- segmented vector moments (sum / mean / cov) ---------------------------------------

struct MomArgs {
  const rmh_vec3* a;
  const rmh_vec3* b;        // NULL: b = a
  const uint8_t* mask;      // NULL: all
  uint64_t seg;             // elements per segment
  uint32_t nseg;
  uint32_t bps;             // blocks per segment
  double* partial;          // bps > 1: nseg*bps*kMom
  double* sum_a;            // 3 per segment (nullable)
  double* mean_a;           // 3 per segment (nullable)
  double* cov;              // 9 per segment: (1/n) sum b a^T, column-major (nullable)
  uint32_t* count;          // nullable
};

__device__ void finalize_moments(const MomArgs& a, uint32_t s, const Acc<kMom>& m) {
  const double n = m.v[0];
  if (a.count) a.count[s] = static_cast<uint32_t>(n);
  if (a.sum_a)
    for (int k = 0; k < 3; k++) a.sum_a[3 * s + k] = m.v[1 + k];
  if (a.mean_a)
    for (int k = 0; k < 3; k++) a.mean_a[3 * s + k] = n > 0 ? m.v[1 + k] / n : 0.0;
  if (a.cov)
    for (int k = 0; k < 9; k++) a.cov[9 * s + k] = n > 0 ? m.v[7 + k] / n : 0.0;
}

__global__ __launch_bounds__(kBlk) void k_moments(const MomArgs a) {
  __shared__ double lds[kWaves][kMom];
  const uint32_t s = blockIdx.y, bb = blockIdx.x;
  Acc<kMom> m;
  zero(m);
  const uint64_t base = static_cast<uint64_t>(s) * a.seg;
  const uint64_t stride = static_cast<uint64_t>(a.bps) * kBlk;
  for (uint64_t i = static_cast<uint64_t>(bb) * kBlk + threadIdx.x; i < a.seg; i += stride) {
    const uint64_t g = base + i;
    if (a.mask && a.mask[g] == 0) continue;
    const rmh_vec3 x = a.a[g];
    const rmh_vec3 y = a.b ? a.b[g] : x;
    acc_pair(m, x.x, x.y, x.z, y.x, y.y, y.z);
  }
  block_sum(m, lds);
  if (threadIdx.x != 0) return;
  if (a.bps == 1) finalize_moments(a, s, m);
  else
    for (int k = 0; k < kMom; k++) a.partial[(static_cast<uint64_t>(s) * a.bps + bb) * kMom + k] = m.v[k];
}

__global__ void k_moments_combine(const MomArgs a) {
  const uint32_t s = blockIdx.x * kBlk + threadIdx.x;
  if (s >= a.nseg) return;
  Acc<kMom> m;
  zero(m);
  for (uint32_t b = 0; b < a.bps; b++)
    for (int k = 0; k < kMom; k++) m.v[k] += a.partial[(static_cast<uint64_t>(s) * a.bps + b) * kMom + k];
  finalize_moments(a, s, m);
}

// ---- correspondence statistics ------------------------------------------------------------

struct StatArgs {
  int kind;                       // 0 p2p, 1 p2l
  const rmh_transform* pre;
  uint32_t npre;
  rmh_cloud d, m;
  uint64_t seg;
  uint32_t nseg, bps;
  rmh_umeyama_constraints prm;
  double* partial;
  rmh_cross_stats* out;
};

struct M3d {
  double m[9];  // row-major rotation of q p q^-1 (valid for any q, |q|^2 scaled)
};

__device__ __forceinline__ double uniform_f64(double v) {
  const uint64_t u = __double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u >> 32));
  return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
}

__device__ __forceinline__ M3d quat_matrix(const rmh_quat& qf) {
  const double x = qf.x, y = qf.y, z = qf.z, w = qf.w;
  M3d r;
  r.m[0] = w * w + x * x - y * y - z * z; r.m[1] = 2.0 * (x * y - w * z);          r.m[2] = 2.0 * (x * z + w * y);
  r.m[3] = 2.0 * (x * y + w * z);         r.m[4] = w * w - x * x + y * y - z * z;  r.m[5] = 2.0 * (y * z - w * x);
  r.m[6] = 2.0 * (x * z - w * y);         r.m[7] = 2.0 * (y * z + w * x);          r.m[8] = w * w - x * x - y * y + z * z;
  return r;
}

__device__ void finalize_stats(rmh_cross_stats* out, const Acc<kMom>& m, const double c[3]) {
  rmh_cross_stats r;
  const double n = m.v[0];
  r.n_meas = static_cast<uint32_t>(n);
  r._pad = 0;
  if (n > 0) {
    double dm[3], mm[3];
    for (int k = 0; k < 3; k++) { dm[k] = m.v[1 + k] / n; mm[k] = m.v[4 + k] / n; }
    for (int col = 0; col < 3; col++)
      for (int row = 0; row < 3; row++)
        r.covariance[col * 3 + row] = m.v[7 + col * 3 + row] / n - mm[row] * dm[col];
    for (int k = 0; k < 3; k++) { r.dataset_mean[k] = dm[k] + c[k]; r.model_mean[k] = mm[k] + c[k]; }
  } else {
    for (int k = 0; k < 3; k++) { r.dataset_mean[k] = 0.0; r.model_mean[k] = 0.0; }
    for (int k = 0; k < 9; k++) r.covariance[k] = 0.0;
  }
  *out = r;
}

// IDX = uint32_t when every segment's offsets fit 32 bits (the usual case): the segment's
// arrays are then addressed as block-uniform base pointer + 32-bit lane offset; uint64_t
// otherwise.  The kernel streams its inputs once, so resident waves are what hides the load
// latency: it is held to 80 VGPRs for six waves per SIMD (a few bytes per lane spill; on
// 2,000 x 32,768 point-to-plane correspondences 0.490 ms at five waves, 0.455 ms at six).
template <typename IDX>
__global__ __launch_bounds__(kBlk) __attribute__((amdgpu_waves_per_eu(6))) void k_stats(const StatArgs a) {
  __shared__ double lds[kWaves][kMom];
  const uint32_t s = blockIdx.y, bb = blockIdx.x;
  const rmh_transform P = a.pre[a.npre == 1 ? 0 : s];
  M3d R = quat_matrix(P.R);
  // block-uniform: held in scalar registers, used as the scalar operand of the FP64 FMAs
  for (int k = 0; k < 9; k++) R.m[k] = uniform_f64(R.m[k]);
  // moments are taken relative to the pre-transform's translation c = t (keeps the
  // cancellation in sum(m d^T)/n - mean_m mean_d^T small); relative to c the transformed
  // dataset point is R d and the model point is m - c
  const double c[3] = {uniform_f64((double)P.t.x), uniform_f64((double)P.t.y), uniform_f64((double)P.t.z)};
  Acc<kMom> m;
  zero(m);
  const uint64_t db = static_cast<uint64_t>(s) * a.d.segment_stride;
  const uint64_t mb = static_cast<uint64_t>(s) * a.m.segment_stride;
  const rmh_vec3* __restrict__ dpts = a.d.points + db;
  const rmh_vec3* __restrict__ mpts = a.m.points + mb;
  const rmh_vec3* __restrict__ mnrm = a.m.normals ? a.m.normals + mb : nullptr;
  const uint8_t* __restrict__ dmask = a.d.mask ? a.d.mask + db : nullptr;
  const uint8_t* __restrict__ mmask = a.m.mask ? a.m.mask + mb : nullptr;
  const uint32_t* __restrict__ dids = a.d.ids ? a.d.ids + db : nullptr;
  const uint32_t* __restrict__ mids = a.m.ids ? a.m.ids + mb : nullptr;
  const IDX stride = static_cast<IDX>(a.bps) * kBlk;
  const IDX seg = static_cast<IDX>(a.seg);
  const double maxd = uniform_f64(a.prm.max_dist);
  for (IDX i = static_cast<IDX>(bb) * kBlk + threadIdx.x; i < seg; i += stride) {
    // statistics.cu:230-235 / statistics.cpp:99-103: mask and id filters
    if (dmask && dmask[i] == 0) continue;
    if (mmask && mmask[i] == 0) continue;
    if (dids && dids[i] != a.prm.dataset_id) continue;
    if (mids && mids[i] != a.prm.model_id) continue;
    const rmh_vec3 dp = dpts[i];
    const rmh_vec3 mp = mpts[i];
    // Di - c = R * dataset point (the pre-transform's rotation, FP64)
    const double dx = R.m[0] * dp.x + R.m[1] * dp.y + R.m[2] * dp.z;
    const double dy = R.m[3] * dp.x + R.m[4] * dp.y + R.m[5] * dp.z;
    const double dz = R.m[6] * dp.x + R.m[7] * dp.y + R.m[8] * dp.z;
    double mx = (double)mp.x - c[0], my = (double)mp.y - c[1], mz = (double)mp.z - c[2];
    if (a.kind == 1) {
      // point to plane: project Di onto the model plane (statistics.cpp:105-119)
      const rmh_vec3 nn = mnrm[i];
      const double spd = (mx - dx) * nn.x + (my - dy) * nn.y + (mz - dz) * nn.z;
      if (!(fabs(spd) < maxd)) continue;
      mx = dx + nn.x * spd;
      my = dy + nn.y * spd;
      mz = dz + nn.z * spd;
    } else {
      // point to point (statistics.cpp:40-50)
      const double ex = mx - dx, ey = my - dy, ez = mz - dz;
      if (!(sqrt(ex * ex + ey * ey + ez * ez) < maxd)) continue;
    }
    acc_pair(m, dx, dy, dz, mx, my, mz);
  }
  block_sum(m, lds);
  if (threadIdx.x != 0) return;
  if (a.bps == 1) finalize_stats(&a.out[s], m, c);
  else
    for (int k = 0; k < kMom; k++) a.partial[(static_cast<uint64_t>(s) * a.bps + bb) * kMom + k] = m.v[k];
}

__global__ void k_stats_combine(const StatArgs a) {
  const uint32_t s = blockIdx.x * kBlk + threadIdx.x;
  if (s >= a.nseg) return;
  const rmh_transform P = a.pre[a.npre == 1 ? 0 : s];
  const double c[3] = {(double)P.t.x, (double)P.t.y, (double)P.t.z};
  Acc<kMom> m;
  zero(m);
  for (uint32_t b = 0; b < a.bps; b++)
    for (int k = 0; k < kMom; k++) m.v[k] += a.partial[(static_cast<uint64_t>(s) * a.bps + b) * kMom + k];
  finalize_stats(&a.out[s], m, c);
}

// ---- objectwise point-to-plane statistics ---------------------------------------------------

struct ObjArgs {
  const rmh_transform* pre;             // one per object
  rmh_cloud d, m;                       // image-shaped scans (segment_stride unused)
  uint32_t width;                       // row stride of the scans
  uint32_t n_points;                    // elements of the scans (the reference's bound N)
  const rmh_umeyama_constraints* prm;   // one per object
  const rmh_aabb* boxes;                // one per object: pixel window (x = column, y = row)
  uint32_t window_only;
  rmh_cross_stats* out;
};

// float -> unsigned as the GPU's saturating conversion gives it (the reference assigns the
// float box corners to unsigned ints): NaN and values <= 0 give 0, values >= 2^32 give
// 2^32 - 1, others are truncated toward zero.
__device__ __forceinline__ uint32_t f2u_sat(float f) {
  if (!(f > 0.0f)) return 0u;
  if (f >= 4294967296.0f) return 0xFFFFFFFFu;
  return static_cast<uint32_t>(f);
}

// statistics_objectwise_p2l_kernel (src/rmagine_cuda/src/math/statistics.cu:278-371): one
// workgroup per object.  The element set is the reference kernel's: window element k
// (row-major inside the box, k < bw * bh) is scan element start + (k / bw) * width + k % bw,
// kept while below n_points.  The reference hands ceil(bw * bh / 512) consecutive window
// elements to each of its 512 threads without bounding k by bw * bh, so its k runs up to
// bw * bh rounded up to a multiple of 512, into the rows below the window; that is the
// default here, window_only = 1 stops at bw * bh.  Each element then passes the mask and id
// filters of the object's constraints and the plane-distance test exactly as in k_stats
// (point-to-plane), in FP64.
__global__ __launch_bounds__(kBlk) void k_stats_objectwise(const ObjArgs a) {
  __shared__ double lds[kWaves][kMom];
  const uint32_t o = blockIdx.x;
  const rmh_aabb bb = a.boxes[o];
  const uint64_t c0 = f2u_sat(bb.min.x), r0 = f2u_sat(bb.min.y);
  const uint64_t c1 = f2u_sat(bb.max.x), r1 = f2u_sat(bb.max.y);
  // an inverted box is empty (the reference's unsigned differences would wrap around)
  const uint64_t bw = c1 > c0 ? c1 - c0 : 0, bh = r1 > r0 ? r1 - r0 : 0;
  const uint64_t n_win = bw * bh;
  uint64_t n_it = a.window_only ? n_win : (n_win + 511) / 512 * 512;
  const uint64_t start = r0 * a.width + c0;
  // rows from row_lim on start at or beyond n_points and contribute nothing
  const uint64_t row_lim = start >= a.n_points ? 0 : (a.n_points - start + a.width - 1) / a.width;
  if (n_it > row_lim * bw) n_it = row_lim * bw;

  const rmh_transform P = a.pre[o];
  M3d R = quat_matrix(P.R);
  for (int k = 0; k < 9; k++) R.m[k] = uniform_f64(R.m[k]);
  const double c[3] = {uniform_f64((double)P.t.x), uniform_f64((double)P.t.y), uniform_f64((double)P.t.z)};
  const rmh_umeyama_constraints prm = a.prm[o];
  const double maxd = uniform_f64(prm.max_dist);
  Acc<kMom> m;
  zero(m);
  for (uint64_t k = threadIdx.x; k < n_it; k += kBlk) {
    const uint64_t row = k / bw, col = k - row * bw;
    const uint64_t i = start + row * a.width + col;
    if (i >= a.n_points) continue;
    if (a.d.mask && a.d.mask[i] == 0) continue;
    if (a.m.mask && a.m.mask[i] == 0) continue;
    if (a.d.ids && a.d.ids[i] != prm.dataset_id) continue;
    if (a.m.ids && a.m.ids[i] != prm.model_id) continue;
    const rmh_vec3 dp = a.d.points[i];
    const rmh_vec3 mp = a.m.points[i];
    const rmh_vec3 nn = a.m.normals[i];
    const double dx = R.m[0] * dp.x + R.m[1] * dp.y + R.m[2] * dp.z;
    const double dy = R.m[3] * dp.x + R.m[4] * dp.y + R.m[5] * dp.z;
    const double dz = R.m[6] * dp.x + R.m[7] * dp.y + R.m[8] * dp.z;
    const double mx = (double)mp.x - c[0], my = (double)mp.y - c[1], mz = (double)mp.z - c[2];
    const double spd = (mx - dx) * nn.x + (my - dy) * nn.y + (mz - dz) * nn.z;
    if (!(fabs(spd) < maxd)) continue;
    acc_pair(m, dx, dy, dz, dx + nn.x * spd, dy + nn.y * spd, dz + nn.z * spd);
  }
  block_sum(m, lds);
  if (threadIdx.x == 0) finalize_stats(&a.out[o], m, c);
}

// ---- point-to-plane Gauss-Newton system -------------------------------------------------------

// accumulators: [0] count, [1..6] J^T w r, [7..27] upper triangle of J^T w J (row-major),
// [28] sum w r^2
constexpr int kLin = 29;

struct LinArgs {
  const rmh_transform* pre;
  uint32_t npre;
  rmh_cloud d, m;
  uint64_t seg;
  uint32_t nseg, bps;
  rmh_umeyama_constraints prm;
  double kscale;                  // GM kernel scale; <= 0: unit weights
  double* partial;
  rmh_p2l_system* out;
};

__device__ void finalize_system(rmh_p2l_system* out, const Acc<kLin>& m) {
  rmh_p2l_system r;
  r.n_meas = static_cast<uint32_t>(m.v[0]);
  r._pad = 0;
  for (int k = 0; k < 6; k++) r.JTwr[k] = m.v[1 + k];
  int u = 7;
  for (int i = 0; i < 6; i++)
    for (int j = i; j < 6; j++, u++) {
      r.JTwJ[i * 6 + j] = m.v[u];
      r.JTwJ[j * 6 + i] = m.v[u];
    }
  r.wr2 = m.v[28];
  *out = r;
}

// build_linear_system_p2l (src/rmagine_core/include/rmagine/math/optimization.h:170-200) with
// jacobian_and_residual_p2l (:113-137) per correspondence, segmented and in FP64: Pd = pre *
// dataset point, r = (Pd - Pm) . Nm, J = [Pd x Nm, Nm] (d r / d(omega, tau) for the left
// perturbation Pd -> Exp(omega) Pd + tau), w = GM_weight(s, r) = s^2 / (s + r)^2 (the
// reference passes the signed residual to GM_weight), and J^T w J, J^T w r, sum w r^2.  The
// mask, id and |r| < max_dist filters of the statistics apply (with no mask and an infinite
// max_dist every correspondence counts, as in the reference's loop).
__global__ __launch_bounds__(kBlk) void k_p2l_system(const LinArgs a) {
  __shared__ double lds[kWaves][kLin];
  const uint32_t s = blockIdx.y, bb = blockIdx.x;
  const rmh_transform P = a.pre[a.npre == 1 ? 0 : s];
  M3d R = quat_matrix(P.R);
  for (int k = 0; k < 9; k++) R.m[k] = uniform_f64(R.m[k]);
  const double t[3] = {uniform_f64((double)P.t.x), uniform_f64((double)P.t.y), uniform_f64((double)P.t.z)};
  Acc<kLin> m;
#pragma unroll
  for (int k = 0; k < kLin; k++) m.v[k] = 0.0;
  const uint64_t db = static_cast<uint64_t>(s) * a.d.segment_stride;
  const uint64_t mb = static_cast<uint64_t>(s) * a.m.segment_stride;
  const uint64_t stride = static_cast<uint64_t>(a.bps) * kBlk;
  const double maxd = a.prm.max_dist, ks = a.kscale;
  for (uint64_t i = static_cast<uint64_t>(bb) * kBlk + threadIdx.x; i < a.seg; i += stride) {
    const uint64_t di = db + i, mi = mb + i;
    if (a.d.mask && a.d.mask[di] == 0) continue;
    if (a.m.mask && a.m.mask[mi] == 0) continue;
    if (a.d.ids && a.d.ids[di] != a.prm.dataset_id) continue;
    if (a.m.ids && a.m.ids[mi] != a.prm.model_id) continue;
    const rmh_vec3 dp = a.d.points[di];
    const rmh_vec3 mp = a.m.points[mi];
    const rmh_vec3 nn = a.m.normals[mi];
    const double px = R.m[0] * dp.x + R.m[1] * dp.y + R.m[2] * dp.z + t[0];
    const double py = R.m[3] * dp.x + R.m[4] * dp.y + R.m[5] * dp.z + t[1];
    const double pz = R.m[6] * dp.x + R.m[7] * dp.y + R.m[8] * dp.z + t[2];
    const double nx = nn.x, ny = nn.y, nz = nn.z;
    const double r = (px - mp.x) * nx + (py - mp.y) * ny + (pz - mp.z) * nz;
    if (!(fabs(r) < maxd)) continue;
    const double J[6] = {py * nz - pz * ny, pz * nx - px * nz, px * ny - py * nx, nx, ny, nz};
    double w = 1.0;
    if (ks > 0.0) {
      const double den = ks + r;
      w = (ks * ks) / (den * den);
    }
    m.v[0] += 1.0;
#pragma unroll
    for (int k = 0; k < 6; k++) m.v[1 + k] += w * J[k] * r;
    int u = 7;
#pragma unroll
    for (int p = 0; p < 6; p++)
#pragma unroll
      for (int q = p; q < 6; q++, u++) m.v[u] += w * J[p] * J[q];
    m.v[28] += w * r * r;
  }
  block_sum(m, lds);
  if (threadIdx.x != 0) return;
  if (a.bps == 1) finalize_system(&a.out[s], m);
  else
    for (int k = 0; k < kLin; k++) a.partial[(static_cast<uint64_t>(s) * a.bps + bb) * kLin + k] = m.v[k];
}

__global__ void k_p2l_system_combine(const LinArgs a) {
  const uint32_t s = blockIdx.x * kBlk + threadIdx.x;
  if (s >= a.nseg) return;
  Acc<kLin> m;
  for (int k = 0; k < kLin; k++) m.v[k] = 0.0;
  for (uint32_t b = 0; b < a.bps; b++)
    for (int k = 0; k < kLin; k++) m.v[k] += a.partial[(static_cast<uint64_t>(s) * a.bps + b) * kLin + k];
  finalize_system(&a.out[s], m);
}

// ---- Umeyama ------------------------------------------------------------------------------

__host__ __device__ __forceinline__ double det3(const double* M) {  // row-major
  return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) + M[2] * (M[3] * M[7] - M[4] * M[6]);
}

// One-sided Jacobi SVD of a 3x3 matrix A (row-major): A = U diag(s) V^T, s descending,
// U and V orthonormal (U completed by cross products for rank-deficient A).
__host__ __device__ void svd3(const double* Ain, double* U, double* s, double* V) {
  double A[9];
  for (int k = 0; k < 9; k++) A[k] = Ain[k];
  for (int k = 0; k < 9; k++) V[k] = (k % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 40; sweep++) {
    bool rotated = false;
    for (int pq = 0; pq < 3; pq++) {
      const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
      double al = 0, be = 0, ga = 0;
      for (int r = 0; r < 3; r++) {
        al += A[r * 3 + p] * A[r * 3 + p];
        be += A[r * 3 + q] * A[r * 3 + q];
        ga += A[r * 3 + p] * A[r * 3 + q];
      }
      if (fabs(ga) <= 1e-15 * sqrt(al * be) || ga == 0.0) continue;
      rotated = true;
      const double zeta = (be - al) / (2.0 * ga);
      const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
      const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
      for (int r = 0; r < 3; r++) {
        const double ap = A[r * 3 + p], aq = A[r * 3 + q];
        A[r * 3 + p] = cs * ap - sn * aq;
        A[r * 3 + q] = sn * ap + cs * aq;
        const double vp = V[r * 3 + p], vq = V[r * 3 + q];
        V[r * 3 + p] = cs * vp - sn * vq;
        V[r * 3 + q] = sn * vp + cs * vq;
      }
    }
    if (!rotated) break;
  }
  double sv[3];
  for (int c = 0; c < 3; c++) sv[c] = sqrt(A[c] * A[c] + A[3 + c] * A[3 + c] + A[6 + c] * A[6 + c]);
  int ord[3] = {0, 1, 2};
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2 - i; j++)
      if (sv[ord[j]] < sv[ord[j + 1]]) { const int t = ord[j]; ord[j] = ord[j + 1]; ord[j + 1] = t; }
  double Vs[9];
  for (int c = 0; c < 3; c++) {
    s[c] = sv[ord[c]];
    for (int r = 0; r < 3; r++) {
      Vs[r * 3 + c] = V[r * 3 + ord[c]];
      U[r * 3 + c] = s[c] > 0 ? A[r * 3 + ord[c]] / s[c] : 0.0;
    }
  }
  for (int k = 0; k < 9; k++) V[k] = Vs[k];
  const double tiny = 1e-13 * s[0];
  if (!(s[1] > tiny)) {
    // rank <= 1: complete U[:,1] orthogonal to U[:,0]
    const double ux = U[0], uy = U[3], uz = U[6];
    double ax = 1, ay = 0, az = 0;
    if (fabs(ux) > 0.9) { ax = 0; ay = 1; }
    double cx = uy * az - uz * ay, cy = uz * ax - ux * az, cz = ux * ay - uy * ax;
    const double l = sqrt(cx * cx + cy * cy + cz * cz);
    U[1] = cx / l; U[4] = cy / l; U[7] = cz / l;
  }
  if (!(s[2] > tiny)) {
    U[2] = U[3] * U[7] - U[6] * U[4];
    U[5] = U[6] * U[1] - U[0] * U[7];
    U[8] = U[0] * U[4] - U[3] * U[1];
  }
}

// Umeyama (optimization.cpp:12-49): C = U S V^T, R = U diag(1, 1, sign(det U det V)) V^T,
// converted to a quaternion with the reference's matrix -> quaternion formula
// (src/rmagine_core/include/rmagine/math/types/MatrixOps.tcc:1492-1529), normalised;
// a non-finite rotation falls back to identity; t = m - R d; n_meas == 0 -> identity.
// The same code runs on the device (k_umeyama) and on the host (rmh_umeyama_host).
// v - v is 0 for finite v and NaN for infinities and NaN (no fast-math here)
__host__ __device__ inline bool finite_hd(double v) { return v - v == 0.0; }

__host__ __device__ void umeyama_one(const rmh_cross_stats& S, double q[4], double t[3]) {
  q[0] = 0.0; q[1] = 0.0; q[2] = 0.0; q[3] = 1.0;
  t[0] = 0.0; t[1] = 0.0; t[2] = 0.0;
  if (S.n_meas == 0) return;
  double C[9];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) C[r * 3 + c] = S.covariance[c * 3 + r];
  double U[9], sv[3], V[9];
  svd3(C, U, sv, V);
  bool ok = sv[0] > 0.0;
  if (ok) {
    const double sd = (det3(U) * det3(V) < 0.0) ? -1.0 : 1.0;
    double R[9];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++)
        R[r * 3 + c] = U[r * 3 + 0] * V[c * 3 + 0] + U[r * 3 + 1] * V[c * 3 + 1] + sd * U[r * 3 + 2] * V[c * 3 + 2];
    const double tr = R[0] + R[4] + R[8];
    double x, y, z, w;
    if (tr > 0) {
      const double s4 = sqrt(tr + 1.0) * 2.0;
      w = 0.25 * s4; x = (R[7] - R[5]) / s4; y = (R[2] - R[6]) / s4; z = (R[3] - R[1]) / s4;
    } else if (R[0] > R[4] && R[0] > R[8]) {
      const double s4 = sqrt(1.0 + R[0] - R[4] - R[8]) * 2.0;
      w = (R[7] - R[5]) / s4; x = 0.25 * s4; y = (R[1] + R[3]) / s4; z = (R[2] + R[6]) / s4;
    } else if (R[4] > R[8]) {
      const double s4 = sqrt(1.0 + R[4] - R[0] - R[8]) * 2.0;
      w = (R[2] - R[6]) / s4; x = (R[1] + R[3]) / s4; y = 0.25 * s4; z = (R[5] + R[7]) / s4;
    } else {
      const double s4 = sqrt(1.0 + R[8] - R[0] - R[4]) * 2.0;
      w = (R[3] - R[1]) / s4; x = (R[2] + R[6]) / s4; y = (R[5] + R[7]) / s4; z = 0.25 * s4;
    }
    const double l = sqrt(x * x + y * y + z * z + w * w);
    x /= l; y /= l; z /= l; w /= l;
    ok = finite_hd(x) && finite_hd(y) && finite_hd(z) && finite_hd(w);
    if (ok) { q[0] = x; q[1] = y; q[2] = z; q[3] = w; }
  }
  // t = m - R * d, R applied as the quaternion sandwich
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  const double d0 = S.dataset_mean[0], d1 = S.dataset_mean[1], d2 = S.dataset_mean[2];
  const double r0 = (w * w + x * x - y * y - z * z) * d0 + 2.0 * (x * y - w * z) * d1 + 2.0 * (x * z + w * y) * d2;
  const double r1 = 2.0 * (x * y + w * z) * d0 + (w * w - x * x + y * y - z * z) * d1 + 2.0 * (y * z - w * x) * d2;
  const double r2 = 2.0 * (x * z - w * y) * d0 + 2.0 * (y * z + w * x) * d1 + (w * w - x * x - y * y + z * z) * d2;
  t[0] = S.model_mean[0] - r0;
  t[1] = S.model_mean[1] - r1;
  t[2] = S.model_mean[2] - r2;
}

__host__ __device__ inline void umeyama_store(const double q[4], const double t[3], uint32_t i, rmh_transform* out,
                                              double* out64) {
  rmh_transform T;
  T.R.x = (float)q[0]; T.R.y = (float)q[1]; T.R.z = (float)q[2]; T.R.w = (float)q[3];
  T.t.x = (float)t[0]; T.t.y = (float)t[1]; T.t.z = (float)t[2];
  T.stamp = 0;
  if (out) out[i] = T;
  if (out64) {
    for (int k = 0; k < 4; k++) out64[7 * static_cast<size_t>(i) + k] = q[k];
    for (int k = 0; k < 3; k++) out64[7 * static_cast<size_t>(i) + 4 + k] = t[k];
  }
}

__global__ void k_umeyama(const rmh_cross_stats* st, uint32_t n, rmh_transform* out, double* out64) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  double q[4], t[3];
  umeyama_one(st[i], q, t);
  umeyama_store(q, t, i, out, out64);
}

// CrossStatistics_<float> (CrossStatistics.hpp:50-60, 64 bytes): dataset mean, model mean,
// covariance (column-major), n_meas
struct CrossStatsF32 {
  float dataset_mean[3], model_mean[3], covariance[9];
  uint32_t n_meas;
};

__global__ void k_stats_to_f32(const rmh_cross_stats* in, CrossStatsF32* out, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  const rmh_cross_stats S = in[i];
  CrossStatsF32 r;
  for (int k = 0; k < 3; k++) { r.dataset_mean[k] = (float)S.dataset_mean[k]; r.model_mean[k] = (float)S.model_mean[k]; }
  for (int k = 0; k < 9; k++) r.covariance[k] = (float)S.covariance[k];
  r.n_meas = S.n_meas;
  out[i] = r;
}

// (dataset means, model means, covariances, counts) -> FP64 statistics; n_meas NULL -> 1
// (the default of umeyama_transform, optimization.h:61-65)
__global__ void k_stats_pack(const rmh_vec3* ds, const rmh_vec3* ms, const float* Cs, const uint32_t* n_meas,
                             rmh_cross_stats* out, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  if (i >= n) return;
  rmh_cross_stats r;
  const rmh_vec3 d = ds[i], m = ms[i];
  r.dataset_mean[0] = d.x; r.dataset_mean[1] = d.y; r.dataset_mean[2] = d.z;
  r.model_mean[0] = m.x; r.model_mean[1] = m.y; r.model_mean[2] = m.z;
  for (int k = 0; k < 9; k++) r.covariance[k] = Cs[9 * static_cast<size_t>(i) + k];
  r.n_meas = n_meas ? n_meas[i] : 1u;
  r._pad = 0;
  out[i] = r;
}

uint32_t blocks_per_segment(uint64_t seg, uint32_t nseg) {
  // enough blocks for several rounds of the chip's resident blocks (>= 8192 in total, so
  // that the last, partly filled round is a small part of the launch), at least 2048
  // elements per block
  const uint64_t by_size = (seg + 2047) / 2048;
  const uint64_t want = (8192 + nseg - 1) / nseg;
  uint64_t b = want < by_size ? want : by_size;
  if (b < 1) b = 1;
  if (b > 256) b = 256;
  return static_cast<uint32_t>(b);
}

// device memory the call allocates itself when the caller passes no workspace
struct TempAlloc {
  void* p = nullptr;
  ~TempAlloc() { if (p) (void)hipFree(p); }
};

}  // namespace

extern "C" {

int rmh_transform_mul(const rmh_transform* A, uint32_t nA, const rmh_transform* B, uint32_t nB, rmh_transform* C,
                      uint32_t n, void* stream) {
  if (n == 0) return RMH_OK;
  if (!A || !B || !C) return fail(RMH_ERR_INVALID, "rmh_transform_mul: NULL pointer");
  if ((nA != 1 && nA != n) || (nB != 1 && nB != n)) return fail(RMH_ERR_INVALID, "rmh_transform_mul: sizes must be 1 or n");
  k_tf_mul<<<grid1(n), kBlk, 0, static_cast<hipStream_t>(stream)>>>(A, nA, B, nB, C, n);
  RMH_HIP_TRY(hipGetLastError());
  return RMH_OK;
}

int rmh_transform_inv(const rmh_transform* A, rmh_transform* C, uint32_t n, void* stream) {
  if (n == 0) return RMH_OK;
  if (!A || !C) return fail(RMH_ERR_INVALID, "rmh_transform_inv: NULL pointer");
  k_tf_inv<<<grid1(n), kBlk, 0, static_cast<hipStream_t>(stream)>>>(A, C, n);
  RMH_HIP_TRY(hipGetLastError());
  return RMH_OK;
}

int rmh_transform_points(const rmh_transform* T, uint32_t nT, const rmh_vec3* p, rmh_vec3* out, uint32_t n, void* stream) {
  if (n == 0) return RMH_OK;
  if (!T || !p || !out) return fail(RMH_ERR_INVALID, "rmh_transform_points: NULL pointer");
  if (nT != 1 && nT != n) return fail(RMH_ERR_INVALID, "rmh_transform_points: nT must be 1 or n");
  k_tf_points<<<grid1(n), kBlk, 0, static_cast<hipStream_t>(stream)>>>(T, nT, p, out, n);
  RMH_HIP_TRY(hipGetLastError());
  return RMH_OK;
}

int rmh_transform_set_identity(rmh_transform* T, uint32_t n, void* stream) {
  if (n == 0) return RMH_OK;
  if (!T) return fail(RMH_ERR_INVALID, "rmh_transform_set_identity: NULL pointer");
  k_set_identity<<<grid1(n), kBlk, 0, static_cast<hipStream_t>(stream)>>>(T, n);
  RMH_HIP_TRY(hipGetLastError());
  return RMH_OK;
}

int rmh_quat_normalize(rmh_transform* T, uint32_t n, void* stream) {
  if (n == 0) return RMH_OK;
  if (!T) return fail(RMH_ERR_INVALID, "rmh_quat_normalize: NULL pointer");
  k_quat_normalize<<<grid1(n), kBlk, 0, static_cast<hipStream_t>(stream)>>>(T, n);
  RMH_HIP_TRY(hipGetLastError());
  return RMH_OK;
}

int rmh_transform_pack_matrix(const float* R, const rmh_vec3* t, rmh_transform* T, uint32_t n, void* stream) {
  if (n == 0) return RMH_OK;
  if (!R || !t || !T) return fail(RMH_ERR_INVALID, "rmh_transform_pack_matrix: NULL pointer");
  k_tf_pack_mat<<<grid1(n), kBlk, 0, static_cast<hipStream_t>(stream)>>>(R, t, T, n);
  RMH_HIP_TRY(hipGetLastError());
  return RMH_OK;
}

int rmh_transform_pack_quat(const rmh_quat* q, const rmh_vec3* t, rmh_transform* T, uint32_t n, void* stream) {
  if (n == 0) return RMH_OK;
  if (!q || !t || !T) return fail(RMH_ERR_INVALID, "rmh_transform_pack_quat: NULL pointer");
  k_tf_pack_quat<<<grid1(n), kBlk, 0, static_cast<hipStream_t>(stream)>>>(q, t, T, n);
  RMH_HIP_TRY(hipGetLastError());
  return RMH_OK;
}

// one workspace serves every segmented reduction: sized for the widest accumulator set
int rmh_reduction_workspace_size(uint64_t segment_size, uint32_t n_segments, size_t* bytes) {
  if (!bytes) return fail(RMH_ERR_INVALID, "rmh_reduction_workspace_size: NULL");
  const uint32_t b = blocks_per_segment(segment_size, n_segments ? n_segments : 1);
  *bytes = b > 1 ? sizeof(double) * kLin * static_cast<size_t>(b) * n_segments : 0;
  return RMH_OK;
}

int rmh_vec3_moments(const rmh_vec3* a, const rmh_vec3* b, const uint8_t* mask, uint64_t segment_size,
                     uint32_t n_segments, double* sum_a, double* mean_a, double* cov, uint32_t* count,
                     void* workspace, size_t workspace_bytes, void* stream_) {
  if (n_segments == 0) return RMH_OK;
  if (!a && segment_size > 0) return fail(RMH_ERR_INVALID, "rmh_vec3_moments: NULL input");
  if (n_segments > 65535u) return fail(RMH_ERR_UNSUPPORTED, "rmh_vec3_moments: at most 65535 segments");
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  MomArgs A{a, b, mask, segment_size, n_segments, blocks_per_segment(segment_size, n_segments), nullptr,
            sum_a, mean_a, cov, count};
  TempAlloc tmp;
  if (A.bps > 1) {
    const size_t need = sizeof(double) * kMom * static_cast<size_t>(A.bps) * n_segments;
    if (workspace) {
      if (workspace_bytes < need) return fail(RMH_ERR_INVALID, "rmh_vec3_moments: workspace too small");
      A.partial = static_cast<double*>(workspace);
    } else {
      RMH_HIP_TRY(hipMalloc(&tmp.p, need));
      A.partial = static_cast<double*>(tmp.p);
    }
  }
  k_moments<<<dim3(A.bps, n_segments), kBlk, 0, stream>>>(A);
  RMH_HIP_TRY(hipGetLastError());
  if (A.bps > 1) {
    k_moments_combine<<<grid1(n_segments), kBlk, 0, stream>>>(A);
    RMH_HIP_TRY(hipGetLastError());
    if (tmp.p) RMH_HIP_TRY(hipStreamSynchronize(stream));
  }
  return RMH_OK;
}

}  // extern "C"

namespace {

// Argument checks and set-up shared by rmh_statistics and rmh_statistics_noalloc; leaves
// A.partial unset.  Returns RMH_OK, or an error whose message starts with `who`.
int stats_prepare(const char* who, int kind, const rmh_transform* pre_transforms, uint32_t n_pre,
                  const rmh_cloud* dataset, const rmh_cloud* model, uint64_t segment_size, uint32_t n_segments,
                  const rmh_umeyama_constraints* params, rmh_cross_stats* stats, StatArgs& A) {
  const std::string w(who);
  if (!pre_transforms || !dataset || !model || !params || !stats) return fail(RMH_ERR_INVALID, w + ": NULL argument");
  if (kind != RMH_STATS_P2P && kind != RMH_STATS_P2L) return fail(RMH_ERR_INVALID, w + ": bad kind");
  if (n_pre != 1 && n_pre != n_segments) return fail(RMH_ERR_INVALID, w + ": n_pre must be 1 or n_segments");
  if (segment_size > 0 && (!dataset->points || !model->points)) return fail(RMH_ERR_INVALID, w + ": points are required");
  if (kind == RMH_STATS_P2L && segment_size > 0 && !model->normals)
    return fail(RMH_ERR_INVALID, w + ": point-to-plane needs model normals");
  if (n_segments > 65535u) return fail(RMH_ERR_UNSUPPORTED, w + ": at most 65535 segments");
  A.kind = kind;
  A.pre = pre_transforms;
  A.npre = n_pre;
  A.d = *dataset;
  A.m = *model;
  A.seg = segment_size;
  A.nseg = n_segments;
  A.bps = blocks_per_segment(segment_size, n_segments);
  A.prm = *params;
  A.partial = nullptr;
  A.out = stats;
  return RMH_OK;
}

size_t stats_partial_bytes(const StatArgs& A) {
  return A.bps > 1 ? sizeof(double) * kMom * static_cast<size_t>(A.bps) * A.nseg : 0;
}

// k_stats, then (segments split over several blocks) k_stats_combine over A.partial
int stats_launch(const StatArgs& A, hipStream_t stream) {
  // 32-bit offsets when i * 12 B (the largest element) and i + stride stay below 2^32
  if (A.seg + static_cast<uint64_t>(A.bps) * kBlk < (1ull << 32) / sizeof(rmh_vec3))
    k_stats<uint32_t><<<dim3(A.bps, A.nseg), kBlk, 0, stream>>>(A);
  else
    k_stats<uint64_t><<<dim3(A.bps, A.nseg), kBlk, 0, stream>>>(A);
  RMH_HIP_TRY(hipGetLastError());
  if (A.bps > 1) {
    k_stats_combine<<<grid1(A.nseg), kBlk, 0, stream>>>(A);
    RMH_HIP_TRY(hipGetLastError());
  }
  return RMH_OK;
}

}  // namespace

extern "C" {

int rmh_statistics(int kind, const rmh_transform* pre_transforms, uint32_t n_pre, const rmh_cloud* dataset,
                   const rmh_cloud* model, uint64_t segment_size, uint32_t n_segments,
                   const rmh_umeyama_constraints* params, rmh_cross_stats* stats, void* workspace,
                   size_t workspace_bytes, void* stream_) {
  if (n_segments == 0) return RMH_OK;
  StatArgs A;
  const int rc = stats_prepare("rmh_statistics", kind, pre_transforms, n_pre, dataset, model, segment_size, n_segments,
                               params, stats, A);
  if (rc != RMH_OK) return rc;
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  TempAlloc tmp;
  if (A.bps > 1) {
    const size_t need = stats_partial_bytes(A);
    if (workspace) {
      if (workspace_bytes < need) return fail(RMH_ERR_INVALID, "rmh_statistics: workspace too small");
      A.partial = static_cast<double*>(workspace);
    } else {
      RMH_HIP_TRY(hipMalloc(&tmp.p, need));
      A.partial = static_cast<double*>(tmp.p);
    }
  }
  const int lc = stats_launch(A, stream);
  if (lc != RMH_OK) return lc;
  if (tmp.p) RMH_HIP_TRY(hipStreamSynchronize(stream));
  return RMH_OK;
}

}  // extern "C"

// ---- allocation-free statistics -----------------------------------------------------------
//
// rmh_statistics_noalloc keeps the per-block partial sums of rmh_statistics in a fixed
// device array instead of a caller workspace or a hipMalloc.  Each GPU the library runs on
// holds its own copy of the array (2 MiB) until the process ends.  Every call on one GPU
// uses the same array: calls that overlap on one GPU (two streams, two host threads) race.
// Nothing in it needs clearing: k_stats writes every partial slot that k_stats_combine reads.

namespace {

constexpr size_t kScratchBytes = size_t(2) << 20;
// the largest partial set rmh_statistics' partition makes is bps * n_segments < 16384 slots
// of kMom doubles (bps > 1 only for n_segments <= 8191, and bps * n_segments < 8192 +
// n_segments), 2 MiB
static_assert(sizeof(CrossStatsF32) == 64, "CrossStatistics_<float> is 64 bytes");
static_assert(sizeof(double) * kMom * 16384 <= kScratchBytes, "scratch holds the largest partial set");

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of this device's `scratch` (the array's name is only a host-side stand-in),
// looked up once per device with that device current.
int scratch_address(void** out) {
  int dev = 0;
  RMH_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(RMH_ERR_UNSUPPORTED, "rmh_statistics_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return RMH_OK;
}

}  // namespace

extern "C" {

int rmh_statistics_noalloc(int kind, const rmh_transform* pre_transforms, uint32_t n_pre, const rmh_cloud* dataset,
                           const rmh_cloud* model, uint64_t segment_size, uint32_t n_segments,
                           const rmh_umeyama_constraints* params, rmh_cross_stats* stats, void* stream) {
  if (n_segments == 0) return RMH_OK;
  StatArgs A;
  const int rc = stats_prepare("rmh_statistics_noalloc", kind, pre_transforms, n_pre, dataset, model, segment_size,
                               n_segments, params, stats, A);
  if (rc != RMH_OK) return rc;
  const size_t need = stats_partial_bytes(A);
  if (need > kScratchBytes)
    return fail(RMH_ERR_UNSUPPORTED, "rmh_statistics_noalloc: partial sums need " + std::to_string(need) +
                                         " bytes, more than the 2 MiB scratch array");
  if (need > 0) {
    void* p = nullptr;
    const int sc = scratch_address(&p);
    if (sc != RMH_OK) return sc;
    A.partial = static_cast<double*>(p);
  }
  return stats_launch(A, static_cast<hipStream_t>(stream));
}

int rmh_statistics_objectwise_p2l(const rmh_transform* pre_transforms, const rmh_cloud* dataset, const rmh_cloud* model,
                                  uint32_t width, uint32_t height, uint32_t n_points,
                                  const rmh_umeyama_constraints* params, const rmh_aabb* bboxes, uint32_t n_objects,
                                  uint32_t flags, rmh_cross_stats* stats, void* stream) {
  (void)height;  // the reference's kernel takes it but addresses rows with width only
  if (n_objects == 0) return RMH_OK;
  if (!pre_transforms || !dataset || !model || !params || !bboxes || !stats)
    return fail(RMH_ERR_INVALID, "rmh_statistics_objectwise_p2l: NULL argument");
  if (flags & ~RMH_OBJECTWISE_WINDOW_ONLY) return fail(RMH_ERR_INVALID, "rmh_statistics_objectwise_p2l: unknown flags");
  if (width == 0) return fail(RMH_ERR_INVALID, "rmh_statistics_objectwise_p2l: width must be > 0");
  if (n_points > 0 && (!dataset->points || !model->points || !model->normals))
    return fail(RMH_ERR_INVALID, "rmh_statistics_objectwise_p2l: dataset points, model points and normals are required");
  if (n_objects > 0x7FFFFFFFu) return fail(RMH_ERR_UNSUPPORTED, "rmh_statistics_objectwise_p2l: too many objects");
  ObjArgs A;
  A.pre = pre_transforms;
  A.d = *dataset;
  A.m = *model;
  A.width = width;
  A.n_points = n_points;
  A.prm = params;
  A.boxes = bboxes;
  A.window_only = (flags & RMH_OBJECTWISE_WINDOW_ONLY) ? 1u : 0u;
  A.out = stats;
  k_stats_objectwise<<<n_objects, kBlk, 0, static_cast<hipStream_t>(stream)>>>(A);
  RMH_HIP_TRY(hipGetLastError());
  return RMH_OK;
}

int rmh_linear_system_p2l(const rmh_transform* pre_transforms, uint32_t n_pre, const rmh_cloud* dataset,
                          const rmh_cloud* model, uint64_t segment_size, uint32_t n_segments,
                          const rmh_umeyama_constraints* params, double kernel_scale, rmh_p2l_system* out,
                          void* workspace, size_t workspace_bytes, void* stream_) {
  if (n_segments == 0) return RMH_OK;
  if (!pre_transforms || !dataset || !model || !params || !out)
    return fail(RMH_ERR_INVALID, "rmh_linear_system_p2l: NULL argument");
  if (n_pre != 1 && n_pre != n_segments) return fail(RMH_ERR_INVALID, "rmh_linear_system_p2l: n_pre must be 1 or n_segments");
  if (segment_size > 0 && (!dataset->points || !model->points || !model->normals))
    return fail(RMH_ERR_INVALID, "rmh_linear_system_p2l: dataset points, model points and normals are required");
  if (n_segments > 65535u) return fail(RMH_ERR_UNSUPPORTED, "rmh_linear_system_p2l: at most 65535 segments");
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  LinArgs A;
  A.pre = pre_transforms;
  A.npre = n_pre;
  A.d = *dataset;
  A.m = *model;
  A.seg = segment_size;
  A.nseg = n_segments;
  A.bps = blocks_per_segment(segment_size, n_segments);
  A.prm = *params;
  A.kscale = kernel_scale;
  A.partial = nullptr;
  A.out = out;
  TempAlloc tmp;
  if (A.bps > 1) {
    const size_t need = sizeof(double) * kLin * static_cast<size_t>(A.bps) * n_segments;
    if (workspace) {
      if (workspace_bytes < need) return fail(RMH_ERR_INVALID, "rmh_linear_system_p2l: workspace too small");
      A.partial = static_cast<double*>(workspace);
    } else {
      RMH_HIP_TRY(hipMalloc(&tmp.p, need));
      A.partial = static_cast<double*>(tmp.p);
    }
  }
  k_p2l_system<<<dim3(A.bps, n_segments), kBlk, 0, stream>>>(A);
  RMH_HIP_TRY(hipGetLastError());
  if (A.bps > 1) {
    k_p2l_system_combine<<<grid1(n_segments), kBlk, 0, stream>>>(A);
    RMH_HIP_TRY(hipGetLastError());
    if (tmp.p) RMH_HIP_TRY(hipStreamSynchronize(stream));
  }
  return RMH_OK;
}

int rmh_umeyama(const rmh_cross_stats* stats, uint32_t n, rmh_transform* out, double* out64, void* stream) {
  if (n == 0) return RMH_OK;
  if (!stats || (!out && !out64)) return fail(RMH_ERR_INVALID, "rmh_umeyama: NULL argument");
  k_umeyama<<<grid1(n), kBlk, 0, static_cast<hipStream_t>(stream)>>>(stats, n, out, out64);
  RMH_HIP_TRY(hipGetLastError());
  return RMH_OK;
}

int rmh_umeyama_host(const rmh_cross_stats* stats, uint32_t n, rmh_transform* out, double* out64) {
  if (n == 0) return RMH_OK;
  if (!stats || (!out && !out64)) return fail(RMH_ERR_INVALID, "rmh_umeyama_host: NULL argument");
  for (uint32_t i = 0; i < n; i++) {
    double q[4], t[3];
    umeyama_one(stats[i], q, t);
    umeyama_store(q, t, i, out, out64);
  }
  return RMH_OK;
}

int rmh_cross_stats_to_f32(const rmh_cross_stats* in, void* out, uint32_t n, void* stream) {
  if (n == 0) return RMH_OK;
  if (!in || !out) return fail(RMH_ERR_INVALID, "rmh_cross_stats_to_f32: NULL argument");
  k_stats_to_f32<<<grid1(n), kBlk, 0, static_cast<hipStream_t>(stream)>>>(in, static_cast<CrossStatsF32*>(out), n);
  RMH_HIP_TRY(hipGetLastError());
  return RMH_OK;
}

int rmh_cross_stats_pack(const rmh_vec3* dataset_means, const rmh_vec3* model_means, const float* covariances,
                         const uint32_t* n_meas, rmh_cross_stats* out, uint32_t n, void* stream) {
  if (n == 0) return RMH_OK;
  if (!dataset_means || !model_means || !covariances || !out)
    return fail(RMH_ERR_INVALID, "rmh_cross_stats_pack: NULL argument");
  k_stats_pack<<<grid1(n), kBlk, 0, static_cast<hipStream_t>(stream)>>>(dataset_means, model_means, covariances,
                                                                           n_meas, out, n);
  RMH_HIP_TRY(hipGetLastError());
  return RMH_OK;
}

}  // extern "C"
