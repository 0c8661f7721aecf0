// Range noise of librmh: the device side of Rmagine's noise models
// (src/rmagine_cuda/src/noise/GaussianNoiseCuda.cu:7-27, RelGaussianNoiseCuda.cu:7-29,
//  UniformDustNoiseCuda.cu:7-45; CPU: src/rmagine_core/src/noise/*.cpp), applied in place
// to a range buffer.
//
// The reference keeps one curandState per element and advances it on every call.  Here the
// random numbers are counter-based: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC 2011) of counter (element index lo/hi, call epoch,
// model) under key (seed, 0x5EED) gives each element four independent 32-bit words per
// call, with no state to store, grow or copy; calls with a new epoch draw new numbers.
// Per element:
//   Gaussian      r <= max_range:  r += n * stddev + mean
//   rel. Gaussian r <= max_range:  r += n * stddev * r^range_exp + mean
//   uniform dust  d = min(r, max_range); with probability 1 - (1 - p_hit)^d a dust particle
//                 at u * d is seen and returned with probability p_return^(u d): r = u d
// n is a standard normal draw (Box-Muller, u1 in (0, 1], u2 in [0, 1)), u uniform; the
// transcendental parts run in FP64 and round once to FP32.  cos(2 pi u2) comes from a
// quadrant reduction in u2 (cos_2pi), r^1 is r, and the dust powers
// of per-call bases are exp(y log base) with log base computed once on the host (pow's
// x^0 = 1 kept): differences from pow far below the FP32 rounding of the result.
#include "rmh_internal.hpp"

#include <cmath>

namespace {

using namespace rmh;

struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t kM0 = 0xD2511F53u, kM1 = 0xCD9E8D57u;  // round multipliers
  constexpr uint32_t kW0 = 0x9E3779B9u, kW1 = 0xBB67AE85u;  // key schedule (Weyl) increments
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = static_cast<uint64_t>(kM0) * c.x;
    const uint64_t p1 = static_cast<uint64_t>(kM1) * c.z;
    const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
    const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += kW0;
    k1 += kW1;
  }
  return c;
}

__device__ __forceinline__ double unit_open0(uint32_t x) { return (static_cast<double>(x >> 8) + 1.0) * 0x1p-24; }  // (0, 1]
__device__ __forceinline__ double unit_open1(uint32_t x) { return static_cast<double>(x >> 8) * 0x1p-24; }          // [0, 1)

struct NoiseArgs {
  float* r;
  uint64_t n;
  double log_miss, log_return;  // dust: log(1 - p_hit), log(p_return)
  uint32_t kind, seed, epoch;
  float a, b, c, max_range;
};

// cos(2 pi u) for u in [0, 1) given with at most 50 significant bits: 4u = k + f exactly
// (k the nearest integer, |f| <= 1/2), x = f pi/2 (one rounding), then cos(k pi/2 + x) from
// the Taylor polynomials of cos x (to x^16) and sin x (to x^17) on |x| <= pi/4, whose
// truncation errors stay below 2e-18.  Within 1e-15 of the cosine of the FP64 product
// 2 pi u, far below the FP32 rounding of a noisy range.
__device__ __forceinline__ double cos_2pi(double u) {
  const double t = 4.0 * u;
  const double k = rint(t);
  const double x = (t - k) * 1.5707963267948966;
  const double x2 = x * x;
  double c = 1.0 / 20922789888000.0;          // 1/16!
  c = fma(c, x2, -1.0 / 87178291200.0);       // 1/14!
  c = fma(c, x2, 1.0 / 479001600.0);          // 1/12!
  c = fma(c, x2, -1.0 / 3628800.0);           // 1/10!
  c = fma(c, x2, 1.0 / 40320.0);              // 1/8!
  c = fma(c, x2, -1.0 / 720.0);
  c = fma(c, x2, 1.0 / 24.0);
  c = fma(c, x2, -0.5);
  c = fma(c, x2, 1.0);
  double sn = 1.0 / 355687428096000.0;        // 1/17!
  sn = fma(sn, x2, -1.0 / 1307674368000.0);   // 1/15!
  sn = fma(sn, x2, 1.0 / 6227020800.0);       // 1/13!
  sn = fma(sn, x2, -1.0 / 39916800.0);        // 1/11!
  sn = fma(sn, x2, 1.0 / 362880.0);           // 1/9!
  sn = fma(sn, x2, -1.0 / 5040.0);
  sn = fma(sn, x2, 1.0 / 120.0);
  sn = fma(sn, x2, -1.0 / 6.0);
  sn = fma(sn, x2, 1.0) * x;
  const int q = static_cast<int>(k) & 3;
  const double r = (q & 1) ? sn : c;
  return (q == 1 || q == 2) ? -r : r;
}

// base^y from log(base) for base >= 0, as pow(base, y) has it for y = 0 (1 for every base);
// other bases (negative, NaN) go to pow
__device__ __forceinline__ double pow_from_log(double base, double log_base, double y) {
  if (y == 0.0) return 1.0;
  if (!(base >= 0.0)) return pow(base, y);
  return exp(y * log_base);
}

// One instantiation per model, so each carries only its own transcendentals (and registers).
template <uint32_t KIND>
__global__ __launch_bounds__(256) void k_noise(const NoiseArgs a) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n; i += stride) {
    const float range = a.r[i];
    const U4 w = philox4x32_10(U4{static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32), a.epoch, KIND},
                               a.seed, 0x5EEDu);
    if constexpr (KIND == RMH_NOISE_UNIFORM_DUST) {
      const float d = range > a.max_range ? a.max_range : range;
      const double p_hit = 1.0 - pow_from_log(1.0 - static_cast<double>(a.a), a.log_miss, static_cast<double>(d));
      if (unit_open1(w.x) < p_hit) {
        const float nr = static_cast<float>(unit_open1(w.y) * static_cast<double>(d));
        const double p_ret = pow_from_log(static_cast<double>(a.b), a.log_return, static_cast<double>(nr));
        if (unit_open1(w.z) < p_ret) a.r[i] = nr;
      }
    } else if (range <= a.max_range) {
      const double n = sqrt(-2.0 * log(unit_open0(w.x))) * cos_2pi(unit_open1(w.y));
      double sd = static_cast<double>(a.b);
      if constexpr (KIND == RMH_NOISE_REL_GAUSSIAN)
        sd *= a.c == 1.0f ? static_cast<double>(range) : pow(static_cast<double>(range), static_cast<double>(a.c));
      a.r[i] = static_cast<float>(static_cast<double>(range) + n * sd + static_cast<double>(a.a));
    }
  }
}

}  // namespace

extern "C" int rmh_noise_apply(const rmh_noise_params* params, float* ranges, uint64_t n, uint32_t epoch,
                               void* stream) {
  if (!params) return fail(RMH_ERR_INVALID, "rmh_noise_apply: NULL params");
  if (params->kind != RMH_NOISE_GAUSSIAN && params->kind != RMH_NOISE_REL_GAUSSIAN &&
      params->kind != RMH_NOISE_UNIFORM_DUST)
    return fail(RMH_ERR_INVALID, "rmh_noise_apply: unknown noise kind");
  if (n == 0) return RMH_OK;
  if (!ranges) return fail(RMH_ERR_INVALID, "rmh_noise_apply: NULL ranges");
  NoiseArgs a;
  a.r = ranges;
  a.n = n;
  a.kind = params->kind;
  a.seed = params->seed;
  a.epoch = epoch;
  a.a = params->p0;
  a.b = params->p1;
  a.c = params->p2;
  a.max_range = params->max_range;
  a.log_miss = std::log(1.0 - static_cast<double>(params->p0));
  a.log_return = std::log(static_cast<double>(params->p1));
  const uint64_t want = (n + 255) / 256;
  const uint32_t blocks = static_cast<uint32_t>(want < 65536 ? want : 65536);
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (a.kind == RMH_NOISE_GAUSSIAN) k_noise<RMH_NOISE_GAUSSIAN><<<blocks, 256, 0, s>>>(a);
  else if (a.kind == RMH_NOISE_REL_GAUSSIAN) k_noise<RMH_NOISE_REL_GAUSSIAN><<<blocks, 256, 0, s>>>(a);
  else k_noise<RMH_NOISE_UNIFORM_DUST><<<blocks, 256, 0, s>>>(a);
  RMH_HIP_TRY(hipGetLastError());
  return RMH_OK;
}
