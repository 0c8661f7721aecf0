// Division of unsigned 32-bit integers by a run-time constant without a divide instruction:
// the round-up method of Granlund and Montgomery ("Division by Invariant Integers using
// Multiplication", PLDI 1994, figure 4.1), exact for every 32-bit dividend and divisor >= 1.
//
// The kernels divide wave-uniform block indices by per-launch constants (tiles per pose,
// tiles per image row).  The compiler expands `n / d` into a float-reciprocal sequence on
// the vector ALU, which leaves the quotient in VGPRs held through the whole traversal; a
// multiply-high, two shifts and two adds on scalar operands stay in SGPRs.
//
// Host-only C++ includes this header too (tests/cpp/fastdiv_check.cpp checks it against `/`).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define RMH_DIV_HD __host__ __device__ inline
#else
#define RMH_DIV_HD inline
#endif

namespace rmh {

struct FastDiv {
  uint32_t m;       // multiplier: floor(2^32 (2^l - d) / d) + 1, l = ceil(log2 d)
  uint32_t s1, s2;  // shifts: min(l, 1) and max(l - 1, 0)
};

RMH_DIV_HD FastDiv make_fastdiv(uint32_t d) {
  uint32_t l = 0;
  while (l < 32 && (uint64_t(1) << l) < d) l++;
  FastDiv f;
  f.m = static_cast<uint32_t>(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
  f.s1 = l < 1 ? l : 1;
  f.s2 = l > 1 ? l - 1 : 0;
  return f;
}

// n / d for the d that `f` was made for
RMH_DIV_HD uint32_t fast_div(uint32_t n, const FastDiv& f) {
  const uint32_t t = static_cast<uint32_t>((static_cast<uint64_t>(f.m) * n) >> 32);
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

}  // namespace rmh
