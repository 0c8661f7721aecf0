// Host check of rmh_fastdiv.hpp against the built-in unsigned division: every divisor
// 1..65536 with dividends around each multiple of it and at the ends of the 32-bit range,
// then powers of two and their neighbours, the largest divisors, and random pairs.
// Prints the number of checked pairs; exit status 0 when every quotient matches.
#include "../../rmagine_amd/csrc/rmh_fastdiv.hpp"

#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

using rmh::fast_div;
using rmh::make_fastdiv;

static uint64_t checked = 0;

static bool check(uint32_t n, uint32_t d, const rmh::FastDiv& f) {
  checked++;
  if (fast_div(n, f) == n / d) return true;
  std::printf("FAILED: %u / %u = %u, fast_div gives %u\n", n, d, n / d, fast_div(n, f));
  return false;
}

static bool check_divisor(uint32_t d, std::mt19937& rng) {
  const rmh::FastDiv f = make_fastdiv(d);
  const uint32_t ends[] = {0u, 1u, 2u, d - 1u, d, d + 1u, 0x7fffffffu, 0x80000000u, 0xfffffffeu, 0xffffffffu};
  for (uint32_t n : ends)
    if (!check(n, d, f)) return false;
  // around multiples q d of the divisor, q spread over the whole quotient range
  const uint32_t qmax = 0xffffffffu / d;
  for (int k = 0; k < 64; k++) {
    const uint32_t q = k < 32 ? static_cast<uint32_t>(k) : static_cast<uint32_t>(rng() % (uint64_t(qmax) + 1));
    const uint64_t base = uint64_t(q) * d;
    for (int64_t o = -1; o <= 1; o++) {
      const int64_t n = static_cast<int64_t>(base) + o;
      if (n >= 0 && n <= 0xffffffffll && !check(static_cast<uint32_t>(n), d, f)) return false;
    }
    if (base + d - 1 <= 0xffffffffull && !check(static_cast<uint32_t>(base + d - 1), d, f)) return false;
  }
  return true;
}

int main() {
  std::mt19937 rng(12345);
  std::vector<uint32_t> divisors;
  for (uint32_t d = 1; d <= 65536; d++) divisors.push_back(d);
  for (int b = 16; b < 32; b++) {
    const uint32_t p = 1u << b;
    divisors.push_back(p - 1);
    divisors.push_back(p);
    divisors.push_back(p + 1);
  }
  divisors.push_back(0xfffffffeu);
  divisors.push_back(0xffffffffu);
  for (int k = 0; k < 20000; k++) divisors.push_back(1u + rng() % 0xffffffffu);
  for (uint32_t d : divisors)
    if (!check_divisor(d, rng)) return 1;
  for (int k = 0; k < 2000000; k++) {
    const uint32_t d = 1u + rng() % 0xffffffffu, n = rng();
    if (!check(n, d, make_fastdiv(d))) return 1;
  }
  // every dividend below 2^20 for the divisors the kernels meet most (tile counts)
  for (uint32_t d : {1u, 3u, 7u, 16u, 20u, 48u, 64u, 100u, 2048u, 4800u, 32768u})
    for (uint32_t n = 0; n < (1u << 20); n++)
      if (!check(n, d, make_fastdiv(d))) return 1;
  std::printf("%llu\n", static_cast<unsigned long long>(checked));
  return 0;
}
