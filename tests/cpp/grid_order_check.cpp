// Host check of rmh_grid_order.hpp: for every grid of 1..70 columns by 1..21 rows and a few
// scanner sizes, the tile order is a permutation of the buffer indices, and when 4 divides
// the rows and 16 the columns every run of 64 slots is one 4 x 16 block of the grid.
// Prints the number of checked grids; exit status 0 when all pass.
#include "../../rmagine_amd/csrc/rmh_grid_order.hpp"

#include <cstdint>
#include <cstdio>
#include <vector>

static bool check(uint32_t W, uint32_t H) {
  const std::vector<uint32_t> o = rmh::grid_tile_order(W, H);
  const size_t n = static_cast<size_t>(W) * H;
  if (o.size() != n) { std::printf("FAILED %ux%u: %zu slots\n", W, H, o.size()); return false; }
  std::vector<char> seen(n, 0);
  for (uint32_t v : o) {
    if (v >= n || seen[v]) { std::printf("FAILED %ux%u: index %u out of range or repeated\n", W, H, v); return false; }
    seen[v] = 1;
  }
  if (W % 16 == 0 && H % 4 == 0)
    for (size_t t = 0; t < n / 64; t++) {
      const uint32_t r0 = o[64 * t] / W, c0 = o[64 * t] % W;
      for (uint32_t l = 0; l < 64; l++) {
        const uint32_t v = o[64 * t + l];
        if (v / W != r0 + l / 16 || v % W != c0 + l % 16) {
          std::printf("FAILED %ux%u: slot %zu is not in the tile of slot %zu\n", W, H, 64 * t + l, 64 * t);
          return false;
        }
      }
    }
  return true;
}

int main() {
  long grids = 0;
  for (uint32_t W = 1; W <= 70; W++)
    for (uint32_t H = 1; H <= 21; H++, grids++)
      if (!check(W, H)) return 1;
  const uint32_t sizes[][2] = {{1024, 128}, {2048, 64}, {900, 16}, {512, 32}, {1, 100000}, {100000, 1}};
  for (const auto& s : sizes) {
    if (!check(s[0], s[1])) return 1;
    grids++;
  }
  std::printf("%ld\n", grids);
  return 0;
}
