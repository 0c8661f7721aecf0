// Traversal order of a ray list that is a W x H grid (buffer index row * W + column): the
// grid in 4 x 16 tiles, the shape image sensors are cast in.  Host code, shared by the
// library (rmh_sensor_set_ray_order) and its host check (tests/cpp/grid_order_check.cpp).
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace rmh {

constexpr uint32_t kGridTileRows = 4, kGridTileCols = 16;

// Buffer indices in tile order: tile rows top to bottom, tiles left to right, each tile row
// by row.  Positions outside the grid are skipped, so every index appears exactly once and,
// when 4 divides H and 16 divides W, each run of 64 slots is one tile.
inline std::vector<uint32_t> grid_tile_order(uint32_t W, uint32_t H) {
  std::vector<uint32_t> o;
  o.reserve(static_cast<size_t>(W) * H);
  for (uint32_t ty = 0; ty < H; ty += kGridTileRows)
    for (uint32_t tx = 0; tx < W; tx += kGridTileCols)
      for (uint32_t r = ty; r < std::min(H, ty + kGridTileRows); r++)
        for (uint32_t c = tx; c < std::min(W, tx + kGridTileCols); c++) o.push_back(r * W + c);
  return o;
}

}  // namespace rmh
