// tile_geom.h -- tile bookkeeping shared by the gfx950 kernels and the host scheduler.
//
// A W x H instance is cut into tiles of TW x TH cells in row-major order: tile (tx, ty) has
// the local id tx + tiles_x * ty (src/eikonal_hcm.cpp:209, block id = x + y * blocks_x) and
// the global id tile_base + local id.  Edge tiles are partial: tile (tx, ty) covers
// min(TW, W - tx TW) x min(TH, H - ty TH) cells (data/eikonal_ghcm_fim.cl:253).  Sides follow
// the reference's numbering (src/eikonal_hcm.cpp:18, data/eikonal_common.cl:361):
// 0 = x-, 1 = y-, 2 = x+, 3 = y+; the opposite of side s is s ^ 2.  A tile whose edge on
// side s changed activates the neighbour across s with the bit of the neighbour's own side
// facing it, 1 << (s ^ 2) (the reference ORs 1 << (8 (s ^ 2)) into a byte mask,
// src/eikonal_hcm.cpp:457-462; here one bit per side, and bit 4 opens every cell).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ghcm {

constexpr uint32_t kNoTile = 0xFFFFFFFFu;

__host__ __device__ inline uint32_t tiles_along(uint32_t cells, uint32_t tile) {
    return (cells + tile - 1) / tile;
}
__host__ __device__ inline int side_dx(int s) { return s == 0 ? -1 : s == 2 ? 1 : 0; }
__host__ __device__ inline int side_dy(int s) { return s == 1 ? -1 : s == 3 ? 1 : 0; }
__host__ __device__ inline int opposite_side(int s) { return s ^ 2; }
// Border bit set on the neighbour across side s of a tile whose edge on s changed.
__host__ __device__ inline uint32_t border_bit_across(int s) { return 1u << opposite_side(s); }
// Cells of a tile along one axis: the last tile of an axis may be partial.
__host__ __device__ inline int tile_extent(int cells, int origin, int tile) {
    return cells - origin < tile ? cells - origin : tile;
}
// Local id -> tile coordinates, row-major.
__host__ __device__ inline void tile_xy(uint32_t local, uint32_t tiles_x, int& tx, int& ty) {
    tx = (int)(local % tiles_x);
    ty = (int)(local / tiles_x);
}
// Global id of the neighbour across side s, or kNoTile at the instance's border.
__host__ __device__ inline uint32_t tile_neighbour(uint32_t tile_base, uint32_t tiles_x,
                                                   uint32_t tiles_y, int tx, int ty, int s) {
    const int nx = tx + side_dx(s), ny = ty + side_dy(s);
    if (nx < 0 || ny < 0 || nx >= (int)tiles_x || ny >= (int)tiles_y) return kNoTile;
    return tile_base + (uint32_t)ny * tiles_x + (uint32_t)nx;
}
// Global id of the tile holding cell (x, y).
__host__ __device__ inline uint32_t tile_of_cell(uint32_t tile_base, uint32_t tiles_x, uint32_t x,
                                                 uint32_t y, uint32_t tw, uint32_t th) {
    return tile_base + (y / th) * tiles_x + x / tw;
}

}  // namespace ghcm
