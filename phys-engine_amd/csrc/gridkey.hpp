// Cell keys of the uniform grid, shared by the grid's key kernel and the
// fused AABB + key kernel of the collision step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pg {

__device__ __forceinline__ uint32_t spread3(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Key of a cell: its Morton code over the low k bits of each coordinate
// (mask = 2^k - 1), within the table's t bits (tmask = 2^t - 1).
__device__ __forceinline__ uint32_t cell_key(int32_t x, int32_t y, int32_t z, uint32_t mask,
                                             uint32_t tmask)
{
    return (spread3(uint32_t(x) & mask) | (spread3(uint32_t(y) & mask) << 1) |
            (spread3(uint32_t(z) & mask) << 2)) &
           tmask;
}

// Per-world key offset (0 for world 0): multiplicative hash to the t table bits.
__device__ __forceinline__ uint32_t world_key(uint32_t world, uint32_t table_bits)
{
    return world ? (world * 0x9E3779B1u) >> (32u - table_bits) : 0u;
}

// Sort-item key of a body that takes the overflow path: the radix pass drops it.
constexpr uint32_t kOverflowKey = 1u << 31;

// Home cell of a body; false when it must use the overflow path.
__device__ __forceinline__ bool home_cell(float x, float y, float z, double inv, int32_t cell[3])
{
    const double v[3] = {floor(double(x) * inv), floor(double(y) * inv), floor(double(z) * inv)};
    for (int a = 0; a < 3; ++a) {
        if (!(v[a] > -1073741824.0 && v[a] < 1073741824.0))
            return false;
        cell[a] = int32_t(v[a]);
    }
    return true;
}


// Key of the home cell (cell of the min corner) of a body with bounds b in
// world `world`, or kOverflowKey when the body is larger than the cell
// allows (some side above cell * (1 - 2^-9)) or a cell coordinate is out of
// range; inv = 1 / cell in double precision.
__device__ __forceinline__ uint32_t grid_cell_key(const float b[6], uint32_t world, float cell,
                                                  double inv, uint32_t bits_per_axis,
                                                  uint32_t table_bits)
{
    const float limit = cell * (1.0f - 1.0f / 512.0f);
    const bool fits = (b[3] - b[0]) <= limit && (b[4] - b[1]) <= limit && (b[5] - b[2]) <= limit;
    int32_t c[3], top[3];
    if (fits && home_cell(b[0], b[1], b[2], inv, c) && home_cell(b[3], b[4], b[5], inv, top))
        return cell_key(c[0], c[1], c[2], (1u << bits_per_axis) - 1u, (1u << table_bits) - 1u) ^
               world_key(world, table_bits);
    return kOverflowKey;
}

}  // namespace pg
