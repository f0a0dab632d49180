// Brick layout inside the HBM pool.
//
// getData() delivers bricks x fastest (what glTexSubImage3D consumes,
// TextureObject.cpp:107-109). In the pool a brick is stored as 4x4x4-voxel
// micro-tiles (64 voxels: one 128-byte cache line for 16-bit data), tiles
// x fastest, voxels x fastest inside a tile, each axis padded to a multiple
// of 4. A wavefront's 8x8 ray patch then touches a handful of cache lines
// from any view direction instead of up to 64 for views along x.
#pragma once

#include <cstdint>

namespace lvr
{
struct BrickLayout
{
    uint32_t dim[3];   // maximumBlockSize (valid voxels per axis)
    uint32_t tiles[3]; // ceil(dim / 4)
};

__host__ __device__ inline BrickLayout makeLayout(const uint32_t* dim)
{
    BrickLayout L;
    for (int a = 0; a < 3; ++a)
    {
        L.dim[a] = dim[a];
        L.tiles[a] = (dim[a] + 3u) / 4u;
    }
    return L;
}

// voxels stored per brick, padding included
__host__ __device__ inline uint32_t layoutVoxels(const BrickLayout& L)
{
    return L.tiles[0] * L.tiles[1] * L.tiles[2] * 64u;
}

__host__ __device__ inline uint32_t tiledIndex(const BrickLayout& L, uint32_t x, uint32_t y,
                                               uint32_t z)
{
    const uint32_t tile = ((z >> 2) * L.tiles[1] + (y >> 2)) * L.tiles[0] + (x >> 2);
    return (tile << 6) | ((z & 3u) << 4) | ((y & 3u) << 2) | (x & 3u);
}
} // namespace lvr
