// Procedural value-noise DataSource: every voxel of every brick (any level,
// halo included) is a pure integer function of (seed, level, voxel
// coordinates), so bricks generated on the host and on the device are
// bit-identical, and the device can fill its HBM pool without PCIe traffic.
//
// Field: sum over octaves o of 2^-o * trilinear value noise with a smoothstep
// fade, lattice of base_cells * 2^o cells across the largest axis, in 16.16
// fixed point. A coarse level samples the field at its voxel centres and
// leaves out octaves finer than twice its voxel size (band limit). The sum
// is remapped through a monotone 4096-segment uint16 table (log-normal
// shaping, built on the host in double precision).
//
// Blobs plus value noise: the noise above blended, in integers, with a
// smooth field of Gaussian blobs (field_core.h) tabulated once on a lattice
// and read by trilinear interpolation in the same 16.16 fixed point, so
// the blend stays a pure integer function of the voxel as well.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <thread>
#include <vector>

#include "field_core.h"
#include "lvr_internal.h"

namespace lvr
{
namespace
{
__host__ __device__ inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__host__ __device__ inline int64_t lattice(uint64_t seed, int o, int64_t x,
                                           int64_t y, int64_t z)
{
    const uint64_t key = (uint64_t(x & 0xFFFFF)) | (uint64_t(y & 0xFFFFF) << 20) |
                         (uint64_t(z & 0xFFFFF) << 40) | (uint64_t(o) << 60);
    return int64_t(splitmix64(seed ^ splitmix64(key)) >> 48); // 0..65535
}

__host__ __device__ inline int64_t fade(int64_t f)
{
    // f^2 (3 - 2f) in 16-bit fixed point
    const uint64_t uf = uint64_t(f);
    return int64_t((uf * uf * (3ull * 65536ull - 2ull * uf)) >> 32);
}

__host__ __device__ inline int64_t lerp16(int64_t a, int64_t b, int64_t s)
{
    return a + (((b - a) * s) >> 16);
}

struct NoiseShape
{
    uint32_t voxels[3];
    uint32_t vmax;
    uint32_t depth;
    int32_t octaves;
    int32_t baseCells;
    uint64_t seed;
    int64_t maxSum;
    int64_t cellMult[8];   // floor(cells_o * 2^32 / vmax): cell coordinate scale
    int64_t cellFinest[8]; // finest voxels per cell of octave o
    int64_t segMult;       // floor(4096 * 2^32 / (maxSum + 1))
};

// Octave sum at a finest-voxel position in 16.16 fixed point.
__host__ __device__ inline int64_t octaveSum(const NoiseShape& s,
                                             const int64_t* ufp,
                                             int octaves)
{
    int64_t sum = 0;
    for (int o = 0; o < octaves; ++o)
    {
        int64_t ci[3], fr[3];
        for (int a = 0; a < 3; ++a)
        {
            const int64_t c = (ufp[a] * s.cellMult[o]) >> 32;
            ci[a] = c >> 16;
            fr[a] = fade(c & 0xFFFF);
        }
        int64_t v[8];
        for (int k = 0; k < 8; ++k)
            v[k] = lattice(s.seed, o, ci[0] + (k & 1), ci[1] + ((k >> 1) & 1),
                           ci[2] + ((k >> 2) & 1));
        const int64_t x00 = lerp16(v[0], v[1], fr[0]);
        const int64_t x10 = lerp16(v[2], v[3], fr[0]);
        const int64_t x01 = lerp16(v[4], v[5], fr[0]);
        const int64_t x11 = lerp16(v[6], v[7], fr[0]);
        const int64_t y0 = lerp16(x00, x10, fr[1]);
        const int64_t y1 = lerp16(x01, x11, fr[1]);
        sum += lerp16(y0, y1, fr[2]) >> o;
    }
    return sum;
}

// Voxel of a level-L brick at level-lattice coordinates g (may lie in the
// halo, outside the volume: clamped to the volume), remapped to 0..65535.
// Centre of level voxel g along one axis in finest voxels, 16.16 fixed
// point, clamped to the volume's first and last voxel centres.
__host__ __device__ inline int64_t voxelCentre16(int64_t g, int64_t S, uint32_t voxels)
{
    int64_t c = (2 * g + 1) * S * 32768; // (g + 0.5) * S in 16.16
    const int64_t lo = 32768;
    const int64_t hi = (2 * int64_t(voxels) - 1) * 32768;
    return c < lo ? lo : (c > hi ? hi : c);
}

__host__ __device__ inline uint32_t noiseVoxel(const NoiseShape& s,
                                               const uint16_t* remap,
                                               uint32_t level, int64_t gx,
                                               int64_t gy, int64_t gz)
{
    const uint32_t shift = s.depth - 1u - level;
    const int64_t S = int64_t(1) << shift; // finest voxels per level voxel
    const int64_t u[3] = {voxelCentre16(gx, S, s.voxels[0]), voxelCentre16(gy, S, s.voxels[1]),
                          voxelCentre16(gz, S, s.voxels[2])};
    // band limit: keep octaves whose cell spans at least 2 level voxels
    int octaves = 0;
    for (int o = 0; o < s.octaves; ++o)
    {
        if (s.cellFinest[o] >= 2 * S || o == 0)
            octaves = o + 1;
    }
    const int64_t sum = octaveSum(s, u, octaves);
    // 4096 segments over [0, maxSum]
    const int64_t t = (sum * s.segMult) >> 16;
    int64_t seg = t >> 16;
    const int64_t fr = t & 0xFFFF;
    if (seg > 4095)
        seg = 4095;
    const int64_t a = remap[seg], b = remap[seg + 1];
    int64_t v = a + (((b - a) * fr) >> 16);
    v = v < 0 ? 0 : (v > 65535 ? 65535 : v);
    return uint32_t(v);
}

template <typename T>
__host__ __device__ inline T convertNoise(uint32_t v);
template <>
__host__ __device__ inline uint8_t convertNoise<uint8_t>(uint32_t v)
{
    return uint8_t(v >> 8);
}
template <>
__host__ __device__ inline uint16_t convertNoise<uint16_t>(uint32_t v)
{
    return uint16_t(v);
}
template <>
__host__ __device__ inline float convertNoise<float>(uint32_t v)
{
    return float(v) * (2.0f / 65535.0f) - 1.0f;
}

void finishShape(NoiseShape& s)
{
    int64_t m = 0;
    for (int o = 0; o < 8; ++o)
    {
        const int64_t cells = int64_t(s.baseCells) << o;
        s.cellMult[o] = int64_t((uint64_t(cells) << 32) / s.vmax);
        s.cellFinest[o] = int64_t(s.vmax) / cells;
        if (o < s.octaves)
            m += 65535 >> o;
    }
    s.maxSum = m;
    // t = sum * 4096 / (maxSum + 1) in 16.16
    s.segMult = int64_t((uint64_t(4096) << 32) / uint64_t(m + 1));
}

NoiseShape makeShape(const lvr_volume_info& vi, const lvr_noise_params& p)
{
    NoiseShape s;
    for (int a = 0; a < 3; ++a)
        s.voxels[a] = vi.voxels[a];
    s.vmax = std::max(vi.voxels[0], std::max(vi.voxels[1], vi.voxels[2]));
    s.depth = vi.depth;
    s.octaves = p.octaves;
    s.baseCells = p.base_cells;
    s.seed = p.seed;
    finishShape(s);
    return s;
}

struct BrickGeo
{
    int64_t origin[3]; // level-lattice coordinate of the brick's first voxel
    uint32_t dim[3];   // maximumBlockSize
    uint32_t level;
};

BrickGeo brickGeo(const lvr_volume_info& vi, uint64_t id)
{
    const NodeKey k = decodeNode(id);
    BrickGeo g;
    const uint32_t pos[3] = {k.x, k.y, k.z};
    for (int a = 0; a < 3; ++a)
    {
        const uint32_t block = vi.max_block[a] - 2 * vi.overlap[a];
        g.origin[a] = int64_t(pos[a]) * block - int64_t(vi.overlap[a]);
        g.dim[a] = vi.max_block[a];
    }
    g.level = k.level;
    return g;
}

// The blob field of the blobs-plus-noise source on a lattice of `cells`
// cells across the largest axis ((cells + 1)^3 uint16 nodes; node i at
// i / cells of the largest axis), and its share of the blend.
struct BlobLattice
{
    int64_t cellMult; // floor(cells * 2^32 / vmax): cell coordinate scale
    int32_t cells;
    int32_t weight; // 0..256
};

// Trilinear interpolation of the lattice at a 16.16 finest-voxel position
// (x, then y, then z, as the noise octaves). u < vmax * 65536 on every axis,
// so the cell index stays below `cells` and its upper node in the table.
__host__ __device__ inline uint32_t blobValue(const BlobLattice& b, const uint16_t* table,
                                              const int64_t* u)
{
    int64_t ci[3], fr[3];
    for (int a = 0; a < 3; ++a)
    {
        const int64_t c = (u[a] * b.cellMult) >> 32;
        ci[a] = c >> 16;
        fr[a] = c & 0xFFFF;
    }
    const int64_t n = b.cells + 1;
    const uint16_t* t0 = table + (ci[2] * n + ci[1]) * n + ci[0];
    const uint16_t* t1 = t0 + n * n;
    const int64_t x00 = lerp16(t0[0], t0[1], fr[0]);
    const int64_t x10 = lerp16(t0[n], t0[n + 1], fr[0]);
    const int64_t x01 = lerp16(t1[0], t1[1], fr[0]);
    const int64_t x11 = lerp16(t1[n], t1[n + 1], fr[0]);
    const int64_t y0 = lerp16(x00, x10, fr[1]);
    const int64_t y1 = lerp16(x01, x11, fr[1]);
    return uint32_t(lerp16(y0, y1, fr[2]));
}

__host__ __device__ inline uint32_t blend(uint32_t blob, uint32_t noise, int32_t w)
{
    return (blob * uint32_t(w) + noise * uint32_t(256 - w) + 128u) >> 8;
}

template <typename T, bool BLOBS>
__global__ void noiseBricksKernel(NoiseShape s, const uint16_t* __restrict__ remap,
                                  BlobLattice b, const uint16_t* __restrict__ table,
                                  const BrickGeo* __restrict__ geos,
                                  const int32_t* __restrict__ slots,
                                  uint8_t* __restrict__ pool, size_t slotStride,
                                  BrickLayout L)
{
    const BrickGeo g = geos[blockIdx.y];
    T* out = reinterpret_cast<T*>(pool + size_t(slots[blockIdx.y]) * slotStride);
    const uint32_t n = g.dim[0] * g.dim[1] * g.dim[2];
    const int64_t S = int64_t(1) << (s.depth - 1u - g.level);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x)
    {
        const uint32_t x = i % g.dim[0];
        const uint32_t y = (i / g.dim[0]) % g.dim[1];
        const uint32_t z = i / (g.dim[0] * g.dim[1]);
        uint32_t v = noiseVoxel(s, remap, g.level, g.origin[0] + x, g.origin[1] + y,
                                g.origin[2] + z);
        if (BLOBS)
        {
            const int64_t u[3] = {voxelCentre16(g.origin[0] + x, S, s.voxels[0]),
                                  voxelCentre16(g.origin[1] + y, S, s.voxels[1]),
                                  voxelCentre16(g.origin[2] + z, S, s.voxels[2])};
            v = blend(blobValue(b, table, u), v, b.weight);
        }
        out[tiledIndex(L, x, y, z)] = convertNoise<T>(v);
    }
}

// The voxel passes of the separable host paths run in 32-bit unsigned
// integers, which the host compiler vectorises: for a, b, s in [0, 65535],
// lerp16(a, b, s) = a + floor((b - a) s / 2^16) = (a (65536 - s) + b s) >> 16
// (65536 a is a multiple of 2^16), and that sum stays below 2^32. Lattice
// values, fade weights and every lerp of them are 16-bit, and an octave sum
// stays below 2^17.
inline uint32_t lerpU16(uint32_t a, uint32_t b, uint32_t s)
{
    return (a * (65536u - s) + b * s) >> 16;
}

// acc[x] += lerp16(p0[x], p1[x], w) >> o over one row of voxels
inline void accumulateLerp(uint32_t* __restrict__ acc, const uint32_t* __restrict__ p0,
                           const uint32_t* __restrict__ p1, uint32_t w, int shift, int n)
{
    const uint32_t iw = 65536u - w;
    for (int x = 0; x < n; ++x)
        acc[x] += (p0[x] * iw + p1[x] * w) >> shift;
}

// out[x] = lerp16(p0[x], p1[x], w) over one row of voxels
inline void lerpRow(uint32_t* __restrict__ out, const uint32_t* __restrict__ p0,
                    const uint32_t* __restrict__ p1, uint32_t w, int n)
{
    const uint32_t iw = 65536u - w;
    for (int x = 0; x < n; ++x)
        out[x] = (p0[x] * iw + p1[x] * w) >> 16;
}

// Host generation of one brick, separable: per octave, the lattice values
// the brick touches are hashed once, interpolated along x into rows, the rows
// along y into planes, and each voxel takes the final lerp along z. Every
// value goes through the same integer operations in the same order as
// noiseVoxel (x, then y, then z), so the result equals the device brick.
template <typename T>
__attribute__((always_inline)) inline void noiseBrickHostBody(const NoiseShape& s,
                                                             const uint16_t* remap,
                                                             const BrickGeo& g, T* out)
{
    const uint32_t shift = s.depth - 1u - g.level;
    const int64_t S = int64_t(1) << shift;
    const int nx = int(g.dim[0]), ny = int(g.dim[1]), nz = int(g.dim[2]);
    int octaves = 0;
    for (int o = 0; o < s.octaves; ++o)
        if (s.cellFinest[o] >= 2 * S || o == 0)
            octaves = o + 1;
    // per axis and octave: lattice cell and fade weight of every voxel
    std::vector<int64_t> ci[3];
    std::vector<uint32_t> fr[3];
    const int n[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a)
    {
        ci[a].resize(size_t(octaves) * n[a]);
        fr[a].resize(size_t(octaves) * n[a]);
        for (int i = 0; i < n[a]; ++i)
        {
            int64_t c = (2 * (g.origin[a] + i) + 1) * S * 32768;
            const int64_t lo = 32768;
            const int64_t hi = (2 * int64_t(s.voxels[a]) - 1) * 32768;
            c = c < lo ? lo : (c > hi ? hi : c);
            for (int o = 0; o < octaves; ++o)
            {
                const int64_t cc = (c * s.cellMult[o]) >> 32;
                ci[a][size_t(o) * n[a] + i] = cc >> 16;
                fr[a][size_t(o) * n[a] + i] = uint32_t(fade(cc & 0xFFFF));
            }
        }
    }
    std::vector<uint32_t> sum(size_t(nx) * ny * nz, 0u);
    std::vector<uint32_t> lat, rows, planes;
    for (int o = 0; o < octaves; ++o)
    {
        const int64_t* cx = &ci[0][size_t(o) * nx];
        const int64_t* cy = &ci[1][size_t(o) * ny];
        const int64_t* cz = &ci[2][size_t(o) * nz];
        const uint32_t* fx = &fr[0][size_t(o) * nx];
        const uint32_t* fy = &fr[1][size_t(o) * ny];
        const uint32_t* fz = &fr[2][size_t(o) * nz];
        // lattice range touched (cells are monotone in the voxel index)
        const int64_t x0 = cx[0], y0 = cy[0], z0 = cz[0];
        const int lx = int(cx[nx - 1] - x0) + 2, ly = int(cy[ny - 1] - y0) + 2,
                  lz = int(cz[nz - 1] - z0) + 2;
        lat.resize(size_t(lx) * ly * lz);
        for (int k = 0; k < lz; ++k)
            for (int j = 0; j < ly; ++j)
                for (int i = 0; i < lx; ++i)
                    lat[(size_t(k) * ly + j) * lx + i] =
                        uint32_t(lattice(s.seed, o, x0 + i, y0 + j, z0 + k));
        // x-interpolated rows for every lattice (y, z) and voxel x
        rows.resize(size_t(lz) * ly * nx);
        for (int k = 0; k < lz; ++k)
            for (int j = 0; j < ly; ++j)
            {
                const uint32_t* L = &lat[(size_t(k) * ly + j) * lx];
                uint32_t* R = &rows[(size_t(k) * ly + j) * nx];
                for (int x = 0; x < nx; ++x)
                {
                    const int i = int(cx[x] - x0);
                    R[x] = lerpU16(L[i], L[i + 1], fx[x]);
                }
            }
        // y-interpolated planes for every lattice z and voxel (x, y)
        planes.resize(size_t(lz) * ny * nx);
        for (int k = 0; k < lz; ++k)
            for (int y = 0; y < ny; ++y)
            {
                const int j = int(cy[y] - y0);
                lerpRow(&planes[(size_t(k) * ny + y) * nx], &rows[(size_t(k) * ly + j) * nx],
                        &rows[(size_t(k) * ly + j + 1) * nx], fy[y], nx);
            }
        // z lerp per voxel, accumulated like octaveSum
        for (int z = 0; z < nz; ++z)
        {
            const int k = int(cz[z] - z0);
            for (int y = 0; y < ny; ++y)
                accumulateLerp(&sum[(size_t(z) * ny + y) * nx],
                               &planes[(size_t(k) * ny + y) * nx],
                               &planes[(size_t(k + 1) * ny + y) * nx], fz[z], 16 + o, nx);
        }
    }
    for (size_t i = 0; i < sum.size(); ++i)
    {
        const int64_t t = (int64_t(sum[i]) * s.segMult) >> 16;
        int64_t seg = t >> 16;
        const int64_t f = t & 0xFFFF;
        if (seg > 4095)
            seg = 4095;
        const int64_t a = remap[seg], b = remap[seg + 1];
        int64_t v = a + (((b - a) * f) >> 16);
        v = v < 0 ? 0 : (v > 65535 ? 65535 : v);
        out[i] = convertNoise<T>(uint32_t(v));
    }
}

// Blob values of a brick on the host, separable as noiseBrickHost: the
// lattice nodes the brick touches are interpolated along x into rows, the
// rows along y into planes, and each voxel takes the lerp along z; the same
// integer operations in the same order as blobValue (in 32-bit form, as
// lerpU16 states).
__attribute__((always_inline)) inline void blobBrickHostBody(const NoiseShape& s,
                                                            const BlobLattice& b,
                                                            const uint16_t* table,
                                                            const BrickGeo& g,
                                                            std::vector<uint32_t>& out)
{
    const int64_t S = int64_t(1) << (s.depth - 1u - g.level);
    const int n[3] = {int(g.dim[0]), int(g.dim[1]), int(g.dim[2])};
    std::vector<int64_t> ci[3];
    std::vector<uint32_t> fr[3];
    for (int a = 0; a < 3; ++a)
    {
        ci[a].resize(size_t(n[a]));
        fr[a].resize(size_t(n[a]));
        for (int i = 0; i < n[a]; ++i)
        {
            const int64_t c = (voxelCentre16(g.origin[a] + i, S, s.voxels[a]) * b.cellMult) >> 32;
            ci[a][size_t(i)] = c >> 16;
            fr[a][size_t(i)] = uint32_t(c & 0xFFFF);
        }
    }
    const int nx = n[0], ny = n[1], nz = n[2];
    const int64_t m = b.cells + 1;
    const int64_t x0 = ci[0][0], y0 = ci[1][0], z0 = ci[2][0];
    const int lx = int(ci[0][size_t(nx - 1)] - x0) + 2, ly = int(ci[1][size_t(ny - 1)] - y0) + 2,
              lz = int(ci[2][size_t(nz - 1)] - z0) + 2;
    std::vector<uint32_t> rows(size_t(lz) * ly * nx), planes(size_t(lz) * ny * nx);
    for (int k = 0; k < lz; ++k)
        for (int j = 0; j < ly; ++j)
        {
            const uint16_t* T = table + ((z0 + k) * m + (y0 + j)) * m + x0;
            uint32_t* R = &rows[(size_t(k) * ly + j) * nx];
            for (int x = 0; x < nx; ++x)
            {
                const int i = int(ci[0][size_t(x)] - x0);
                R[x] = lerpU16(T[i], T[i + 1], fr[0][size_t(x)]);
            }
        }
    for (int k = 0; k < lz; ++k)
        for (int y = 0; y < ny; ++y)
        {
            const int j = int(ci[1][size_t(y)] - y0);
            const uint32_t* R0 = &rows[(size_t(k) * ly + j) * nx];
            lerpRow(&planes[(size_t(k) * ny + y) * nx], R0, R0 + nx, fr[1][size_t(y)], nx);
        }
    out.resize(size_t(nx) * ny * nz);
    for (int z = 0; z < nz; ++z)
    {
        const int k = int(ci[2][size_t(z)] - z0);
        for (int y = 0; y < ny; ++y)
            lerpRow(&out[(size_t(z) * ny + y) * nx], &planes[(size_t(k) * ny + y) * nx],
                    &planes[(size_t(k + 1) * ny + y) * nx], fr[2][size_t(z)], nx);
    }
}

// Host brick of the blobs-plus-noise source: the separable noise brick (as
// 16-bit values, which convertNoise leaves unchanged) blended with the
// separable blob values.
template <typename T>
__attribute__((always_inline)) inline void blobNoiseBrickHostBody(
    const NoiseShape& s, const uint16_t* remap, const BlobLattice& b, const uint16_t* table,
    const BrickGeo& g, T* out)
{
    std::vector<uint16_t> noise(size_t(g.dim[0]) * g.dim[1] * g.dim[2]);
    noiseBrickHostBody(s, remap, g, noise.data());
    std::vector<uint32_t> blob;
    blobBrickHostBody(s, b, table, g, blob);
    for (size_t i = 0; i < noise.size(); ++i)
        out[i] = convertNoise<T>(blend(blob[i], noise[i], b.weight));
}

// The host bricks compiled twice, for the x86-64 baseline and with AVX2
// (eight 32-bit lanes per voxel pass, about 1.4x faster per brick), picked at
// run time from the CPU; both give the same integers.
#if defined(__HIP_DEVICE_COMPILE__)
#define LVR_HOST_AVX2
#else
#define LVR_HOST_AVX2 __attribute__((target("avx2")))
#endif

// LIVRE_AMD_HOST_BASELINE=1 selects the baseline build on any CPU (tests
// compare the two).
inline bool detectHostAvx2()
{
#if !defined(__HIP_DEVICE_COMPILE__)
    // the CPU model may not be set up yet in a library opened at run time
    // (host pass only: the device pass parses this host function too)
    __builtin_cpu_init();
#endif
    return __builtin_cpu_supports("avx2") && !std::getenv("LIVRE_AMD_HOST_BASELINE");
}

inline bool hostHasAvx2()
{
    static const bool has = detectHostAvx2();
    return has;
}

template <typename T>
LVR_HOST_AVX2 void noiseBrickHostAvx2(const NoiseShape& s, const uint16_t* remap,
                                      const BrickGeo& g, T* out)
{
    noiseBrickHostBody(s, remap, g, out);
}

template <typename T>
void noiseBrickHost(const NoiseShape& s, const uint16_t* remap, const BrickGeo& g, T* out)
{
    if (hostHasAvx2())
        noiseBrickHostAvx2(s, remap, g, out);
    else
        noiseBrickHostBody(s, remap, g, out);
}

template <typename T>
LVR_HOST_AVX2 void blobNoiseBrickHostAvx2(const NoiseShape& s, const uint16_t* remap,
                                          const BlobLattice& b, const uint16_t* table,
                                          const BrickGeo& g, T* out)
{
    blobNoiseBrickHostBody(s, remap, b, table, g, out);
}

template <typename T>
void blobNoiseBrickHost(const NoiseShape& s, const uint16_t* remap, const BlobLattice& b,
                        const uint16_t* table, const BrickGeo& g, T* out)
{
    if (hostHasAvx2())
        blobNoiseBrickHostAvx2(s, remap, b, table, g, out);
    else
        blobNoiseBrickHostBody(s, remap, b, table, g, out);
}

} // namespace
} // namespace lvr

// Blobs plus value noise: the volume it was made for, the noise parameters,
// and the blob field tabulated on its lattice.
struct lvr_blobnoise
{
    lvr_volume_info vi;
    lvr_noise_params noise;
    lvr::BlobLattice lattice;
    std::vector<uint16_t> table;
};

namespace lvr
{
// Device generation of noise bricks (bn == nullptr) or blobs-plus-noise
// bricks straight into pool slots, in chunks of 4096.
int noiseBricksDevice(lvr_ctx* ctx, const lvr_noise_params& np, const lvr_blobnoise* bn,
                      const uint64_t* node_ids, int n)
{
    const lvr_noise_params* p = &np;
    const int dt = ctx->vi.data_type;
    if (dt != LVR_DT_UINT8 && dt != LVR_DT_UINT16 && dt != LVR_DT_FLOAT)
        return fail(ctx, LVR_EINVAL, "noise source supports u8, u16, f32");
    int rc = ensureDevice(ctx);
    if (rc)
        return rc;
    const NoiseShape s = makeShape(ctx->vi, *p);
    const int chunk = 4096;
    const size_t remapBytes = (sizeof p->remap + 255) & ~size_t(255);
    const size_t geoBytes = (sizeof(BrickGeo) * chunk + 255) & ~size_t(255);
    const size_t tableBytes = bn ? (sizeof(uint16_t) * bn->table.size() + 255) & ~size_t(255) : 0;
    uint8_t* scratch = static_cast<uint8_t*>(
        genScratch(ctx, remapBytes + tableBytes + geoBytes + sizeof(int32_t) * chunk));
    if (!scratch)
        return LVR_ENOMEM;
    uint16_t* d_remap = reinterpret_cast<uint16_t*>(scratch);
    uint16_t* d_table = reinterpret_cast<uint16_t*>(scratch + remapBytes);
    BrickGeo* d_geo = reinterpret_cast<BrickGeo*>(scratch + remapBytes + tableBytes);
    int32_t* d_slots =
        reinterpret_cast<int32_t*>(scratch + remapBytes + tableBytes + geoBytes);
    const BlobLattice lat = bn ? bn->lattice : BlobLattice{0, 0, 0};
    // slots taken here may still be read by batches on other streams
    rc = waitPoolReaders(ctx, ctx->stream);
    if (rc)
        return rc;
    LVR_HIP(ctx, hipMemcpyAsync(d_remap, p->remap, sizeof p->remap, hipMemcpyHostToDevice,
                                ctx->stream));
    if (bn)
        LVR_HIP(ctx, hipMemcpyAsync(d_table, bn->table.data(), sizeof(uint16_t) * bn->table.size(),
                                    hipMemcpyHostToDevice, ctx->stream));
    std::vector<BrickGeo> geo;
    std::vector<int32_t> slots;
    for (int base = 0; base < n && rc == LVR_OK; base += chunk)
    {
        const int cnt = std::min(chunk, n - base);
        geo.resize(cnt);
        slots.resize(cnt);
        int taken = 0;
        for (int i = 0; i < cnt && rc == LVR_OK; ++i)
        {
            if (decodeNode(node_ids[base + i]).level >= ctx->vi.depth)
                rc = fail(ctx, LVR_EINVAL, "node id outside the octree");
            else
            {
                geo[i] = brickGeo(ctx->vi, node_ids[base + i]);
                rc = acquireSlotFor(ctx, node_ids[base + i], &slots[i]);
                taken += rc == LVR_OK;
            }
        }
        if (rc == LVR_OK &&
            (hipMemcpyAsync(d_geo, geo.data(), sizeof(BrickGeo) * cnt,
                            hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
             hipMemcpyAsync(d_slots, slots.data(), sizeof(int32_t) * cnt,
                            hipMemcpyHostToDevice, ctx->stream) != hipSuccess))
            rc = fail(ctx, LVR_EHIP, "noise upload");
        if (rc)
        {
            // bricks of this chunk never received their voxels
            const std::string msg = ctx->error;
            for (int i = 0; i < taken; ++i)
                (void)releaseSlot(ctx, node_ids[base + i]);
            ctx->error = msg;
            break;
        }
        const dim3 grid(64, unsigned(cnt));
#define LVR_NOISE(T, B)                                                                     \
    hipLaunchKernelGGL((noiseBricksKernel<T, B>), grid, dim3(256), 0, ctx->stream, s,       \
                       d_remap, lat, d_table, d_geo, d_slots, ctx->d_pool, ctx->slotStride,     \
                       ctx->layout)
        if (dt == LVR_DT_UINT8)
        {
            if (bn)
                LVR_NOISE(uint8_t, true);
            else
                LVR_NOISE(uint8_t, false);
        }
        else if (dt == LVR_DT_UINT16)
        {
            if (bn)
                LVR_NOISE(uint16_t, true);
            else
                LVR_NOISE(uint16_t, false);
        }
        else
        {
            if (bn)
                LVR_NOISE(float, true);
            else
                LVR_NOISE(float, false);
        }
#undef LVR_NOISE
        if (hipGetLastError() != hipSuccess)
            rc = fail(ctx, LVR_EHIP, "noise kernel");
        // each brick's histogram into the cache as it lands in its slot
        if (rc == LVR_OK)
            rc = cacheBrickHistograms(ctx, node_ids + base, slots.data(), cnt, ctx->stream);
        // the host vectors and the scratch are reused by the next chunk
        if (rc == LVR_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = fail(ctx, LVR_EHIP, "noise kernel");
    }
    return rc;
}
} // namespace lvr

extern "C" {

int lvr_noise_params_init(lvr_noise_params* p, uint64_t seed, int octaves,
                          int base_cells, float gain)
{
    if (!p || octaves < 1 || octaves > 8 || base_cells < 1)
        return LVR_EINVAL;
    std::memset(p, 0, sizeof *p);
    p->seed = seed;
    p->octaves = octaves;
    p->base_cells = base_cells;
    // Statistics of the octave sum from 65536 lattice-independent sample
    // points of a unit-shaped volume, then y = exp(gain * z) scaled to
    // [0, 65535] between the 0.01% and 99.99% points.
    lvr::NoiseShape s;
    s.voxels[0] = s.voxels[1] = s.voxels[2] = 1u << 20;
    s.vmax = 1u << 20;
    s.depth = 1;
    s.octaves = octaves;
    s.baseCells = base_cells;
    s.seed = seed;
    lvr::finishShape(s);
    const int64_t m = s.maxSum;
    std::vector<double> samples;
    samples.reserve(65536);
    uint64_t r = seed ^ 0xA5A5A5A5ull;
    for (int i = 0; i < 65536; ++i)
    {
        int64_t u[3];
        for (int a = 0; a < 3; ++a)
        {
            r = r * 6364136223846793005ull + 1442695040888963407ull;
            u[a] = int64_t((r >> 24) % (uint64_t(s.vmax) << 16));
        }
        samples.push_back(double(lvr::octaveSum(s, u, octaves)));
    }
    double mean = 0, var = 0;
    for (double v : samples)
        mean += v;
    mean /= samples.size();
    for (double v : samples)
        var += (v - mean) * (v - mean);
    const double sd = std::sqrt(var / samples.size()) + 1e-9;
    std::sort(samples.begin(), samples.end());
    const double lo = samples[samples.size() / 10000];
    const double hi = samples[samples.size() - 1 - samples.size() / 10000];
    const double elo = std::exp(gain * (lo - mean) / sd);
    const double ehi = std::exp(gain * (hi - mean) / sd);
    for (int k = 0; k <= 4096; ++k)
    {
        const double x = double(k) * double(m + 1) / 4096.0;
        const double xc = std::min(hi, std::max(lo, x));
        const double y = (std::exp(gain * (xc - mean) / sd) - elo) / (ehi - elo);
        p->remap[k] = uint16_t(std::lround(std::min(1.0, std::max(0.0, y)) * 65535.0));
    }
    return LVR_OK;
}

int lvr_noise_brick_host(const lvr_volume_info* info, const lvr_noise_params* p,
                         uint64_t node_id, void* out, size_t bytes)
{
    if (!info || !p || !out)
        return LVR_EINVAL;
    lvr_volume_info vi = *info;
    if (lvr_fill_regular_volume_info(&vi) != LVR_OK)
        return LVR_EINVAL;
    const size_t n = size_t(vi.max_block[0]) * vi.max_block[1] * vi.max_block[2];
    if (bytes != n * lvr::bytesPerVoxel(vi.data_type))
        return LVR_EINVAL;
    if (lvr::decodeNode(node_id).level >= vi.depth)
        return LVR_EINVAL;
    const lvr::NoiseShape s = lvr::makeShape(vi, *p);
    const lvr::BrickGeo g = lvr::brickGeo(vi, node_id);
    switch (vi.data_type)
    {
    case LVR_DT_UINT8:
        lvr::noiseBrickHost(s, p->remap, g, static_cast<uint8_t*>(out));
        return LVR_OK;
    case LVR_DT_UINT16:
        lvr::noiseBrickHost(s, p->remap, g, static_cast<uint16_t*>(out));
        return LVR_OK;
    case LVR_DT_FLOAT:
        lvr::noiseBrickHost(s, p->remap, g, static_cast<float*>(out));
        return LVR_OK;
    default:
        return LVR_EINVAL;
    }
}

int lvr_noise_bricks_device(lvr_ctx* ctx, const lvr_noise_params* p,
                            const uint64_t* node_ids, int n)
{
    if (!ctx || !p || n < 0 || (n > 0 && !node_ids))
        return LVR_EINVAL;
    return lvr::noiseBricksDevice(ctx, *p, nullptr, node_ids, n);
}

int lvr_blobnoise_create(const lvr_volume_info* info, const lvr_noise_params* noise,
                         uint64_t blob_seed, int blob_count, int blob_cells, int blob_weight,
                         lvr_blobnoise** out)
{
    if (!info || !noise || !out || blob_count < 1 || blob_count > 4096 || blob_cells < 1 ||
        blob_cells > 1024 || blob_weight < 0 || blob_weight > 256)
        return LVR_EINVAL;
    *out = nullptr;
    lvr_volume_info vi = *info;
    if (lvr_fill_regular_volume_info(&vi) != LVR_OK)
        return LVR_EINVAL;
    const int dt = vi.data_type;
    if (dt != LVR_DT_UINT8 && dt != LVR_DT_UINT16 && dt != LVR_DT_FLOAT)
        return LVR_EINVAL;
    lvr_blobnoise* bn = nullptr;
    try
    {
        bn = new lvr_blobnoise;
        bn->vi = vi;
        bn->noise = *noise;
        const uint32_t vmax = std::max(vi.voxels[0], std::max(vi.voxels[1], vi.voxels[2]));
        bn->lattice.cells = blob_cells;
        bn->lattice.weight = blob_weight;
        bn->lattice.cellMult = int64_t((uint64_t(blob_cells) << 32) / vmax);
        lvr_field_params fp;
        std::memset(&fp, 0, sizeof fp);
        fp.kind = LVR_FIELD_BLOBS;
        fp.count = blob_count;
        fp.seed = blob_seed;
        fp.noise = 0.f;
        const lvr::field::FieldSetup f = lvr::field::makeField(vi, fp);
        const size_t m = size_t(blob_cells) + 1;
        bn->table.resize(m * m * m);
        // node (i, j, k) at (i, j, k) / cells of the largest axis; threads
        // over z planes
        const unsigned nt =
            std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> workers;
        for (unsigned t = 0; t < nt; ++t)
            workers.emplace_back([&f, bn, m, t, nt, blob_cells] {
                for (size_t k = t; k < m; k += nt)
                    for (size_t j = 0; j < m; ++j)
                        for (size_t i = 0; i < m; ++i)
                        {
                            const float q[3] = {float(i) / float(blob_cells),
                                                float(j) / float(blob_cells),
                                                float(k) / float(blob_cells)};
                            bn->table[(k * m + j) * m + i] = lvr::field::quantise<uint16_t>(
                                f.shape, lvr::field::fieldValue(f.shape, f.blobs.data(), nullptr,
                                                                q, 0));
                        }
            });
        for (std::thread& w : workers)
            w.join();
    }
    catch (...)
    {
        delete bn;
        return LVR_ENOMEM;
    }
    *out = bn;
    return LVR_OK;
}

void lvr_blobnoise_destroy(lvr_blobnoise* bn)
{
    delete bn;
}

int lvr_blobnoise_brick_host(const lvr_blobnoise* bn, uint64_t node_id, void* out,
                             size_t bytes)
{
    if (!bn || !out)
        return LVR_EINVAL;
    const lvr_volume_info& vi = bn->vi;
    const size_t n = size_t(vi.max_block[0]) * vi.max_block[1] * vi.max_block[2];
    if (bytes != n * lvr::bytesPerVoxel(vi.data_type) || lvr::decodeNode(node_id).level >= vi.depth)
        return LVR_EINVAL;
    const lvr::NoiseShape s = lvr::makeShape(vi, bn->noise);
    const lvr::BrickGeo g = lvr::brickGeo(vi, node_id);
    const uint16_t* remap = bn->noise.remap;
    switch (vi.data_type)
    {
    case LVR_DT_UINT8:
        lvr::blobNoiseBrickHost(s, remap, bn->lattice, bn->table.data(), g,
                                static_cast<uint8_t*>(out));
        return LVR_OK;
    case LVR_DT_UINT16:
        lvr::blobNoiseBrickHost(s, remap, bn->lattice, bn->table.data(), g,
                                static_cast<uint16_t*>(out));
        return LVR_OK;
    default:
        lvr::blobNoiseBrickHost(s, remap, bn->lattice, bn->table.data(), g,
                                static_cast<float*>(out));
        return LVR_OK;
    }
}

int lvr_blobnoise_bricks_device(lvr_ctx* ctx, const lvr_blobnoise* bn, const uint64_t* node_ids,
                                int n)
{
    if (!ctx || !bn || n < 0 || (n > 0 && !node_ids))
        return LVR_EINVAL;
    const lvr_volume_info &a = ctx->vi, &b = bn->vi;
    bool same = a.data_type == b.data_type && a.depth == b.depth;
    for (int i = 0; i < 3; ++i)
        same = same && a.voxels[i] == b.voxels[i] && a.max_block[i] == b.max_block[i] &&
               a.overlap[i] == b.overlap[i];
    if (!same)
        return lvr::fail(ctx, LVR_EINVAL, "blob-noise source made for another volume");
    return lvr::noiseBricksDevice(ctx, bn->noise, bn, node_ids, n);
}

int lvr_blobnoise_source(void* user, uint64_t node_id, void* out, size_t bytes)
{
    return lvr_blobnoise_brick_host(static_cast<const lvr_blobnoise*>(user), node_id, out,
                                    bytes);
}
}
