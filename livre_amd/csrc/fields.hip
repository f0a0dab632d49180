// Procedural DataSource plugins for the blob and smooth-field volumes.
//
// Like the value-noise source, every voxel of every brick (any level, halo
// included) is a pure function of (parameters, level, voxel coordinates) and
// is evaluated at the voxel's centre, so bricks are independent and can be
// produced on the device straight into the pool. Only +, -, *, / and floor
// are used (exp and sin are evaluated by polynomials written here, not by
// the math library), compiled without contraction, so a brick computed on the
// host equals the one computed on the device.
//
//  - blobs: sum of anisotropic, rotated Gaussians a * exp(-d^2 / 2) (d the
//    distance in units of sigma along the blob's principal axes, cut off at
//    d^2 = 100), plus uniform noise from a voxel hash, scaled so that the
//    maximum found on a probe grid maps to the type's top.
//  - sines: mean of `count` plane waves sin(k . p + phi) in [-1, 1], with a
//    polynomial sine after range reduction.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "lvr_internal.h"

namespace lvr
{
namespace
{
struct Blob
{
    float c[3];
    float rot[9]; // rows: principal axes
    float inv[3]; // 1 / sigma per axis
    float amp;
};

struct Wave
{
    float k[3];
    float phi;
};

struct FieldShape
{
    int32_t kind;
    int32_t count;
    uint32_t voxels[3];
    uint32_t vmax;
    uint32_t depth;
    float noise;
    float scale; // blobs: value -> type range
    uint64_t seed;
};

__host__ __device__ inline uint64_t mix64(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    return x ^ (x >> 33);
}

// exp(x) for x in [-50, 0]: Cody-Waite reduction x = n ln2 + r, |r| <= ln2/2,
// degree-7 Taylor polynomial of exp(r), scaled by 2^n (relative error about
// 1e-7 over the range, float rounding included)
__host__ __device__ inline float polyExpNeg(float x)
{
    const float n = floorf(x * 1.44269504088896341f + 0.5f);
    const float r = (x - n * 0.693145751953125f) - n * 1.42860682030941723e-06f;
    float p = 1.98412698412698e-04f;
    p = p * r + 1.38888888888889e-03f;
    p = p * r + 8.33333333333333e-03f;
    p = p * r + 4.16666666666667e-02f;
    p = p * r + 1.66666666666667e-01f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    return ldexpf(p, int(n));
}

// sin(x) for any finite x: reduce to [-pi, pi], odd Taylor polynomial to x^11
__host__ __device__ inline float polySin(float x)
{
    const float twoPi = 6.28318530717958647692f;
    const float invTwoPi = 0.15915494309189533577f;
    x = x - twoPi * floorf(x * invTwoPi + 0.5f);
    const float x2 = x * x;
    float p = -2.50521083854417e-08f;
    p = p * x2 + 2.75573192239859e-06f;
    p = p * x2 - 1.98412698412698e-04f;
    p = p * x2 + 8.33333333333333e-03f;
    p = p * x2 - 1.66666666666667e-01f;
    return x + x * x2 * p;
}

__host__ __device__ inline float fieldValue(const FieldShape& s, const Blob* blobs,
                                            const Wave* waves, const float p[3],
                                            uint64_t voxelKey)
{
    if (s.kind == LVR_FIELD_SINES)
    {
        float sum = 0.f;
        for (int i = 0; i < s.count; ++i)
        {
            const Wave& w = waves[i];
            sum = sum + polySin(w.k[0] * p[0] + w.k[1] * p[1] + w.k[2] * p[2] + w.phi);
        }
        return sum / float(s.count);
    }
    float v = 0.f;
    for (int i = 0; i < s.count; ++i)
    {
        const Blob& b = blobs[i];
        const float d[3] = {p[0] - b.c[0], p[1] - b.c[1], p[2] - b.c[2]};
        float r2 = 0.f;
        for (int a = 0; a < 3; ++a)
        {
            const float q = (b.rot[3 * a] * d[0] + b.rot[3 * a + 1] * d[1] +
                             b.rot[3 * a + 2] * d[2]) *
                            b.inv[a];
            r2 = r2 + q * q;
        }
        if (r2 < 100.f)
            v = v + b.amp * polyExpNeg(-0.5f * r2);
    }
    const float u = float(mix64(s.seed ^ voxelKey) >> 40) / 16777216.0f; // [0, 1)
    return v + s.noise * u;
}

template <typename T>
__host__ __device__ inline T quantise(const FieldShape& s, float v);
template <>
__host__ __device__ inline uint8_t quantise<uint8_t>(const FieldShape& s, float v)
{
    const float q = floorf(v * s.scale * 255.0f + 0.5f);
    return uint8_t(q < 0.f ? 0.f : (q > 255.f ? 255.f : q));
}
template <>
__host__ __device__ inline uint16_t quantise<uint16_t>(const FieldShape& s, float v)
{
    const float q = floorf(v * s.scale * 65535.0f + 0.5f);
    return uint16_t(q < 0.f ? 0.f : (q > 65535.f ? 65535.f : q));
}
template <>
__host__ __device__ inline float quantise<float>(const FieldShape& s, float v)
{
    return s.kind == LVR_FIELD_SINES ? v : v * s.scale;
}

// Voxel (level, level-lattice coordinates g, may lie in the halo) of the
// field, sampled at its centre, clamped to the volume.
template <typename T>
__host__ __device__ inline T fieldVoxel(const FieldShape& s, const Blob* blobs,
                                        const Wave* waves, uint32_t level, const int64_t* g)
{
    const int64_t S = int64_t(1) << (s.depth - 1u - level);
    float p[3];
    uint64_t key = 0;
    for (int a = 0; a < 3; ++a)
    {
        int64_t u2 = (2 * g[a] + 1) * S; // twice the centre in finest voxels
        const int64_t hi = 2 * int64_t(s.voxels[a]) - 1;
        u2 = u2 < 1 ? 1 : (u2 > hi ? hi : u2);
        p[a] = (float(u2) * 0.5f) / float(s.vmax);
        key = key * 0x9E3779B97F4A7C15ull + uint64_t(u2) + (uint64_t(level) << 56);
    }
    return quantise<T>(s, fieldValue(s, blobs, waves, p, key));
}

struct FieldSetup
{
    FieldShape shape;
    std::vector<Blob> blobs;
    std::vector<Wave> waves;
};

FieldSetup makeField(const lvr_volume_info& vi, const lvr_field_params& p)
{
    FieldSetup f;
    std::memset(&f.shape, 0, sizeof f.shape);
    f.shape.kind = p.kind;
    f.shape.count = p.count;
    for (int a = 0; a < 3; ++a)
        f.shape.voxels[a] = vi.voxels[a];
    f.shape.vmax = std::max(vi.voxels[0], std::max(vi.voxels[1], vi.voxels[2]));
    f.shape.depth = vi.depth;
    f.shape.noise = p.noise;
    f.shape.seed = p.seed;
    f.shape.scale = 1.f;
    std::mt19937_64 rng(p.seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::normal_distribution<double> N(0.0, 1.0);
    const double ext[3] = {double(vi.voxels[0]) / f.shape.vmax,
                           double(vi.voxels[1]) / f.shape.vmax,
                           double(vi.voxels[2]) / f.shape.vmax};
    if (p.kind == LVR_FIELD_SINES)
    {
        for (int i = 0; i < p.count; ++i)
        {
            Wave w;
            double k[3], len = 0;
            for (int a = 0; a < 3; ++a)
            {
                k[a] = N(rng);
                len += k[a] * k[a];
            }
            len = std::sqrt(len);
            const double freq = 2.0 * M_PI * (0.5 + 2.5 * U(rng)); // low frequencies
            for (int a = 0; a < 3; ++a)
                w.k[a] = float(k[a] / len * freq);
            w.phi = float(2.0 * M_PI * U(rng));
            f.waves.push_back(w);
        }
        return f;
    }
    for (int i = 0; i < p.count; ++i)
    {
        Blob b;
        for (int a = 0; a < 3; ++a)
        {
            b.c[a] = float((0.1 + 0.8 * U(rng)) * ext[a]);
            b.inv[a] = float(1.0 / (0.02 + 0.10 * U(rng)));
        }
        // random rotation: Gram-Schmidt of a Gaussian matrix
        double m[3][3];
        for (auto& row : m)
            for (double& v : row)
                v = N(rng);
        for (int r = 0; r < 3; ++r)
        {
            for (int q = 0; q < r; ++q)
            {
                double d = 0;
                for (int a = 0; a < 3; ++a)
                    d += m[r][a] * m[q][a];
                for (int a = 0; a < 3; ++a)
                    m[r][a] -= d * m[q][a];
            }
            double l = 0;
            for (int a = 0; a < 3; ++a)
                l += m[r][a] * m[r][a];
            l = std::sqrt(l);
            for (int a = 0; a < 3; ++a)
                b.rot[3 * r + a] = float(m[r][a] / l);
        }
        b.amp = float(0.3 + 0.7 * U(rng));
        f.blobs.push_back(b);
    }
    // scale so that the largest value on a 48^3 probe grid maps to 1
    float vmax = 0.f;
    const int G = 48;
    for (int z = 0; z < G; ++z)
        for (int y = 0; y < G; ++y)
            for (int x = 0; x < G; ++x)
            {
                const float pt[3] = {float((x + 0.5) / G * ext[0]), float((y + 0.5) / G * ext[1]),
                                     float((z + 0.5) / G * ext[2])};
                vmax = std::max(vmax, fieldValue(f.shape, f.blobs.data(), nullptr, pt, 0) -
                                          f.shape.noise);
            }
    f.shape.scale = 1.f / (vmax + f.shape.noise);
    return f;
}

struct Geo
{
    int64_t origin[3];
    uint32_t level;
};

Geo brickOrigin(const lvr_volume_info& vi, uint64_t id)
{
    const NodeKey k = decodeNode(id);
    const uint32_t pos[3] = {k.x, k.y, k.z};
    Geo g;
    for (int a = 0; a < 3; ++a)
        g.origin[a] = int64_t(pos[a]) * (vi.max_block[a] - 2 * vi.overlap[a]) -
                      int64_t(vi.overlap[a]);
    g.level = k.level;
    return g;
}

template <typename T>
__global__ void fieldBricksKernel(FieldShape s, const Blob* __restrict__ blobs,
                                  const Wave* __restrict__ waves, const Geo* __restrict__ geos,
                                  const int32_t* __restrict__ slots, uint8_t* __restrict__ pool,
                                  size_t slotStride, BrickLayout L)
{
    const Geo g = geos[blockIdx.y];
    T* out = reinterpret_cast<T*>(pool + size_t(slots[blockIdx.y]) * slotStride);
    const uint32_t n = L.dim[0] * L.dim[1] * L.dim[2];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const uint32_t x = i % L.dim[0];
        const uint32_t y = (i / L.dim[0]) % L.dim[1];
        const uint32_t z = i / (L.dim[0] * L.dim[1]);
        const int64_t c[3] = {g.origin[0] + x, g.origin[1] + y, g.origin[2] + z};
        out[tiledIndex(L, x, y, z)] = fieldVoxel<T>(s, blobs, waves, g.level, c);
    }
}
} // namespace
} // namespace lvr

extern "C" {

int lvr_field_brick_host(const lvr_volume_info* info, const lvr_field_params* p,
                         uint64_t node_id, void* out, size_t bytes)
{
    if (!info || !p || !out || p->count < 1 || p->count > 4096 ||
        (p->kind != LVR_FIELD_BLOBS && p->kind != LVR_FIELD_SINES))
        return LVR_EINVAL;
    lvr_volume_info vi = *info;
    if (lvr_fill_regular_volume_info(&vi) != LVR_OK)
        return LVR_EINVAL;
    const int dt = vi.data_type;
    if (dt != LVR_DT_UINT8 && dt != LVR_DT_UINT16 && dt != LVR_DT_FLOAT)
        return LVR_EINVAL;
    const size_t n = size_t(vi.max_block[0]) * vi.max_block[1] * vi.max_block[2];
    if (bytes != n * lvr::bytesPerVoxel(dt) || lvr::decodeNode(node_id).level >= vi.depth)
        return LVR_EINVAL;
    const lvr::FieldSetup f = lvr::makeField(vi, *p);
    const lvr::Geo g = lvr::brickOrigin(vi, node_id);
    size_t i = 0;
    for (uint32_t z = 0; z < vi.max_block[2]; ++z)
        for (uint32_t y = 0; y < vi.max_block[1]; ++y)
            for (uint32_t x = 0; x < vi.max_block[0]; ++x, ++i)
            {
                const int64_t c[3] = {g.origin[0] + x, g.origin[1] + y, g.origin[2] + z};
                if (dt == LVR_DT_UINT8)
                    static_cast<uint8_t*>(out)[i] = lvr::fieldVoxel<uint8_t>(
                        f.shape, f.blobs.data(), f.waves.data(), g.level, c);
                else if (dt == LVR_DT_UINT16)
                    static_cast<uint16_t*>(out)[i] = lvr::fieldVoxel<uint16_t>(
                        f.shape, f.blobs.data(), f.waves.data(), g.level, c);
                else
                    static_cast<float*>(out)[i] = lvr::fieldVoxel<float>(
                        f.shape, f.blobs.data(), f.waves.data(), g.level, c);
            }
    return LVR_OK;
}

int lvr_field_bricks_device(lvr_ctx* ctx, const lvr_field_params* p, const uint64_t* node_ids,
                            int n)
{
    if (!ctx || !p || n < 0 || (n > 0 && !node_ids) || p->count < 1 || p->count > 4096 ||
        (p->kind != LVR_FIELD_BLOBS && p->kind != LVR_FIELD_SINES))
        return LVR_EINVAL;
    const int dt = ctx->vi.data_type;
    if (dt != LVR_DT_UINT8 && dt != LVR_DT_UINT16 && dt != LVR_DT_FLOAT)
        return lvr::fail(ctx, LVR_EINVAL, "field sources support u8, u16, f32");
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    const lvr::FieldSetup f = lvr::makeField(ctx->vi, *p);
    const int chunk = 4096;
    auto up256 = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t blobBytes = up256(sizeof(lvr::Blob) * (f.blobs.size() + 1));
    const size_t waveBytes = up256(sizeof(lvr::Wave) * (f.waves.size() + 1));
    const size_t geoBytes = up256(sizeof(lvr::Geo) * chunk);
    uint8_t* scratch = static_cast<uint8_t*>(
        lvr::genScratch(ctx, blobBytes + waveBytes + geoBytes + sizeof(int32_t) * chunk));
    if (!scratch)
        return LVR_ENOMEM;
    lvr::Blob* d_blobs = reinterpret_cast<lvr::Blob*>(scratch);
    lvr::Wave* d_waves = reinterpret_cast<lvr::Wave*>(scratch + blobBytes);
    lvr::Geo* d_geo = reinterpret_cast<lvr::Geo*>(scratch + blobBytes + waveBytes);
    int32_t* d_slots = reinterpret_cast<int32_t*>(scratch + blobBytes + waveBytes + geoBytes);
    // slots taken here may still be read by batches on other streams
    rc = lvr::waitPoolReaders(ctx, ctx->stream);
    if (rc)
        return rc;
    if ((!f.blobs.empty() &&
         hipMemcpyAsync(d_blobs, f.blobs.data(), sizeof(lvr::Blob) * f.blobs.size(),
                        hipMemcpyHostToDevice, ctx->stream) != hipSuccess) ||
        (!f.waves.empty() &&
         hipMemcpyAsync(d_waves, f.waves.data(), sizeof(lvr::Wave) * f.waves.size(),
                        hipMemcpyHostToDevice, ctx->stream) != hipSuccess))
        return lvr::fail(ctx, LVR_EHIP, "field source upload");
    std::vector<lvr::Geo> geo;
    std::vector<int32_t> slots;
    for (int base = 0; base < n && rc == LVR_OK; base += chunk)
    {
        const int cnt = std::min(chunk, n - base);
        geo.resize(size_t(cnt));
        slots.resize(size_t(cnt));
        int taken = 0;
        for (int i = 0; i < cnt && rc == LVR_OK; ++i)
        {
            if (lvr::decodeNode(node_ids[base + i]).level >= ctx->vi.depth)
                rc = lvr::fail(ctx, LVR_EINVAL, "node id outside the octree");
            else
            {
                geo[size_t(i)] = lvr::brickOrigin(ctx->vi, node_ids[base + i]);
                rc = lvr::acquireSlotFor(ctx, node_ids[base + i], &slots[size_t(i)]);
                taken += rc == LVR_OK;
            }
        }
        if (rc == LVR_OK &&
            (hipMemcpyAsync(d_geo, geo.data(), sizeof(lvr::Geo) * cnt, hipMemcpyHostToDevice,
                            ctx->stream) != hipSuccess ||
             hipMemcpyAsync(d_slots, slots.data(), sizeof(int32_t) * cnt, hipMemcpyHostToDevice,
                            ctx->stream) != hipSuccess))
            rc = lvr::fail(ctx, LVR_EHIP, "field source upload");
        if (rc)
        {
            // bricks of this chunk never received their voxels
            const std::string msg = ctx->error;
            for (int i = 0; i < taken; ++i)
                (void)lvr::releaseSlot(ctx, node_ids[base + i]);
            ctx->error = msg;
            break;
        }
        const dim3 grid(64, unsigned(cnt));
        if (dt == LVR_DT_UINT8)
            hipLaunchKernelGGL(lvr::fieldBricksKernel<uint8_t>, grid, dim3(256), 0, ctx->stream,
                               f.shape, d_blobs, d_waves, d_geo, d_slots, ctx->d_pool,
                               ctx->slotStride, ctx->layout);
        else if (dt == LVR_DT_UINT16)
            hipLaunchKernelGGL(lvr::fieldBricksKernel<uint16_t>, grid, dim3(256), 0, ctx->stream,
                               f.shape, d_blobs, d_waves, d_geo, d_slots, ctx->d_pool,
                               ctx->slotStride, ctx->layout);
        else
            hipLaunchKernelGGL(lvr::fieldBricksKernel<float>, grid, dim3(256), 0, ctx->stream,
                               f.shape, d_blobs, d_waves, d_geo, d_slots, ctx->d_pool,
                               ctx->slotStride, ctx->layout);
        if (hipGetLastError() != hipSuccess)
            rc = lvr::fail(ctx, LVR_EHIP, "field source kernel");
        // each brick's histogram into the cache as it lands in its slot
        if (rc == LVR_OK)
            rc = lvr::cacheBrickHistograms(ctx, node_ids + base, slots.data(), cnt, ctx->stream);
        // the host vectors and the scratch are reused by the next chunk
        if (rc == LVR_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = lvr::fail(ctx, LVR_EHIP, "field source kernel");
    }
    return rc;
}
}
