// Blob and plane-wave fields shared by the procedural sources: every value
// is computed with +, -, *, / and floor only (exp and sin by polynomials
// written here), compiled without contraction, so host and device agree bit
// for bit.
//  - blobs: sum of anisotropic, rotated Gaussians a * exp(-d^2 / 2) (d the
//    distance in units of sigma along the blob's principal axes, cut off at
//    d^2 = 100), plus uniform noise from a voxel hash, scaled so that the
//    maximum found on a probe grid maps to the type's top.
//  - sines: mean of `count` plane waves sin(k . p + phi) in [-1, 1], with a
//    polynomial sine after range reduction.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "lvr_internal.h"

namespace lvr
{
namespace field
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

inline FieldSetup makeField(const lvr_volume_info& vi, const lvr_field_params& p)
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
} // namespace field
} // namespace lvr
