// Device code of the ray caster shared by the single-frame kernel and the
// batched (many-request) kernel: per-frame uniforms, the per-tile march that
// restates fragRayCast.glsl:113-207 for every brick of a tile's ordered list,
// and the tile/work-item numbering.
#pragma once

#include "lvr_internal.h"

namespace lvr
{
constexpr int kWave = 64;
// work items are wavefront tiles in kSuper x kSuper super-tile order
constexpr int kSuper = 4;

__device__ inline void intersectBoxInv(const float o[3], const float inv[3],
                                       const float mn[3], const float mx[3], float& t0,
                                       float& t1)
{
    float tmin[3], tmax[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const float tb = inv[i] * (mn[i] - o[i]);
        const float tt = inv[i] * (mx[i] - o[i]);
        tmin[i] = fminf(tt, tb);
        tmax[i] = fmaxf(tt, tb);
    }
    t0 = fmaxf(fmaxf(tmin[0], tmin[1]), fmaxf(tmin[0], tmin[2]));
    t1 = fminf(fminf(tmax[0], tmax[1]), fminf(tmax[0], tmax[2]));
}

__device__ inline int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

template <typename T>
__device__ inline float loadVoxel(const T* v, uint32_t idx)
{
    return float(v[idx]);
}

// Texel index along one axis for GL_NEAREST: floor(u * N) clamped to
// [0, N-1]. Clamping the float to [0, N-1] first and truncating gives the same
// integer for every input (truncation equals floor on non-negative values, and
// everything below 0 clamps to 0 either way), in two instructions.
__device__ inline uint32_t nearestTexel(float tpN, float nMinus1)
{
    return uint32_t(int(__builtin_amdgcn_fmed3f(tpN, 0.0f, nMinus1)));
}

// tiledIndex() with 24-bit multiplies (coordinates and tile counts are far
// below 2^24), so the address takes no 64-bit multiply-add
__device__ inline uint32_t tiledIndexDev(const BrickLayout& L, uint32_t x, uint32_t y,
                                         uint32_t z)
{
    const uint32_t tile = __umul24(z >> 2, L.tiles[0] * L.tiles[1]) +
                          __umul24(y >> 2, L.tiles[0]) + (x >> 2);
    return (tile << 6) | ((z & 3u) << 4) | ((y & 3u) << 2) | (x & 3u);
}

// Byte offset of voxel (x, y, z) in a pool slot: tiledIndexDev() times the
// voxel size, in 9 integer operations. x + 4 y + 16 z places the voxel inside
// its micro-tile but advances only 4, 16 and 64 per tile step in x, y and z;
// the per-axis constants U.tileK add the rest of 64, 64 T0 and 64 T0 T1.
template <typename T>
__device__ inline uint32_t tiledOffsetDev(const FrameUniforms& U, uint32_t x, uint32_t y,
                                          uint32_t z)
{
    constexpr uint32_t lb = sizeof(T) == 1 ? 0u : (sizeof(T) == 2 ? 1u : 2u);
    uint32_t a = __umul24(x >> 2, U.tileK[0]);
    a = __umul24(y >> 2, U.tileK[1]) + a;
    a = __umul24(z >> 2, U.tileK[2]) + a;
    a = (x << lb) + a;
    a = (y << (lb + 2)) + a;
    a = (z << (lb + 4)) + a;
    return a;
}

// The byte offset of a voxel in a slot is a sum of one term per axis
// (tiledOffsetDev); with TAB the three terms come from per-axis tables in LDS
// (offsetTables) instead of being computed per sample.
template <typename T, bool TAB>
__device__ __forceinline__ uint32_t slotOffset(const FrameUniforms& U, const uint32_t* tab,
                                               uint32_t ix, uint32_t iy, uint32_t iz)
{
    if (TAB)
        return tab[ix] + tab[uint32_t(U.texDim[0]) + iy] +
               tab[uint32_t(U.texDim[0] + U.texDim[1]) + iz];
    return tiledOffsetDev<T>(U, ix, iy, iz);
}

// The per-axis terms of tiledOffsetDev for every texel index of the brick:
// texDim[0] + texDim[1] + texDim[2] words, entry i written by the thread
// with i = first (mod stride).
template <typename T>
__device__ inline void offsetTables(const FrameUniforms& U, uint32_t* tab, int first,
                                    int stride)
{
    constexpr uint32_t lb = sizeof(T) == 1 ? 0u : (sizeof(T) == 2 ? 1u : 2u);
    const int n0 = U.texDim[0], n1 = U.texDim[1], n2 = U.texDim[2];
    for (int i = first; i < n0 + n1 + n2; i += stride)
    {
        const int a = i < n0 ? 0 : (i < n0 + n1 ? 1 : 2);
        const uint32_t x = uint32_t(i - (a == 0 ? 0 : (a == 1 ? n0 : n0 + n1)));
        tab[i] = __umul24(x >> 2, U.tileK[a]) + (x << (lb + 2u * uint32_t(a)));
    }
}

// Voxel loads through a buffer resource over one pool slot: a 32-bit byte
// offset per lane instead of a 64-bit address, and the hardware returns 0
// for any offset past the slot.
__device__ inline __amdgpu_buffer_rsrc_t slotResource(const void* slot, uint32_t bytes)
{
    // word 3 of the buffer descriptor for raw (untyped) loads on gfx9 parts
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(slot), short(0), int(bytes),
                                             0x00020000);
}

template <typename T>
__device__ inline T loadSlot(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <>
__device__ inline uint8_t loadSlot<uint8_t>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
}
template <>
__device__ inline int8_t loadSlot<int8_t>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return int8_t(__builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0));
}
template <>
__device__ inline uint16_t loadSlot<uint16_t>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
}
template <>
__device__ inline int16_t loadSlot<int16_t>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return int16_t(__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0));
}
template <>
__device__ inline uint32_t loadSlot<uint32_t>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}
template <>
__device__ inline int32_t loadSlot<int32_t>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return int32_t(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <>
__device__ inline float loadSlot<float>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

// texture() on the brick: GL_NEAREST picks floor(u * N) clamped; GL_LINEAR
// blends the 8 texels around u * N - 0.5 with clamp-to-edge.
template <typename T, bool LINEAR>
__device__ inline float sampleBrick(const T* __restrict__ vox, const FrameUniforms& U,
                                    const float tp[3])
{
    const int nx = U.texDim[0], ny = U.texDim[1], nz = U.texDim[2];
    if (!LINEAR)
    {
        const uint32_t ix = nearestTexel(tp[0] * float(nx), float(nx - 1));
        const uint32_t iy = nearestTexel(tp[1] * float(ny), float(ny - 1));
        const uint32_t iz = nearestTexel(tp[2] * float(nz), float(nz - 1));
        return loadVoxel(vox, tiledIndexDev(U.layout, ix, iy, iz));
    }
    const float u[3] = {tp[0] * float(nx) - 0.5f, tp[1] * float(ny) - 0.5f,
                        tp[2] * float(nz) - 0.5f};
    const int n[3] = {nx, ny, nz};
    int i0[3], i1[3];
    float w[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        const float fl = floorf(u[a]);
        w[a] = u[a] - fl;
        const int k = int(fl);
        i0[a] = clampi(k, 0, n[a] - 1);
        i1[a] = clampi(k + 1, 0, n[a] - 1);
    }
    auto at = [&](int x, int y, int z) {
        return loadVoxel(vox, tiledIndexDev(U.layout, uint32_t(x), uint32_t(y), uint32_t(z)));
    };
    const float c00 = at(i0[0], i0[1], i0[2]) * (1.f - w[0]) + at(i1[0], i0[1], i0[2]) * w[0];
    const float c10 = at(i0[0], i1[1], i0[2]) * (1.f - w[0]) + at(i1[0], i1[1], i0[2]) * w[0];
    const float c01 = at(i0[0], i0[1], i1[2]) * (1.f - w[0]) + at(i1[0], i0[1], i1[2]) * w[0];
    const float c11 = at(i0[0], i1[1], i1[2]) * (1.f - w[0]) + at(i1[0], i1[1], i1[2]) * w[0];
    const float c0 = c00 * (1.f - w[1]) + c10 * w[1];
    const float c1 = c01 * (1.f - w[1]) + c11 * w[1];
    return c0 * (1.f - w[2]) + c1 * w[2];
}

// calcTexturePositionFromAABBPos (fragRayCast.glsl:72-76) for a sample
// position: (pos - aabbMin) / (aabbMax - aabbMin) correctly rounded, from the
// IEEE reciprocal of the divisor and one FMA correction step (Markstein),
// then * (texMax - texMin) + texMin.
// With P2 (the brick's extents are powers of two, BrickEntry::kBrickPow2)
// the quotient is exact and rcp holds (1 / ext) * texExt, so one multiply
// gives the same rounded product.
template <bool P2>
__device__ __forceinline__ void texturePositionOf(const FrameUniforms& U, const float bmin[3],
                                                  const float rcp[3], const float ext[3],
                                                  const float texExt[3], const float pos[3],
                                                  float tp[3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        const float x = pos[a] - bmin[a];
        if (P2)
            tp[a] = x * rcp[a] + U.texMin[a];
        else
        {
            const float q = x * rcp[a];
            const float r = __builtin_fmaf(-q, ext[a], x);
            const float quot = __builtin_fmaf(r, rcp[a], q);
            tp[a] = quot * texExt[a] + U.texMin[a];
        }
    }
}

template <bool P2>
__device__ __forceinline__ void texturePosition(const FrameUniforms& U, const BrickEntry& b,
                                                const float texExt[3], const float pos[3],
                                                float tp[3])
{
    texturePositionOf<P2>(U, b.bmin, b.rcp, b.ext, texExt, pos, tp);
}

// Byte offset in the slot of the GL_NEAREST texel of a sample position
// (with TAB from the per-axis offset tables `otab`).
template <typename T, bool P2, bool TAB = false>
__device__ __forceinline__ uint32_t nearestOffset(const FrameUniforms& U, const BrickEntry& b,
                                                  const float texExt[3], const float pos[3],
                                                  const uint32_t* otab = nullptr)
{
    float tp[3];
    texturePosition<P2>(U, b, texExt, pos, tp);
    const uint32_t ix = nearestTexel(tp[0] * U.texDimF[0], U.texLastF[0]);
    const uint32_t iy = nearestTexel(tp[1] * U.texDimF[1], U.texLastF[1]);
    const uint32_t iz = nearestTexel(tp[2] * U.texDimF[2], U.texLastF[2]);
    return slotOffset<T, TAB>(U, otab, ix, iy, iz);
}

// Classification of one sample value (fragRayCast.glsl:184-204 before the
// blend): texture(transferFnTex, density) with linear filtering and clamp to
// edge, then the opacity-corrected alpha. Returns (r, g, b, alpha): the
// filtered colour and the corrected alpha the blend weighs it with.
struct Classify
{
    float mult, add, nTfF, alphaCorrection, maxA;
    int32_t nTfLast;

    __device__ __forceinline__ float4 operator()(const float4* tf, float value) const
    {
        const float density = value * mult + add;
        // the two entries are floor(u) and floor(u) + 1 clamped to
        // [0, nTf - 1]. The table carries a copy of its first entry before
        // it and of its last after it (tf[0] and tf[nTf + 1]; entry j of the
        // function is tf[j + 1]), so floor(u) clamped to [-1, nTf - 1] picks
        // both with one index: below the range both reads give entry 0,
        // above it both give entry nTf - 1, exactly as clamping each index.
        const float u = density * nTfF - 0.5f;
        // floor and conversion in one instruction (saturating, so the
        // integer clamp keeps every input, NaN and infinities included,
        // inside the table); the weight u - floor(u) as v_fract_f32, which
        // equals it except for u in (-2^-25, 0), where it returns the largest
        // float below 1 instead of 1, and there both entries are the first
        // one. For infinities and NaN both give NaN.
        int fl;
        asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(fl) : "v"(u));
        int i;
        asm("v_med3_i32 %0, %1, -1, %2" : "=v"(i) : "v"(fl), "s"(nTfLast));
        const float w = __builtin_amdgcn_fractf(u);
        const float4 ca = tf[i + 1];
        const float4 cb = tf[i + 2];
        // the linear filter as a + w (b - a), one FMA per channel (a GLSL
        // compiler may contract the same way; the texel and TF indices above
        // stay exact, only colour roundings move by an ulp)
        const float sr = __builtin_fmaf(w, cb.x - ca.x, ca.x);
        const float sg = __builtin_fmaf(w, cb.y - ca.y, ca.y);
        const float sb = __builtin_fmaf(w, cb.z - ca.z, ca.z);
        const float sa = __builtin_fmaf(w, cb.w - ca.w, ca.w);
        // pow(x, c) for x in [1/256, 1] as exp2(c * log2 x) on the hardware
        // transcendental units (1 ulp each)
        const float base = 1.0f - fminf(sa, maxA);
        const float alpha =
            1.0f - __builtin_amdgcn_exp2f(alphaCorrection * __builtin_amdgcn_logf(base));
        return make_float4(sr, sg, sb, alpha);
    }
};

__device__ __forceinline__ Classify makeClassify(const FrameUniforms& U)
{
    Classify c;
    c.mult = U.mult;
    c.add = U.add;
    c.nTfF = U.nTfF;
    c.alphaCorrection = U.alphaCorrection;
    c.maxA = 1.0f - 1.0f / 256.0f;
    c.nTfLast = U.nTf - 1;
    return c;
}

// Front-to-back "over" of classified samples (fragRayCast.glsl:103-111):
// dst.rgb += src.rgb * alpha * (1 - dst.a); dst.a += alpha * (1 - dst.a).
// Returns true when the ray reached the early-exit opacity.
struct Composite
{
    Classify cls;
    float acc0, acc1, acc2, acc3;

    __device__ __forceinline__ bool blend(const FrameUniforms& U, float4 c)
    {
        // the colour weight alpha * (1 - dst.a) once, then one FMA per
        // channel (the shader's products in another association: colours
        // move by an ulp; dst.a, and so the early exit, is rounded as before)
        const float rem = 1.0f - acc3;
        const float wgt = c.w * rem;
        acc0 = __builtin_fmaf(c.x, wgt, acc0);
        acc1 = __builtin_fmaf(c.y, wgt, acc1);
        acc2 = __builtin_fmaf(c.z, wgt, acc2);
        acc3 = __builtin_fmaf(c.w, rem, acc3);
        return acc3 > U.exitA;
    }

    __device__ __forceinline__ bool add(const FrameUniforms& U, const float4* tf, float value)
    {
        return blend(U, cls(tf, value));
    }
};

__device__ inline int itemToTile(const FrameUniforms& U, int item)
{
    const int superX = (U.tilesX + kSuper - 1) / kSuper;
    int s = item / (kSuper * kSuper);
    const int w = item % (kSuper * kSuper);
    if (U.superOrder)
        s = U.superOrder[s];
    const int tx = (s % superX) * kSuper + w % kSuper;
    const int ty = (s / superX) * kSuper + w / kSuper;
    return (tx < U.tilesX && ty < U.tilesY) ? ty * U.tilesX + tx : -1;
}

// The march of one ray through one brick segment with GL_NEAREST fetches
// (fragRayCast.glsl:184-204): `travel` starts at the segment length and
// drops by stepSize per composite step (it ends: see kMaxSegment).
//
// D samples per thread are in the pipeline: while sample k is classified and
// composited, the voxels of samples k+1 .. k+D-1 are being fetched. The fetch
// position runs D - 1 steps ahead of the composited sample and advances by
// exactly the shader's `pos += step`, so sample k reads the voxel of the k-th
// position. Raw voxels sit in a ring of D - 1 registers: the oldest is
// converted, then its register takes the next load, so no register copy waits
// on a load. Fetches past the last composited sample (end of segment or early
// exit) are discarded; their texel indices are clamped into the brick like
// every other.
template <typename T, int D, bool CLS, bool P2, bool TAB = false>
__device__ __forceinline__ void marchNearest(__amdgpu_buffer_rsrc_t vox, const FrameUniforms& U,
                                             const BrickEntry& b, const float texExt[3],
                                             float pos[3], const float step[3], float travel,
                                             float stepSize, Composite& C, const float4* tf,
                                             uint32_t& nsamp, const uint32_t* otab = nullptr)
{
    constexpr int S = D > 1 ? D - 1 : 1;
    if (!(travel > 0.0f))
        return;
    T raw[S];
    raw[0] = loadSlot<T>(vox, nearestOffset<T, P2, TAB>(U, b, texExt, pos, otab));
#pragma unroll
    for (int i = 1; i < S; ++i)
    {
        pos[0] = pos[0] + step[0];
        pos[1] = pos[1] + step[1];
        pos[2] = pos[2] + step[2];
        raw[i] = loadSlot<T>(vox, nearestOffset<T, P2, TAB>(U, b, texExt, pos, otab));
    }
    for (bool more = true; more;)
    {
#pragma unroll
        for (int s = 0; s < S; ++s)
        {
            travel -= stepSize;
            float4 cls = make_float4(0.f, 0.f, 0.f, 0.f);
            float value = 0.f;
            if (CLS)
            {
                // classified table: one read per sample, issued before the
                // next fetch's address arithmetic so the two overlap
                cls = tf[uint32_t(raw[s])];
                __builtin_amdgcn_sched_barrier(0);
            }
            else
                value = float(raw[s]);
            if (D > 1)
            {
                pos[0] = pos[0] + step[0];
                pos[1] = pos[1] + step[1];
                pos[2] = pos[2] + step[2];
                const uint32_t next = nearestOffset<T, P2, TAB>(U, b, texExt, pos, otab);
                // the oldest voxel has been used before its register takes
                // the next load; the load stays above this sample's blend
                __builtin_amdgcn_sched_barrier(0);
                raw[s] = loadSlot<T>(vox, next);
                __builtin_amdgcn_sched_barrier(0);
            }
            ++nsamp;
            const bool opaque = CLS ? C.blend(U, cls) : C.add(U, tf, value);
            if (opaque || !(travel > 0.0f))
            {
                more = false;
                break;
            }
            if (D == 1)
            {
                pos[0] = pos[0] + step[0];
                pos[1] = pos[1] + step[1];
                pos[2] = pos[2] + step[2];
                raw[0] = loadSlot<T>(vox, nearestOffset<T, P2, TAB>(U, b, texExt, pos, otab));
            }
        }
    }
}

// The march with a classified table (integer voxels sampled NEAREST: the
// classification is a function of the raw value, tabulated per value). The
// table read is a dependent load, so it is pipelined too. Step k blends
// sample k with the table entry read during step k - 1, then reads the
// entry of sample k + 1 (whose voxel was fetched S - 1 steps earlier) into
// the same register, then fetches the voxel of sample k + S (S = D + 1).
// Loads complete in issue order, so the blend waits only for the entry and
// the entry read only for a voxel two steps old. Voxels sit in a ring of S
// registers, so no register copy waits on a load.
//
// The S steps of a pass have no branches, so the compiler keeps every load
// where it is issued: a ray that ends its segment (or turns opaque) inside a
// pass blends with weight 0 for the pass's remaining steps (its colour stays
// exactly as it was; table entries are finite) and leaves the loop at the end
// of the pass. Its extra fetches are clamped into the brick like every other.
// Per sample this leaves the position step, the texel address and the blend;
// the table holds exactly what the per-sample classification computes.
template <typename T, int D, bool P2, bool TAB = false>
__device__ __forceinline__ void marchTable(__amdgpu_buffer_rsrc_t vox, const FrameUniforms& U,
                                           const BrickEntry& b, const float texExt[3],
                                           float pos[3], const float step[3], float travel,
                                           float stepSize, Composite& C,
                                           const float4* __restrict__ tab, uint32_t& nsamp,
                                           const uint32_t* otab = nullptr)
{
    constexpr int S = D + 1;
    if (!(travel > 0.0f))
        return;
    const uint32_t first =
        loadSlot<T>(vox, nearestOffset<T, P2, TAB>(U, b, texExt, pos, otab));
    // the voxel of sample m (m >= 1) of the segment is raw[(m - 1) % S]; one
    // full register each (two 16-bit voxels packed in one register would
    // have to wait for both loads to merge them)
    uint32_t raw[S];
#pragma unroll
    for (int i = 0; i < S - 1; ++i)
    {
        pos[0] = pos[0] + step[0];
        pos[1] = pos[1] + step[1];
        pos[2] = pos[2] + step[2];
        raw[i] = loadSlot<T>(vox, nearestOffset<T, P2, TAB>(U, b, texExt, pos, otab));
    }
    float4 cls = tab[first];
    bool alive = true;
    for (;;)
    {
#pragma unroll
        for (int u = 0; u < S; ++u)
        {
            // blend (fragRayCast.glsl:103-111), weight 0 once the ray is done
            const float rem = alive ? 1.0f - C.acc3 : 0.0f;
            const float wgt = cls.w * rem;
            C.acc0 = __builtin_fmaf(cls.x, wgt, C.acc0);
            C.acc1 = __builtin_fmaf(cls.y, wgt, C.acc1);
            C.acc2 = __builtin_fmaf(cls.z, wgt, C.acc2);
            C.acc3 = __builtin_fmaf(cls.w, rem, C.acc3);
            nsamp += alive ? 1u : 0u;
            travel -= stepSize;
            // bitwise: a short-circuit would put branches in the pass
            alive = alive & !(C.acc3 > U.exitA) & (travel > 0.0f);
            cls = tab[raw[u]];
            __builtin_amdgcn_sched_barrier(0);
            pos[0] = pos[0] + step[0];
            pos[1] = pos[1] + step[1];
            pos[2] = pos[2] + step[2];
            const uint32_t next = nearestOffset<T, P2, TAB>(U, b, texExt, pos, otab);
            __builtin_amdgcn_sched_barrier(0);
            raw[(u + S - 1) % S] = loadSlot<T>(vox, next);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (!alive)
            break;
    }
}

// Longest segment a march takes, in world units. A brick box lies within 3
// world units per axis (the octree's root blocks cover less than three times
// the volume's longest axis, which is 1), so a real segment is shorter than
// 6 and is never cut; the bound keeps every march finite for any input:
// with at most kMaxSamplesPerRay samples per ray (the host rejects more) the
// step is at least 2^-20, more than half an ulp of every travel up to 16, so
// `travel -= stepSize` decreases strictly and a segment takes at most 2^24
// steps.
constexpr float kMaxSegment = 16.0f;
constexpr int32_t kMaxSamplesPerRay = 1 << 20;

// A pixel's ray (fragRayCast.glsl:121-140) and what every brick fragment of
// it recomputes identically: the direction, 1 / dir with zero components
// replaced by EPSILON (intersectBox), the near-plane distance along the ray,
// the view cosine of the exit depth and the entry into the volume's box.
struct PixelRay
{
    float dir[3], inv[3];
    float tNearPlane, viewCos, tnearGlobal;
};

__device__ __forceinline__ void pixelRay(const FrameUniforms& U, int px, int py, PixelRay& R)
{
    const float ndcx = (float(px) + 0.5f) / float(U.width) * 2.0f - 1.0f;
    const float ndcy = (float(py) + 0.5f) / float(U.height) * 2.0f - 1.0f;
    const float* P = U.invProj;
    float ep[3];
    ep[0] = P[0] * ndcx + P[4] * ndcy + P[8] * -1.0f + P[12] * 1.0f;
    ep[1] = P[1] * ndcx + P[5] * ndcy + P[9] * -1.0f + P[13] * 1.0f;
    ep[2] = P[2] * ndcx + P[6] * ndcy + P[10] * -1.0f + P[14] * 1.0f;
    const float ew = P[3] * ndcx + P[7] * ndcy + P[11] * -1.0f + P[15] * 1.0f;
    ep[0] = ep[0] / ew;
    ep[1] = ep[1] / ew;
    ep[2] = ep[2] / ew;
    const float* M = U.invMV;
    float wp[3];
    wp[0] = M[0] * ep[0] + M[4] * ep[1] + M[8] * ep[2] + M[12] * 1.0f;
    wp[1] = M[1] * ep[0] + M[5] * ep[1] + M[9] * ep[2] + M[13] * 1.0f;
    wp[2] = M[2] * ep[0] + M[6] * ep[1] + M[10] * ep[2] + M[14] * 1.0f;
    float* dir = R.dir;
    dir[0] = wp[0] - U.eye[0];
    dir[1] = wp[1] - U.eye[1];
    dir[2] = wp[2] - U.eye[2];
    {
        const float l = sqrtf(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        dir[0] = dir[0] / l;
        dir[1] = dir[1] / l;
        dir[2] = dir[2] / l;
    }
    float nz;
    {
        const float l = sqrtf(ep[0] * ep[0] + ep[1] * ep[1] + ep[2] * ep[2]);
        nz = ep[2] / l;
    }
    R.tNearPlane = -U.nearPlane / nz;
    R.viewCos = -nz;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        R.inv[i] = 1.0f / (dir[i] == 0.f ? 0.0000000001f : dir[i]);
    float tfarGlobal;
    intersectBoxInv(U.eye, R.inv, U.gmin, U.gmax, R.tnearGlobal, tfarGlobal);
}

// The segment one brick's back-face fragment marches for a ray
// (fragRayCast.glsl:141-183): the fragment exists where the exit point lies
// between the near and far clip planes; the entry is moved to the near plane
// and onto the ray's global sample grid, then cut by the clip planes. Returns
// false when the brick has no fragment for the ray or the march takes no
// step; else the start, the step vector and the travel (the segment length,
// at most kMaxSegment).
__device__ __forceinline__ bool brickSegment(const FrameUniforms& U, const PixelRay& R,
                                             const BrickEntry& b, float rs[3], float step[3],
                                             float& travel)
{
    const float* dir = R.dir;
    const float stepSize = U.stepSize;
    float tnear, tfar;
    intersectBoxInv(U.eye, R.inv, b.bmin, b.bmax, tnear, tfar);
    const float exitDepth = tfar * R.viewCos;
    if (!(exitDepth >= U.nearPlane && exitDepth <= U.farPlane))
        return false;
    if (tnear < R.tNearPlane)
        tnear = R.tNearPlane;
    {
        // GLSL mod(x, y) = x - y * floor(x / y)
        const float x = tnear - R.tnearGlobal;
        const float residu = x - stepSize * floorf(x / stepSize);
        if (residu > 0.0f)
            tnear += stepSize - residu;
    }
    if (tnear > tfar)
        return false;
    for (int c = 0; c < U.nClip; ++c)
    {
        const float* cp = U.clip + 4 * c;
        float rn = dir[0] * cp[0] + dir[1] * cp[1] + dir[2] * cp[2];
        if (rn == 0.f)
            rn = 0.0000000001f;
        const float t =
            -((cp[0] * U.eye[0] + cp[1] * U.eye[1] + cp[2] * U.eye[2]) + cp[3]) / rn;
        if (rn > 0.0f)
            tnear = fmaxf(tnear, t);
        else
            tfar = fminf(tfar, t);
    }
    if (tnear > tfar)
        return false;
    rs[0] = U.eye[0] + dir[0] * tnear;
    rs[1] = U.eye[1] + dir[1] * tnear;
    rs[2] = U.eye[2] + dir[2] * tnear;
    const float re[3] = {U.eye[0] + dir[0] * tfar, U.eye[1] + dir[1] * tfar,
                         U.eye[2] + dir[2] * tfar};
    const float seg[3] = {re[0] - rs[0], re[1] - rs[1], re[2] - rs[2]};
    const float segLen = sqrtf(seg[0] * seg[0] + seg[1] * seg[1] + seg[2] * seg[2]);
    step[0] = seg[0] / segLen * stepSize;
    step[1] = seg[1] / segLen * stepSize;
    step[2] = seg[2] / segLen * stepSize;
    travel = segLen > kMaxSegment ? kMaxSegment : segLen;
    return travel > 0.0f; // false for an empty segment (and a NaN one)
}

// One axis's term of a voxel's byte offset in its slot (tiledOffsetDev is
// the sum of the three): from the per-axis LDS table (offsetTables), or
// computed as offsetTables computes it.
template <typename T, bool TAB>
__device__ __forceinline__ uint32_t axisOffset(const FrameUniforms& U, const uint32_t* tab,
                                               int axis, uint32_t base, uint32_t x)
{
    constexpr uint32_t lb = sizeof(T) == 1 ? 0u : (sizeof(T) == 2 ? 1u : 2u);
    if (TAB)
        return tab[base + x];
    return __umul24(x >> 2, U.tileK[axis]) + (x << (lb + 2u * uint32_t(axis)));
}

// The eight texels of a GL_LINEAR sample and its three blend weights:
// sampleBrick<T, true>'s indices (u = tp N - 0.5, floor(u) and floor(u) + 1
// clamped to the brick) and weights (u - floor(u)), the texels read through
// the slot's buffer resource. v[x + 2 y + 4 z] is the texel at (i_x, j_y, k_z).
template <typename T>
struct LinearTexels
{
    T v[8];
    float w[3];
};

// The byte offsets of the eight texels of a GL_LINEAR sample at texture
// position tp (off[x + 2 y + 4 z]) and its blend weights.
template <typename T, bool TAB>
__device__ __forceinline__ void linearOffsets(const FrameUniforms& U, const uint32_t* tab,
                                              const float tp[3], uint32_t off[8], float w[3])
{
    uint32_t o0[3], o1[3];
    uint32_t base = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        const float u = tp[a] * U.texDimF[a] - 0.5f;
        const float fl = floorf(u);
        w[a] = u - fl;
        const int k = int(fl);
        const int last = U.texDim[a] - 1;
        o0[a] = axisOffset<T, TAB>(U, tab, a, base, uint32_t(clampi(k, 0, last)));
        o1[a] = axisOffset<T, TAB>(U, tab, a, base, uint32_t(clampi(k + 1, 0, last)));
        base += uint32_t(U.texDim[a]);
    }
    const uint32_t xy[4] = {o0[0] + o0[1], o1[0] + o0[1], o0[0] + o1[1], o1[0] + o1[1]};
#pragma unroll
    for (int c = 0; c < 4; ++c)
    {
        off[c] = xy[c] + o0[2];
        off[c + 4] = xy[c] + o1[2];
    }
}

template <typename T, bool P2, bool TAB>
__device__ __forceinline__ void linearTexels(__amdgpu_buffer_rsrc_t vox, const FrameUniforms& U,
                                             const BrickEntry& b, const float texExt[3],
                                             const float pos[3], const uint32_t* tab,
                                             LinearTexels<T>& t)
{
    float tp[3];
    texturePosition<P2>(U, b, texExt, pos, tp);
    uint32_t off[8];
    linearOffsets<T, TAB>(U, tab, tp, off, t.w);
#pragma unroll
    for (int c = 0; c < 8; ++c)
        t.v[c] = loadSlot<T>(vox, off[c]);
}

// sampleBrick<T, true>'s blend of the eight texels, expression by expression
template <typename T>
__device__ __forceinline__ float linearBlend(const LinearTexels<T>& t)
{
    const float w0 = t.w[0], w1 = t.w[1], w2 = t.w[2];
    const float c00 = float(t.v[0]) * (1.f - w0) + float(t.v[1]) * w0;
    const float c10 = float(t.v[2]) * (1.f - w0) + float(t.v[3]) * w0;
    const float c01 = float(t.v[4]) * (1.f - w0) + float(t.v[5]) * w0;
    const float c11 = float(t.v[6]) * (1.f - w0) + float(t.v[7]) * w0;
    const float c0 = c00 * (1.f - w1) + c10 * w1;
    const float c1 = c01 * (1.f - w1) + c11 * w1;
    return c0 * (1.f - w2) + c1 * w2;
}

// The GL_LINEAR march of one ray through one brick segment: the texels of
// sample k + 1 are fetched before sample k is blended, classified and
// composited, so eight gathers are always in flight per lane. The two texel
// sets alternate (the loop is unrolled twice), so no register copy waits on
// a gather. Sample positions advance by the shader's `pos += step` and
// `travel` as in castTile's loop; the texels fetched past the last sample
// are discarded (their indices are clamped into the brick like every other).
template <typename T, bool P2, bool TAB>
__device__ __forceinline__ bool linearStep(__amdgpu_buffer_rsrc_t vox, const FrameUniforms& U,
                                           const BrickEntry& b, const float texExt[3],
                                           float pos[3], const float step[3], float& travel,
                                           float stepSize, Composite& C, const float4* tf,
                                           uint32_t& nsamp, const uint32_t* tab,
                                           const LinearTexels<T>& cur, LinearTexels<T>& next)
{
    pos[0] = pos[0] + step[0];
    pos[1] = pos[1] + step[1];
    pos[2] = pos[2] + step[2];
    linearTexels<T, P2, TAB>(vox, U, b, texExt, pos, tab, next);
    // the next sample's gathers stay above this sample's blend
    __builtin_amdgcn_sched_barrier(0);
    ++nsamp;
    if (C.add(U, tf, linearBlend(cur)))
        return false;
    travel -= stepSize;
    return travel > 0.0f;
}

template <typename T, bool P2, bool TAB>
__device__ __forceinline__ void marchLinear(__amdgpu_buffer_rsrc_t vox, const FrameUniforms& U,
                                            const BrickEntry& b, const float texExt[3],
                                            float pos[3], const float step[3], float travel,
                                            float stepSize, Composite& C, const float4* tf,
                                            uint32_t& nsamp, const uint32_t* tab)
{
    if (!(travel > 0.0f))
        return;
    LinearTexels<T> t0, t1;
    linearTexels<T, P2, TAB>(vox, U, b, texExt, pos, tab, t0);
    while (linearStep<T, P2, TAB>(vox, U, b, texExt, pos, step, travel, stepSize, C, tf, nsamp,
                                  tab, t0, t1) &&
           linearStep<T, P2, TAB>(vox, U, b, texExt, pos, step, travel, stepSize, C, tf, nsamp,
                                  tab, t1, t0))
    {
    }
}

// fragTexCopy.glsl into the RGBA8 framebuffer, round(clamp(c) * 255), or the
// unconverted RGBA32F accumulation into a sort-last layer.
__device__ __forceinline__ void storePixel(const FrameUniforms& U, int px, int py,
                                           const Composite& C)
{
    if (U.layer)
    {
        U.layer[size_t(py) * U.width + px] = make_float4(C.acc0, C.acc1, C.acc2, C.acc3);
        return;
    }
    const float c[4] = {C.acc0, C.acc1, C.acc2, C.acc3};
    uint32_t packed = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
        const float v = fminf(1.0f, fmaxf(0.0f, c[i]));
        packed |= uint32_t(floorf(v * 255.0f + 0.5f)) << (8 * i);
    }
    U.frame[size_t(py) * U.width + px] = packed;
}

// One wavefront tile of rays: fragRayCast.glsl for each brick of the tile's
// ordered list, composited in registers, then the RGBA8 store. `tf` is the
// transfer function (LDS or global, resolved at the call site), or with CLS
// the classified table indexed by the raw voxel value. The wavefront marches
// one brick at a time; castTileWindow below lets lanes run one brick ahead.
template <typename T, bool LINEAR, int D, bool CLS = false, bool TAB = false>
__device__ __forceinline__ void castTile(const FrameUniforms& U, int tile, int lane,
                                         const BrickEntry* __restrict__ bricks,
                                         const int32_t* __restrict__ offsets,
                                         const int32_t* __restrict__ counts,
                                         const int32_t* __restrict__ lists,
                                         const float4* tf, uint32_t& nsamp,
                                         const uint32_t* otab = nullptr)
{
    const int px = (tile % U.tilesX) * U.tileW + lane % U.tileW;
    const int py = (tile / U.tilesX) * U.tileH + lane / U.tileW;
    const bool inside = px < U.width && py < U.height;
    PixelRay R;
    pixelRay(U, px, py, R);

    const float stepSize = U.stepSize;
    const float* texExt = U.texExt;
    Composite C;
    C.cls = makeClassify(U);
    C.acc0 = C.acc1 = C.acc2 = C.acc3 = 0.f;
    bool alive = inside;

    const int gtile = U.tileBase + tile;
    const int beg = __builtin_amdgcn_readfirstlane(offsets[gtile]);
    const int cnt = __builtin_amdgcn_readfirstlane(counts[gtile]);
    for (int k = 0; k < cnt; ++k)
    {
        if (__ballot(alive) == 0)
            break;
        const int bi = __builtin_amdgcn_readfirstlane(lists[beg + k]);
        const BrickEntry b = bricks[bi];
        if (!alive)
            continue;
        // the slot the record pass read from the device page table; -1 for
        // a brick that is not resident is skipped, as the reference skips a
        // missing texture (RayCastRenderer.cpp:517-522)
        const int slot = b.slot;
        if (slot < 0)
            continue;
        if (C.acc3 > U.exitA) // imageLoad + discard (fragRayCast.glsl:115-117)
        {
            alive = false;
            continue;
        }
        float pos[3], step[3], segLen;
        if (!brickSegment(U, R, b, pos, step, segLen))
            continue;
        const T* vox = reinterpret_cast<const T*>(U.pool + size_t(slot) * U.slotStride);

        // the brick's flags are wave-uniform (scalar loads): no divergence
        const bool p2 = (b.flags & kBrickPow2) != 0;
        if (!LINEAR && CLS)
        {
            if (p2)
                marchTable<T, D, true, TAB>(slotResource(vox, uint32_t(U.slotStride)), U, b,
                                            texExt, pos, step, segLen, stepSize, C, tf, nsamp,
                                            otab);
            else
                marchTable<T, D, false, TAB>(slotResource(vox, uint32_t(U.slotStride)), U, b,
                                             texExt, pos, step, segLen, stepSize, C, tf, nsamp,
                                             otab);
        }
        else if (!LINEAR)
        {
            if (p2)
                marchNearest<T, D, false, true, TAB>(slotResource(vox, uint32_t(U.slotStride)),
                                                     U, b, texExt, pos, step, segLen, stepSize,
                                                     C, tf, nsamp, otab);
            else
                marchNearest<T, D, false, false, TAB>(slotResource(vox, uint32_t(U.slotStride)),
                                                      U, b, texExt, pos, step, segLen, stepSize,
                                                      C, tf, nsamp, otab);
        }
        else
        {
            if (p2)
                marchLinear<T, true, TAB>(slotResource(vox, uint32_t(U.slotStride)), U, b,
                                          texExt, pos, step, segLen, stepSize, C, tf, nsamp,
                                          otab);
            else
                marchLinear<T, false, TAB>(slotResource(vox, uint32_t(U.slotStride)), U, b,
                                           texExt, pos, step, segLen, stepSize, C, tf, nsamp,
                                           otab);
        }
    }
    if (inside)
        storePixel(U, px, py, C);
}

// The brick a lane of castTileWindow marches, in registers: its box minimum
// and extent, its slot, and the texture-position reciprocal in both forms:
// rcpP, BrickEntry::rcp of a power-of-two brick ((1 / ext) * texExt, read by
// the P2 march), and rcpG, 1 / ext correctly rounded (read by the general
// march; for a power-of-two brick its quotient is then exact and equals the
// P2 product bit for bit). A window marches with the P2 form when both of
// its bricks are powers of two, so a lane always finds the form it needs.
struct LaneBrick
{
    float bmin[3], ext[3], rcpG[3], rcpP[3];
    const uint8_t* vox;
};

// 1 / ext for the bricks of a window (wave-uniform): BrickEntry::rcp, or
// for a power-of-two brick (normal, with a normal reciprocal: kBrickPow2)
// the reciprocal 2^-e of ext = 2^e, exact, by negating the exponent.
__device__ __forceinline__ void generalReciprocal(const BrickEntry& b, float rcpG[3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
        rcpG[a] = (b.flags & kBrickPow2)
                      ? __uint_as_float(0x7F000000u - __float_as_uint(b.ext[a]))
                      : b.rcp[a];
}

__device__ __forceinline__ void laneBrick(const BrickEntry& b, const float rcpG[3],
                                          const uint8_t* vox, LaneBrick& L)
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        L.bmin[a] = b.bmin[a];
        L.ext[a] = b.ext[a];
        L.rcpG[a] = rcpG[a];
        L.rcpP[a] = b.rcp[a];
    }
    L.vox = vox;
}

template <typename T, bool P2, bool TAB>
__device__ __forceinline__ T laneFetch(const FrameUniforms& U, const LaneBrick& L,
                                       const uint32_t* tab, const float pos[3])
{
    float tp[3];
    texturePositionOf<P2>(U, L.bmin, P2 ? L.rcpP : L.rcpG, L.ext, U.texExt, pos, tp);
    const uint32_t ix = nearestTexel(tp[0] * U.texDimF[0], U.texLastF[0]);
    const uint32_t iy = nearestTexel(tp[1] * U.texDimF[1], U.texLastF[1]);
    const uint32_t iz = nearestTexel(tp[2] * U.texDimF[2], U.texLastF[2]);
    // the texel indices are clamped into the brick, so the offset stays
    // inside the slot
    return *reinterpret_cast<const T*>(L.vox + slotOffset<T, TAB>(U, tab, ix, iy, iz));
}

// laneFetch for GL_LINEAR: the eight texels and weights of the sample at pos
// in the lane's brick (linearTexels through the lane's slot pointer)
template <typename T, bool P2, bool TAB>
__device__ __forceinline__ void laneLinear(const FrameUniforms& U, const LaneBrick& L,
                                           const uint32_t* tab, const float pos[3],
                                           LinearTexels<T>& t)
{
    float tp[3];
    texturePositionOf<P2>(U, L.bmin, P2 ? L.rcpP : L.rcpG, L.ext, U.texExt, pos, tp);
    uint32_t off[8];
    linearOffsets<T, TAB>(U, tab, tp, off, t.w);
#pragma unroll
    for (int c = 0; c < 8; ++c)
        t.v[c] = *reinterpret_cast<const T*>(L.vox + off[c]);
}

// A lane of castTileWindow. The wavefront holds two bricks of the tile's
// list: A, the oldest one a lane still marches, and B, the next one with a
// segment for some lane (list positions kA < kB; `seg` is the position of
// the brick the lane marches). A lane marches while travel > 0 and it is not
// opaque; it has a B segment while travelB > 0 (set up for every lane when B
// was chosen). A lane that ends its A segment goes on with its B segment at
// once; one that ends its B segment, or has none, waits. When no lane
// marches A any more, B becomes A and the next brick becomes B. Each lane so
// composites its bricks in list order, sample by sample exactly as castTile
// does; only the wavefront's schedule differs: the lanes of a tile no longer
// wait for the longest segment of every brick.
struct WindowLane
{
    float pos[3], step[3], travel;
    float posB[3], stepB[3], travelB;
    int32_t seg;
    LaneBrick L;
};

// The march of the lanes that are marching, until none marches A, with one
// sample's fetch in flight per lane: the voxel of the next sample (GL_NEAREST)
// or its eight texels (LIN, GL_LINEAR) are fetched before the current sample
// is classified and blended. For LIN the current sample's texels are blended
// first and the next sample's are loaded into the same registers.
template <typename T, bool P2, bool TAB, bool LIN>
__device__ __forceinline__ void marchWindow(const FrameUniforms& U, const float4* tf,
                                            const uint32_t* tab, Composite& C, WindowLane& W,
                                            T& raw, LinearTexels<T>& tex, int kA, int kB,
                                            const BrickEntry& bB, const float rcpGB[3],
                                            const uint8_t* voxB, uint32_t& nsamp)
{
    const float stepSize = U.stepSize;
    for (;;)
    {
        // uniform exit: the lanes still here march A or B
        if (__ballot(W.seg == kA) == 0)
            break;
        const float value = LIN ? linearBlend(tex) : float(raw);
        W.travel -= stepSize;
        W.pos[0] = W.pos[0] + W.step[0];
        W.pos[1] = W.pos[1] + W.step[1];
        W.pos[2] = W.pos[2] + W.step[2];
        bool cont = W.travel > 0.0f;
        // a lane marching B has travelB == 0 (its B segment is taken), so
        // travelB > 0 means an A lane with a B segment
        if (!cont && W.travelB > 0.0f)
        {
            // the A segment ended: the B segment starts with this fetch
#pragma unroll
            for (int a = 0; a < 3; ++a)
            {
                W.pos[a] = W.posB[a];
                W.step[a] = W.stepB[a];
            }
            W.travel = W.travelB;
            W.travelB = 0.0f;
            W.seg = kB;
            laneBrick(bB, rcpGB, voxB, W.L);
            cont = true;
        }
        if (LIN)
        {
            laneLinear<T, P2, TAB>(U, W.L, tab, W.pos, tex);
            __builtin_amdgcn_sched_barrier(0);
        }
        else
        {
            const T next = laneFetch<T, P2, TAB>(U, W.L, tab, W.pos);
            __builtin_amdgcn_sched_barrier(0);
            raw = next;
        }
        ++nsamp;
        const bool opaque = C.add(U, tf, value);
        if (opaque || !cont)
            break;
    }
}

// castTile with the two-brick window (WindowLane) for the GL_NEAREST march
// with per-sample classification, or (LIN) the GL_LINEAR march: identical
// samples and frames, fewer idle lanes per wavefront step.
template <typename T, bool TAB, bool LIN = false>
__device__ __forceinline__ void castTileWindow(const FrameUniforms& U, int tile, int lane,
                                               const BrickEntry* __restrict__ bricks,
                                               const int32_t* __restrict__ offsets,
                                               const int32_t* __restrict__ counts,
                                               const int32_t* __restrict__ lists,
                                               const float4* tf, const uint32_t* tab,
                                               uint32_t& nsamp)
{
    const int px = (tile % U.tilesX) * U.tileW + lane % U.tileW;
    const int py = (tile / U.tilesX) * U.tileH + lane / U.tileW;
    const bool inside = px < U.width && py < U.height;
    PixelRay R;
    pixelRay(U, px, py, R);
    Composite C;
    C.cls = makeClassify(U);
    C.acc0 = C.acc1 = C.acc2 = C.acc3 = 0.f;

    WindowLane W;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        W.pos[a] = W.step[a] = W.posB[a] = W.stepB[a] = 0.f;
        W.L.bmin[a] = W.L.ext[a] = W.L.rcpG[a] = W.L.rcpP[a] = 0.f;
    }
    W.travel = W.travelB = 0.f;
    W.seg = -1;
    W.L.vox = U.pool;
    T raw = T(0);
    LinearTexels<T> tex;
#pragma unroll
    for (int c = 0; c < 8; ++c)
        tex.v[c] = T(0);
    tex.w[0] = tex.w[1] = tex.w[2] = 0.f;

    const int gtile = U.tileBase + tile;
    const int beg = __builtin_amdgcn_readfirstlane(offsets[gtile]);
    const int cnt = __builtin_amdgcn_readfirstlane(counts[gtile]);
    BrickEntry bB; // wave-uniform
    bB.flags = kBrickPow2;
    float rcpGB[3] = {0.f, 0.f, 0.f};
    const uint8_t* voxB = U.pool;
    int kA = -1, kB = -1;
    bool p2B = true;
    int k = 0;
    for (;;)
    {
        // a ray ends when it is opaque (imageLoad + discard,
        // fragRayCast.glsl:115-117)
        const bool alive = inside && !(C.acc3 > U.exitA);
        if (__ballot(alive) == 0)
            break;
        // the lanes still marching are in B: it becomes A
        kA = kB;
        const bool p2A = p2B;
        // B: the next brick of the list with a segment for some lane; bricks
        // that are not resident are skipped (RayCastRenderer.cpp:517-522)
        bool found = false;
        while (k < cnt)
        {
            const int bi = __builtin_amdgcn_readfirstlane(lists[beg + k]);
            ++k;
            const BrickEntry cand = bricks[bi];
            if (cand.slot < 0)
                continue;
            W.travelB = 0.0f;
            if (alive)
                brickSegment(U, R, cand, W.posB, W.stepB, W.travelB);
            if (__ballot(W.travelB > 0.0f) != 0)
            {
                bB = cand;
                kB = k - 1;
                found = true;
                break;
            }
        }
        const bool marching = alive && W.travel > 0.0f;
        if (!found)
        {
            W.travelB = 0.0f;
            if (__ballot(marching) == 0)
                break;
        }
        else
        {
            voxB = U.pool + size_t(bB.slot) * U.slotStride;
            p2B = (bB.flags & kBrickPow2) != 0;
            generalReciprocal(bB, rcpGB);
        }
        // the march is built for power-of-two bricks or for any brick
        const bool p2 = p2A && p2B;
        // waiting lanes start their B segment now
        if (alive && !marching && W.travelB > 0.0f)
        {
#pragma unroll
            for (int a = 0; a < 3; ++a)
            {
                W.pos[a] = W.posB[a];
                W.step[a] = W.stepB[a];
            }
            W.travel = W.travelB;
            W.travelB = 0.0f;
            W.seg = kB;
            laneBrick(bB, rcpGB, voxB, W.L);
            if (LIN)
            {
                if (p2)
                    laneLinear<T, true, TAB>(U, W.L, tab, W.pos, tex);
                else
                    laneLinear<T, false, TAB>(U, W.L, tab, W.pos, tex);
            }
            else
                raw = p2 ? laneFetch<T, true, TAB>(U, W.L, tab, W.pos)
                         : laneFetch<T, false, TAB>(U, W.L, tab, W.pos);
        }
        if (alive && W.travel > 0.0f)
        {
            if (p2)
                marchWindow<T, true, TAB, LIN>(U, tf, tab, C, W, raw, tex, kA, kB, bB, rcpGB,
                                               voxB, nsamp);
            else
                marchWindow<T, false, TAB, LIN>(U, tf, tab, C, W, raw, tex, kA, kB, bB,
                                                rcpGB, voxB, nsamp);
        }
    }
    if (inside)
        storePixel(U, px, py, C);
}

// The GL_NEAREST march with per-sample classification at a fetch depth D:
// depth 2 through the two-brick window, 1 and 3 brick by brick.
template <typename T, int D, bool TAB = false>
__device__ __forceinline__ void castNearestTile(const FrameUniforms& U, int tile, int lane,
                                                const BrickEntry* __restrict__ bricks,
                                                const int32_t* __restrict__ offsets,
                                                const int32_t* __restrict__ counts,
                                                const int32_t* __restrict__ lists,
                                                const float4* tf, uint32_t& nsamp,
                                                const uint32_t* tab = nullptr)
{
    if (D == 2)
        castTileWindow<T, TAB>(U, tile, lane, bricks, offsets, counts, lists, tf, tab, nsamp);
    else
        castTile<T, false, D, false, TAB>(U, tile, lane, bricks, offsets, counts, lists, tf,
                                          nsamp, tab);
}

// The per-brick record of a frame from its key (the device half of frame
// preparation): the pool slot from the device page table (the reference looks
// each brick's texture up by NodeId when it draws it, RayCastRenderer.cpp:
// 510-522), the world box as internalNodeToLODNode computes it
// (DataSourcePlugin.cpp:40-71), its extent, the correctly rounded reciprocal
// of the extent, and for power-of-two extents that reciprocal times the
// texture extent (exact) with kBrickPow2 set.
__device__ inline BrickEntry makeBrickEntry(const FrameUniforms& U, const BrickKey& key)
{
    BrickEntry e;
    const NodeKey k = decodeNode(key.id);
    nodeWorldBox(U.worldSize, U.rootBlocks, k, e.bmin, e.bmax);
    bool pow2 = true;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        e.ext[a] = e.bmax[a] - e.bmin[a];
        e.rcp[a] = 1.0f / e.ext[a];
        // a normal power of two: biased exponent in [1, 254], mantissa 0;
        // the scaled reciprocal must stay normal too
        const uint32_t u = __float_as_uint(e.ext[a]);
        const uint32_t ex = (u >> 23) & 0xFFu;
        const uint32_t sc = (__float_as_uint(e.rcp[a] * U.texExt[a]) >> 23) & 0xFFu;
        pow2 = pow2 && (u & 0x807FFFFFu) == 0u && ex != 0u && ex != 0xFFu && sc != 0u &&
               sc != 0xFFu;
    }
    e.flags = 0u;
    if (pow2)
    {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            e.rcp[a] = e.rcp[a] * U.texExt[a];
        e.flags = kBrickPow2;
    }
    e.slot = key.page >= 0 ? U.pageTable[key.page] : -1;
    e.level = k.level;
    e.pad = 0u;
    return e;
}

__device__ inline unsigned long long waveSum(unsigned long long s)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_xor(s, off, kWave);
    return s;
}

// Request of a global tile / work item: requests are sorted by tileBase and
// itemBase; binary search over the (wave-uniform) request table.
__device__ inline int findByTile(const FrameUniforms* __restrict__ reqs, int nReq, int gtile)
{
    int lo = 0, hi = nReq - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (reqs[mid].tileBase <= gtile)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ inline int findByItem(const FrameUniforms* __restrict__ reqs, int nReq, int item)
{
    int lo = 0, hi = nReq - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (reqs[mid].itemBase <= item)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}
} // namespace lvr
