// Ray casting over the ordered brick list on gfx950.
//
// The reference draws the back faces of every brick in front-to-back order,
// one draw call plus a full memory barrier per brick, and each fragment runs
// fragRayCast.glsl over its brick into an RGBA32F image
// (RayCastRenderer.cpp:467-571, fragRayCast.glsl:113-207). A pixel therefore
// composites the bricks its ray crosses in list order. Here the same per-pixel
// result is produced in one launch:
//   1. the host projects each brick to a conservative range of screen tiles;
//   2. binCount / scan / binFill build, per tile, the list of brick indices
//      in global order (one wavefront per tile, ballot compaction keeps the
//      order without sorting);
//   3. raycast: one wavefront per tile of 64 rays, the transfer function in
//      LDS, brick records read with scalar loads (wave-uniform), FP32
//      compositing in registers, early exit per ray and per wavefront
//      (ballot), and the RGBA8 store fused into the epilogue (no RGBA32F
//      round trip, fragTexCopy.glsl:11-15).
// The per-sample arithmetic follows the shader expression by expression with
// floating-point contraction off, so sample positions and texel choices are
// those of the FP32 restatement of the shader.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "lvr_internal.h"

namespace lvr
{
struct FrameUniforms
{
    float invProj[16];
    float invMV[16];
    float eye[3];
    float nearPlane, farPlane;
    int32_t width, height;
    float gmin[3], gmax[3];
    float texMin[3], texMax[3];
    int32_t texDim[3];
    int32_t nSamples, maxSamples;
    float exitA;
    float mult, add;
    int32_t nTf;
    int32_t tileW, tileH, tilesX, tilesY;
    int32_t nBricks;
    int32_t linear;
    int32_t nClip;
    float clip[24];
    uint64_t slotStride;
    BrickLayout layout; // tiled voxel order inside a pool slot
};

namespace
{
constexpr int kWave = 64;

__device__ inline void intersectBox(const float o[3], const float d0[3],
                                    const float mn[3], const float mx[3],
                                    float& t0, float& t1)
{
    float d[3] = {d0[0], d0[1], d0[2]};
    const float eps = 0.0000000001f;
    if (d[0] == 0.f)
        d[0] = eps;
    if (d[1] == 0.f)
        d[1] = eps;
    if (d[2] == 0.f)
        d[2] = eps;
    float tmin[3], tmax[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const float inv = 1.0f / d[i];
        const float tb = inv * (mn[i] - o[i]);
        const float tt = inv * (mx[i] - o[i]);
        tmin[i] = fminf(tt, tb);
        tmax[i] = fmaxf(tt, tb);
    }
    t0 = fmaxf(fmaxf(tmin[0], tmin[1]), fmaxf(tmin[0], tmin[2]));
    t1 = fminf(fminf(tmax[0], tmax[1]), fminf(tmax[0], tmax[2]));
}

// intersectBox with the reciprocal direction already formed
__device__ inline void intersectBoxInv(const float o[3], const float inv[3],
                                       const float mn[3], const float mx[3],
                                       float& t0, float& t1)
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

template <typename T>
__device__ inline float loadVoxel(const T* v, uint32_t idx)
{
    return float(v[idx]);
}

__device__ inline int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

template <typename T, bool LINEAR>
__device__ inline float sampleBrick(const T* __restrict__ vox,
                                    const FrameUniforms& U, const float tp[3])
{
    const int nx = U.texDim[0], ny = U.texDim[1], nz = U.texDim[2];
    if (!LINEAR)
    {
        const int ix = clampi(int(floorf(tp[0] * float(nx))), 0, nx - 1);
        const int iy = clampi(int(floorf(tp[1] * float(ny))), 0, ny - 1);
        const int iz = clampi(int(floorf(tp[2] * float(nz))), 0, nz - 1);
        return loadVoxel(vox, tiledIndex(U.layout, uint32_t(ix), uint32_t(iy), uint32_t(iz)));
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
        return loadVoxel(vox, tiledIndex(U.layout, uint32_t(x), uint32_t(y), uint32_t(z)));
    };
    const float c00 = at(i0[0], i0[1], i0[2]) * (1.f - w[0]) + at(i1[0], i0[1], i0[2]) * w[0];
    const float c10 = at(i0[0], i1[1], i0[2]) * (1.f - w[0]) + at(i1[0], i1[1], i0[2]) * w[0];
    const float c01 = at(i0[0], i0[1], i1[2]) * (1.f - w[0]) + at(i1[0], i0[1], i1[2]) * w[0];
    const float c11 = at(i0[0], i1[1], i1[2]) * (1.f - w[0]) + at(i1[0], i1[1], i1[2]) * w[0];
    const float c0 = c00 * (1.f - w[1]) + c10 * w[1];
    const float c1 = c01 * (1.f - w[1]) + c11 * w[1];
    return c0 * (1.f - w[2]) + c1 * w[2];
}

// One wavefront per tile: count the bricks whose tile range covers the tile.
__global__ __launch_bounds__(256) void binCountKernel(
    const int4* __restrict__ ranges, int nBricks, int tilesX, int nTiles,
    int32_t* __restrict__ counts)
{
    const int tile = __builtin_amdgcn_readfirstlane(
        int(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave));
    if (tile >= nTiles)
        return;
    const int lane = threadIdx.x & (kWave - 1);
    const int tx = tile % tilesX, ty = tile / tilesX;
    int count = 0;
    for (int base = 0; base < nBricks; base += kWave)
    {
        const int i = base + lane;
        bool hit = false;
        if (i < nBricks)
        {
            const int4 r = ranges[i];
            hit = tx >= r.x && tx <= r.z && ty >= r.y && ty <= r.w;
        }
        count += __popcll(__ballot(hit));
    }
    if (lane == 0)
        counts[tile] = count;
}

// Exclusive scan of the tile counts, one workgroup of 1024 threads.
// Also clears the frame's counters (composite steps, ray-cast work queue) for
// the kernels that follow on the stream.
__global__ __launch_bounds__(1024) void scanKernel(const int32_t* __restrict__ counts,
                                                   int n, int32_t* __restrict__ offsets,
                                                   unsigned long long* __restrict__ counters)
{
    __shared__ int32_t partial[1024];
    if (threadIdx.x < 2)
        counters[threadIdx.x] = 0;
    const int per = (n + 1023) / 1024;
    const int begin = threadIdx.x * per;
    const int end = min(n, begin + per);
    int s = 0;
    for (int i = begin; i < end; ++i)
        s += counts[i];
    partial[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1)
    {
        const int v = threadIdx.x >= off ? partial[threadIdx.x - off] : 0;
        __syncthreads();
        partial[threadIdx.x] += v;
        __syncthreads();
    }
    int run = threadIdx.x ? partial[threadIdx.x - 1] : 0;
    for (int i = begin; i < end; ++i)
    {
        offsets[i] = run;
        run += counts[i];
    }
}

// Same walk as binCountKernel, writing the brick indices in global order.
__global__ __launch_bounds__(256) void binFillKernel(
    const int4* __restrict__ ranges, int nBricks, int tilesX, int nTiles,
    const int32_t* __restrict__ offsets, int32_t* __restrict__ lists)
{
    const int tile = __builtin_amdgcn_readfirstlane(
        int(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave));
    if (tile >= nTiles)
        return;
    const int lane = threadIdx.x & (kWave - 1);
    const int tx = tile % tilesX, ty = tile / tilesX;
    int out = offsets[tile];
    const uint64_t below = (uint64_t(1) << lane) - 1;
    for (int base = 0; base < nBricks; base += kWave)
    {
        const int i = base + lane;
        bool hit = false;
        if (i < nBricks)
        {
            const int4 r = ranges[i];
            hit = tx >= r.x && tx <= r.z && ty >= r.y && ty <= r.w;
        }
        const uint64_t mask = __ballot(hit);
        if (hit)
            lists[out + __popcll(mask & below)] = i;
        out += __popcll(mask);
    }
}

// Work items are wavefront tiles in super-tile order: the 16 tiles of a
// kSuper x kSuper block of tiles are consecutive items, so they run at the
// same time on neighbouring wavefronts and share bricks in L1/L2.
constexpr int kSuper = 4;

__device__ inline int itemToTile(const FrameUniforms& U, int item)
{
    const int superX = (U.tilesX + kSuper - 1) / kSuper;
    const int s = item / (kSuper * kSuper), w = item % (kSuper * kSuper);
    const int tx = (s % superX) * kSuper + w % kSuper;
    const int ty = (s / superX) * kSuper + w / kSuper;
    return (tx < U.tilesX && ty < U.tilesY) ? ty * U.tilesX + tx : -1;
}

// One wavefront tile of rays: fragRayCast.glsl for each brick of the tile's
// ordered list, composited in registers, then the RGBA8 store.
template <typename T, bool LINEAR>
__device__ __forceinline__ void castTile(
    const FrameUniforms& U, int tile, int lane, const BrickEntry* __restrict__ bricks,
    const int32_t* __restrict__ offsets, const int32_t* __restrict__ counts,
    const int32_t* __restrict__ lists, const uint8_t* __restrict__ pool,
    const float4* sTf, uint32_t* __restrict__ frame, uint32_t& nsamp)
{
    const int px = (tile % U.tilesX) * U.tileW + lane % U.tileW;
    const int py = (tile / U.tilesX) * U.tileH + lane / U.tileW;
    const bool inside = px < U.width && py < U.height;

    // pixel ray (fragRayCast.glsl:121-140)
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
    float dir[3] = {wp[0] - U.eye[0], wp[1] - U.eye[1], wp[2] - U.eye[2]};
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
    const float tNearPlane = -U.nearPlane / nz;
    const float viewCos = -nz;
    // 1 / r.Dir with zero components replaced by EPSILON, as intersectBox
    // computes it for every brick of this ray (same value each time)
    float inv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        inv[i] = 1.0f / (dir[i] == 0.f ? 0.0000000001f : dir[i]);
    float tnearGlobal, tfarGlobal;
    intersectBoxInv(U.eye, inv, U.gmin, U.gmax, tnearGlobal, tfarGlobal);

    const float stepSize = 1.0f / float(U.nSamples);
    const float alphaCorrection = float(U.maxSamples) / float(U.nSamples);
    const float maxA = 1.0f - 1.0f / 256.0f;
    const float texExt[3] = {U.texMax[0] - U.texMin[0], U.texMax[1] - U.texMin[1],
                             U.texMax[2] - U.texMin[2]};
    const float nTfF = float(U.nTf);

    float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f, acc3 = 0.f;
    bool alive = inside;

    const int beg = __builtin_amdgcn_readfirstlane(offsets[tile]);
    const int cnt = __builtin_amdgcn_readfirstlane(counts[tile]);
    for (int k = 0; k < cnt; ++k)
    {
        if (__ballot(alive) == 0)
            break;
        const int bi = __builtin_amdgcn_readfirstlane(lists[beg + k]);
        const BrickEntry b = bricks[bi];
        if (!alive)
            continue;
        if (acc3 > U.exitA)
        {
            alive = false;
            continue;
        }
        float tnear, tfar;
        intersectBoxInv(U.eye, inv, b.bmin, b.bmax, tnear, tfar);
        const float exitDepth = tfar * viewCos;
        if (!(exitDepth >= U.nearPlane && exitDepth <= U.farPlane))
            continue;
        if (tnear < tNearPlane)
            tnear = tNearPlane;
        {
            // GLSL mod(x, y) = x - y * floor(x / y)
            const float x = tnear - tnearGlobal;
            const float residu = x - stepSize * floorf(x / stepSize);
            if (residu > 0.0f)
                tnear += stepSize - residu;
        }
        if (tnear > tfar)
            continue;
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
            continue;

        const float rs[3] = {U.eye[0] + dir[0] * tnear, U.eye[1] + dir[1] * tnear,
                             U.eye[2] + dir[2] * tnear};
        const float re[3] = {U.eye[0] + dir[0] * tfar, U.eye[1] + dir[1] * tfar,
                             U.eye[2] + dir[2] * tfar};
        float seg[3] = {re[0] - rs[0], re[1] - rs[1], re[2] - rs[2]};
        const float segLen = sqrtf(seg[0] * seg[0] + seg[1] * seg[1] + seg[2] * seg[2]);
        const float step[3] = {seg[0] / segLen * stepSize, seg[1] / segLen * stepSize,
                               seg[2] / segLen * stepSize};
        float pos[3] = {rs[0], rs[1], rs[2]};
        const T* vox = reinterpret_cast<const T*>(pool + size_t(b.slot) * U.slotStride);

        // A segment lies inside one brick, so it holds at most
        // diagonal * nSamples steps; the guard only bounds the loop should
        // `travel - stepSize` ever round back to `travel`.
        int guard = 1 << 22;
        for (float travel = segLen; travel > 0.0f && guard > 0; travel -= stepSize, --guard)
        {
            float tp[3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
            {
                // (pos - aabbMin) / (aabbMax - aabbMin), correctly rounded:
                // reciprocal of the divisor (host, IEEE) and one FMA
                // correction step (Markstein) give the IEEE quotient.
                const float x = pos[a] - b.bmin[a];
                const float q = x * b.rcp[a];
                const float r = __builtin_fmaf(-q, b.ext[a], x);
                const float quot = __builtin_fmaf(r, b.rcp[a], q);
                tp[a] = quot * texExt[a] + U.texMin[a];
            }
            const float value = sampleBrick<T, LINEAR>(vox, U, tp);
            const float density = value * U.mult + U.add;
            // texture(transferFnTex, density): linear, clamp to edge
            const float u = density * nTfF - 0.5f;
            const float fl = floorf(u);
            const float w = u - fl;
            const float flc = fminf(fmaxf(fl, -1.0f), nTfF);
            const int kk = int(flc);
            const int ia = clampi(kk, 0, U.nTf - 1);
            const int ib = clampi(kk + 1, 0, U.nTf - 1);
            const float4 ca = sTf[ia];
            const float4 cb = sTf[ib];
            const float sr = ca.x * (1.0f - w) + cb.x * w;
            const float sg = ca.y * (1.0f - w) + cb.y * w;
            const float sb = ca.z * (1.0f - w) + cb.z * w;
            const float sa = ca.w * (1.0f - w) + cb.w * w;
            // composite (fragRayCast.glsl:103-111)
            // pow(x, c) for x in [1/256, 1] as exp2(c * log2 x) on the
            // hardware transcendental units (1 ulp each)
            const float base = 1.0f - fminf(sa, maxA);
            const float alpha =
                1.0f - __builtin_amdgcn_exp2f(alphaCorrection * __builtin_amdgcn_logf(base));
            const float rem = 1.0f - acc3;
            acc0 += sr * alpha * rem;
            acc1 += sg * alpha * rem;
            acc2 += sb * alpha * rem;
            acc3 += alpha * rem;
            ++nsamp;
            if (acc3 > U.exitA)
                break;
            pos[0] = pos[0] + step[0];
            pos[1] = pos[1] + step[1];
            pos[2] = pos[2] + step[2];
        }
    }

    if (inside)
    {
        const float c[4] = {acc0, acc1, acc2, acc3};
        uint32_t packed = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
            const float v = fminf(1.0f, fmaxf(0.0f, c[i]));
            packed |= uint32_t(floorf(v * 255.0f + 0.5f)) << (8 * i);
        }
        frame[size_t(py) * U.width + px] = packed;
    }
}

// Persistent ray-cast kernel: each workgroup (16 wavefronts) stages the
// transfer function in LDS once; then every wavefront pulls tiles from a
// frame-wide atomic counter until none are left. Tiles differ by orders of
// magnitude in work (empty sky vs rays through the whole volume), so pulling
// them dynamically keeps every CU busy until the end of the frame. The queue
// counter is reset by scanKernel earlier on the same stream.
template <typename T, bool LINEAR>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void raycastKernel(
    const FrameUniforms U, const BrickEntry* __restrict__ bricks,
    const int32_t* __restrict__ offsets, const int32_t* __restrict__ counts,
    const int32_t* __restrict__ lists, const uint8_t* __restrict__ pool,
    const float4* __restrict__ tf, uint32_t* __restrict__ frame,
    unsigned long long* __restrict__ counters)
{
    extern __shared__ float4 sTf[];
    for (int i = threadIdx.x; i < U.nTf; i += blockDim.x)
        sTf[i] = tf[i];
    __syncthreads();

    const int lane = threadIdx.x & (kWave - 1);
    const int superX = (U.tilesX + kSuper - 1) / kSuper;
    const int superY = (U.tilesY + kSuper - 1) / kSuper;
    const int nItems = superX * superY * kSuper * kSuper;
    uint32_t nsamp = 0;
    for (;;)
    {
        unsigned long long got = 0;
        if (lane == 0)
            got = atomicAdd(&counters[1], 1ull);
        const int item = __builtin_amdgcn_readfirstlane(int(__shfl(got, 0, kWave)));
        if (item >= nItems)
            break;
        const int tile = itemToTile(U, item);
        if (tile >= 0)
            castTile<T, LINEAR>(U, tile, lane, bricks, offsets, counts, lists, pool, sTf,
                                frame, nsamp);
    }
    // wavefront sum of executed composite steps
    unsigned long long s = nsamp;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_xor(s, off, kWave);
    if (lane == 0 && s)
        atomicAdd(&counters[0], s);
}

template <typename T>
void launchRaycast(int blocks, int threads, size_t lds, hipStream_t st,
                   const FrameUniforms& U, const BrickEntry* bricks,
                   const int32_t* offsets, const int32_t* counts,
                   const int32_t* lists, const uint8_t* pool, const float4* tf,
                   uint32_t* frame, unsigned long long* counters)
{
    if (U.linear)
    {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&raycastKernel<T, true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        hipLaunchKernelGGL((raycastKernel<T, true>), dim3(blocks), dim3(threads), lds, st,
                           U, bricks, offsets, counts, lists, pool, tf, frame, counters);
    }
    else
    {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&raycastKernel<T, false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        hipLaunchKernelGGL((raycastKernel<T, false>), dim3(blocks), dim3(threads), lds, st,
                           U, bricks, offsets, counts, lists, pool, tf, frame, counters);
    }
}
} // namespace

// Data range rule of RayCastRenderer::Impl::onFrameStart
// (RayCastRenderer.cpp:236-285): the type's full range, replaced by the
// transfer function's range when that range is valid.
static void dataRange(int dt, float tf0, float tf1, float* r0, float* r1)
{
    switch (dt)
    {
    case LVR_DT_UINT8:
        *r0 = 0.f;
        *r1 = 255.f;
        break;
    case LVR_DT_UINT16:
        *r0 = 0.f;
        *r1 = 65535.f;
        break;
    case LVR_DT_UINT32:
        *r0 = 0.f;
        *r1 = 4294967295.f;
        break;
    case LVR_DT_INT8:
        *r0 = -128.f;
        *r1 = 127.f;
        break;
    case LVR_DT_INT16:
        *r0 = -32768.f;
        *r1 = 32767.f;
        break;
    case LVR_DT_INT32:
        *r0 = -2147483648.f;
        *r1 = 2147483647.f;
        break;
    default:
        *r0 = 1.17549435e-38f; // numeric_limits<float>::min()
        *r1 = 3.40282347e+38f;
        break;
    }
    if (tf1 > 0.f && tf1 - tf0 > 0.f)
    {
        *r0 = tf0;
        *r1 = tf1;
    }
}

template <typename T>
static int growBuffer(lvr_ctx* ctx, T** ptr, size_t* cap, size_t need)
{
    if (need <= *cap && *ptr)
        return LVR_OK;
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(*ptr);
    *ptr = nullptr;
    size_t n = std::max<size_t>(need, 1);
    n = n + n / 4;
    LVR_HIP(ctx, hipMalloc((void**)ptr, n * sizeof(T)));
    *cap = n;
    return LVR_OK;
}
} // namespace lvr

extern "C" {

int lvr_render(lvr_ctx* ctx, const lvr_frustum* f, const int viewport[4],
               const uint64_t* ordered_ids, int n,
               const lvr_render_params* params)
{
    using namespace lvr;
    if (!ctx || !f || !viewport || !params || n < 0 || (n > 0 && !ordered_ids))
        return LVR_EINVAL;
    if (!ctx->tfSet)
        return fail(ctx, LVR_EINVAL, "no transfer function set");
    const int W = viewport[2], H = viewport[3];
    if (W <= 0 || H <= 0 || W > 16384 || H > 16384)
        return fail(ctx, LVR_EINVAL, "invalid viewport");
    if (params->n_samples_per_ray <= 0 || params->n_clip_planes < 0 ||
        params->n_clip_planes > 6)
        return fail(ctx, LVR_EINVAL, "invalid render parameters");
    int tw = params->tile_w, th = params->tile_h;
    if (tw == 0 && th == 0)
    {
        tw = 8;
        th = 8;
    }
    if (tw <= 0 || th <= 0 || tw * th != 64)
        return fail(ctx, LVR_EINVAL, "ray tile must hold 64 rays");
    int rc = ensureDevice(ctx);
    if (rc)
        return rc;

    FrameUniforms U;
    std::memset(&U, 0, sizeof U);
    std::memcpy(U.invProj, f->inv_proj, sizeof U.invProj);
    std::memcpy(U.invMV, f->inv_mv, sizeof U.invMV);
    for (int i = 0; i < 3; ++i)
        U.eye[i] = f->eye[i];
    U.nearPlane = f->near_plane;
    U.farPlane = f->far_plane;
    U.width = W;
    U.height = H;
    const lvr_volume_info& vi = ctx->vi;
    for (int i = 0; i < 3; ++i)
    {
        const float half = vi.world_size[i] / 2.0f;
        U.gmin[i] = -half;
        U.gmax[i] = half;
        // TextureObject::Impl::initialize (TextureObject.cpp:76-92)
        const float ov = float(vi.overlap[i]) / float(vi.max_block[i]);
        const float sz = float(vi.max_block[i] - 2u * vi.overlap[i]) /
                         float(vi.max_block[i]);
        U.texMin[i] = ov;
        U.texMax[i] = ov + sz;
        U.texDim[i] = int32_t(vi.max_block[i]);
    }
    U.nSamples = params->n_samples_per_ray;
    U.maxSamples = params->max_samples_per_ray;
    U.exitA = params->early_exit;
    float r0, r1;
    dataRange(vi.data_type, ctx->range0, ctx->range1, &r0, &r1);
    U.mult = 1.0f / (r1 - r0);
    U.add = -r0 / (r1 - r0);
    U.nTf = ctx->nTf;
    U.tileW = tw;
    U.tileH = th;
    U.tilesX = (W + tw - 1) / tw;
    U.tilesY = (H + th - 1) / th;
    U.linear = params->linear_filter ? 1 : 0;
    U.nClip = params->n_clip_planes;
    std::memcpy(U.clip, params->clip_planes, sizeof U.clip);
    U.slotStride = ctx->slotStride;
    U.layout = ctx->layout;
    const int nTiles = U.tilesX * U.tilesY;

    // brick records and conservative tile ranges
    if (n > ctx->brickCapacity)
    {
        LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        (void)hipFree(ctx->d_bricks);
        (void)hipFree(ctx->d_tileRange);
        if (ctx->h_bricks)
            (void)hipHostFree(ctx->h_bricks);
        ctx->d_bricks = nullptr;
        ctx->d_tileRange = nullptr;
        ctx->h_bricks = nullptr;
        const int cap = std::max(n + n / 4, 256);
        LVR_HIP(ctx, hipMalloc((void**)&ctx->d_bricks, sizeof(BrickEntry) * cap));
        LVR_HIP(ctx, hipMalloc((void**)&ctx->d_tileRange, sizeof(int32_t) * 4 * cap));
        LVR_HIP(ctx, hipHostMalloc((void**)&ctx->h_bricks,
                                   (sizeof(BrickEntry) + 16) * cap,
                                   hipHostMallocDefault));
        ctx->brickCapacity = cap;
    }
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream)); // pinned records reused
    BrickEntry* hb = ctx->h_bricks;
    int32_t* hr = reinterpret_cast<int32_t*>(ctx->h_bricks + ctx->brickCapacity);
    double mvp[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
        {
            double s = 0;
            for (int k = 0; k < 4; ++k)
                s += double(f->proj[k * 4 + i]) * double(f->mv[j * 4 + k]);
            mvp[j * 4 + i] = s;
        }
    uint64_t listTotal = 0, onScreen = 0;
    for (int b = 0; b < n; ++b)
    {
        const NodeKey key = decodeNode(ordered_ids[b]);
        BrickEntry& e = hb[b];
        std::memset(&e, 0, sizeof e);
        int32_t* r = hr + 4 * b;
        r[0] = 1;
        r[1] = 1;
        r[2] = 0;
        r[3] = 0; // empty range
        if (key.level >= vi.depth)
            return fail(ctx, LVR_EINVAL, "node id outside the octree");
        nodeWorldBox(vi, key, e.bmin, e.bmax);
        for (int a = 0; a < 3; ++a)
        {
            e.ext[a] = e.bmax[a] - e.bmin[a];
            e.rcp[a] = 1.0f / e.ext[a];
        }
        e.level = key.level;
        e.slot = slotOf(ctx, ordered_ids[b]);
        if (e.slot < 0)
            continue; // not resident: skipped (RayCastRenderer.cpp:517-522)
        touch(ctx, ordered_ids[b]);
        double mnx = 1e300, mny = 1e300, mxx = -1e300, mxy = -1e300;
        bool full = false;
        for (int c = 0; c < 8 && !full; ++c)
        {
            const double v[3] = {(c & 1) ? e.bmax[0] : e.bmin[0],
                                 (c & 2) ? e.bmax[1] : e.bmin[1],
                                 (c & 4) ? e.bmax[2] : e.bmin[2]};
            double clip[4];
            for (int i = 0; i < 4; ++i)
                clip[i] = mvp[i] * v[0] + mvp[4 + i] * v[1] + mvp[8 + i] * v[2] +
                          mvp[12 + i];
            if (clip[3] <= 1e-6)
            {
                full = true;
                break;
            }
            const double sx = (clip[0] / clip[3] * 0.5 + 0.5) * W;
            const double sy = (clip[1] / clip[3] * 0.5 + 0.5) * H;
            mnx = std::min(mnx, sx);
            mny = std::min(mny, sy);
            mxx = std::max(mxx, sx);
            mxy = std::max(mxy, sy);
        }
        int x0, y0, x1, y1;
        if (full)
        {
            x0 = 0;
            y0 = 0;
            x1 = W - 1;
            y1 = H - 1;
        }
        else
        {
            if (mxx < -2.0 || mxy < -2.0 || mnx > W + 2.0 || mny > H + 2.0)
                continue; // entirely off screen
            x0 = int(std::max(0.0, std::floor(mnx) - 2));
            y0 = int(std::max(0.0, std::floor(mny) - 2));
            x1 = int(std::min(double(W - 1), std::ceil(mxx) + 2));
            y1 = int(std::min(double(H - 1), std::ceil(mxy) + 2));
        }
        r[0] = x0 / tw;
        r[1] = y0 / th;
        r[2] = x1 / tw;
        r[3] = y1 / th;
        listTotal += uint64_t(r[2] - r[0] + 1) * uint64_t(r[3] - r[1] + 1);
        ++onScreen;
    }
    if (listTotal > (1ull << 31))
        return fail(ctx, LVR_ENOMEM, "tile lists exceed 2^31 entries");

    // framebuffer and binning buffers
    const size_t pixels = size_t(W) * H;
    rc = growBuffer(ctx, &ctx->d_frame, &ctx->fbCapacity, pixels);
    if (rc)
        return rc;
    ctx->fbW = W;
    ctx->fbH = H;
    size_t tileCap = ctx->tileCapacity;
    if (size_t(nTiles) > tileCap || !ctx->d_tileCount)
    {
        LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        (void)hipFree(ctx->d_tileCount);
        (void)hipFree(ctx->d_tileOffset);
        ctx->d_tileCount = nullptr;
        ctx->d_tileOffset = nullptr;
        tileCap = size_t(nTiles) + size_t(nTiles) / 4;
        LVR_HIP(ctx, hipMalloc((void**)&ctx->d_tileCount, sizeof(int32_t) * tileCap));
        LVR_HIP(ctx, hipMalloc((void**)&ctx->d_tileOffset, sizeof(int32_t) * tileCap));
        ctx->tileCapacity = tileCap;
    }
    rc = growBuffer(ctx, &ctx->d_tileList, &ctx->tileListCapacity,
                    size_t(std::max<uint64_t>(listTotal, 1)));
    if (rc)
        return rc;

    if (n > 0)
    {
        LVR_HIP(ctx, hipMemcpyAsync(ctx->d_bricks, hb, sizeof(BrickEntry) * n,
                                    hipMemcpyHostToDevice, ctx->stream));
        LVR_HIP(ctx, hipMemcpyAsync(ctx->d_tileRange, hr, sizeof(int32_t) * 4 * n,
                                    hipMemcpyHostToDevice, ctx->stream));
    }
    const int binBlocks = (nTiles + 3) / 4;
    ctx->lastBricks = onScreen;
    ctx->lastEntries = listTotal;
    if (ctx->timing)
        LVR_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    hipLaunchKernelGGL(binCountKernel, dim3(binBlocks), dim3(256), 0, ctx->stream,
                       reinterpret_cast<const int4*>(ctx->d_tileRange), n,
                       U.tilesX, nTiles, ctx->d_tileCount);
    hipLaunchKernelGGL(scanKernel, dim3(1), dim3(1024), 0, ctx->stream,
                       ctx->d_tileCount, nTiles, ctx->d_tileOffset, ctx->d_counters);
    hipLaunchKernelGGL(binFillKernel, dim3(binBlocks), dim3(256), 0, ctx->stream,
                       reinterpret_cast<const int4*>(ctx->d_tileRange), n,
                       U.tilesX, nTiles, ctx->d_tileOffset, ctx->d_tileList);
    if (ctx->timing)
        LVR_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    // persistent: two 1024-ray workgroups per CU are resident (LDS and VGPR
    // budget), more would only find the queue empty
    const int threads = kSuper * kSuper * 64;
    const int superTiles =
        ((U.tilesX + kSuper - 1) / kSuper) * ((U.tilesY + kSuper - 1) / kSuper);
    const int blocks = std::max(1, std::min(superTiles, 2 * ctx->numCUs));
    const size_t lds = size_t(U.nTf) * sizeof(float4);
    const float4* tf = reinterpret_cast<const float4*>(ctx->d_tf);
    switch (vi.data_type)
    {
    case LVR_DT_UINT8:
        launchRaycast<uint8_t>(blocks, threads, lds, ctx->stream, U, ctx->d_bricks,
                               ctx->d_tileOffset, ctx->d_tileCount, ctx->d_tileList,
                               ctx->d_pool, tf, ctx->d_frame, ctx->d_counters);
        break;
    case LVR_DT_UINT16:
        launchRaycast<uint16_t>(blocks, threads, lds, ctx->stream, U, ctx->d_bricks,
                                ctx->d_tileOffset, ctx->d_tileCount, ctx->d_tileList,
                                ctx->d_pool, tf, ctx->d_frame, ctx->d_counters);
        break;
    case LVR_DT_UINT32:
        launchRaycast<uint32_t>(blocks, threads, lds, ctx->stream, U, ctx->d_bricks,
                                ctx->d_tileOffset, ctx->d_tileCount, ctx->d_tileList,
                                ctx->d_pool, tf, ctx->d_frame, ctx->d_counters);
        break;
    case LVR_DT_INT8:
        launchRaycast<int8_t>(blocks, threads, lds, ctx->stream, U, ctx->d_bricks,
                              ctx->d_tileOffset, ctx->d_tileCount, ctx->d_tileList,
                              ctx->d_pool, tf, ctx->d_frame, ctx->d_counters);
        break;
    case LVR_DT_INT16:
        launchRaycast<int16_t>(blocks, threads, lds, ctx->stream, U, ctx->d_bricks,
                               ctx->d_tileOffset, ctx->d_tileCount, ctx->d_tileList,
                               ctx->d_pool, tf, ctx->d_frame, ctx->d_counters);
        break;
    case LVR_DT_INT32:
        launchRaycast<int32_t>(blocks, threads, lds, ctx->stream, U, ctx->d_bricks,
                               ctx->d_tileOffset, ctx->d_tileCount, ctx->d_tileList,
                               ctx->d_pool, tf, ctx->d_frame, ctx->d_counters);
        break;
    default:
        launchRaycast<float>(blocks, threads, lds, ctx->stream, U, ctx->d_bricks,
                             ctx->d_tileOffset, ctx->d_tileCount, ctx->d_tileList,
                             ctx->d_pool, tf, ctx->d_frame, ctx->d_counters);
        break;
    }
    if (ctx->timing)
        LVR_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    LVR_HIP(ctx, hipGetLastError());
    return LVR_OK;
}

int lvr_frame_read(lvr_ctx* ctx, uint8_t* out_rgba8, size_t bytes,
                   uint64_t* samples)
{
    if (!ctx)
        return LVR_EINVAL;
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    const size_t need = size_t(ctx->fbW) * ctx->fbH * 4;
    if (out_rgba8)
    {
        if (bytes < need)
            return lvr::fail(ctx, LVR_ERANGE, "frame buffer too small");
        LVR_HIP(ctx, hipMemcpyAsync(out_rgba8, ctx->d_frame, need,
                                    hipMemcpyDeviceToHost, ctx->stream));
    }
    LVR_HIP(ctx, hipMemcpyAsync(ctx->h_counters, ctx->d_counters,
                                sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                ctx->stream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (samples)
        *samples = ctx->h_counters[0];
    return LVR_OK;
}

int lvr_set_timing(lvr_ctx* ctx, int enable)
{
    if (!ctx)
        return LVR_EINVAL;
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    if (enable && !ctx->ev[0])
        for (auto& e : ctx->ev)
            LVR_HIP(ctx, hipEventCreate(&e));
    ctx->timing = enable != 0;
    return LVR_OK;
}

int lvr_frame_stats_get(lvr_ctx* ctx, lvr_frame_stats* out)
{
    if (!ctx || !out)
        return LVR_EINVAL;
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    std::memset(out, 0, sizeof *out);
    LVR_HIP(ctx, hipMemcpyAsync(ctx->h_counters, ctx->d_counters,
                                sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                ctx->stream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    out->samples = ctx->h_counters[0];
    out->bricks = ctx->lastBricks;
    out->tile_entries = ctx->lastEntries;
    out->brick_bytes = ctx->brickBytes;
    if (ctx->timing)
    {
        LVR_HIP(ctx, hipEventElapsedTime(&out->bin_ms, ctx->ev[0], ctx->ev[1]));
        LVR_HIP(ctx, hipEventElapsedTime(&out->raycast_ms, ctx->ev[1], ctx->ev[2]));
        if (ctx->histTimed)
            LVR_HIP(ctx, hipEventElapsedTime(&out->histogram_ms, ctx->ev[3],
                                             ctx->ev[4]));
    }
    return LVR_OK;
}

int lvr_frame_device_ptr(lvr_ctx* ctx, void** ptr, int* width, int* height)
{
    if (!ctx || !ptr)
        return LVR_EINVAL;
    *ptr = ctx->d_frame;
    if (width)
        *width = ctx->fbW;
    if (height)
        *height = ctx->fbH;
    return LVR_OK;
}
}
