// Per-frame histogram of the rendered bricks (HistogramObject.cpp:140-224,
// HistogramFilter.cpp:78-149): interior voxels of each brick are counted and
// scaled by 8^(depth - level - 1), the number of finest voxels a voxel of
// that level stands for. Integer arithmetic throughout, so the result is
// exact and independent of the order of the atomics.
//
// The reference caches one histogram per brick and sums the cached
// histograms of the rendered set every frame. Here the device keeps the
// running sum of the last set instead: a new frame adds the bricks that
// entered the set and subtracts those that left it (two's-complement u64
// adds, exact modulo 2^64, and the true result is never negative), and
// recomputes from scratch when that is cheaper or when any brick of the
// pool changed since.
//
// Bricks are grouped by level on the host; a workgroup takes a run of bricks
// of one level and one part of the bin range (all 256 bins for 8-bit and
// float data, one of two 32768-bin halves for 16-bit data), counts into an
// LDS histogram of 32-bit counters (at most 128 KiB, within gfx950's
// 160 KiB), and flushes the non-zero bins to the u64 result with one atomic
// each, shifted by 3 * (depth - level - 1) and negated for leaving bricks.
//
// A thread takes whole 4x4x4 micro-tiles of the pool layout (64 voxels, one
// 128-byte line for 16-bit data) with 16-byte loads, and tests the interior
// with three 4-bit masks per tile. Consecutive threads take tiles far apart
// in the brick (a multiplicative permutation of the tile index): the values
// of neighbouring voxels are close, and 64 lanes adding to one LDS bin in
// the same instruction would be serialized.
//
// The job list is written into pinned host memory that the kernel reads in
// place (fine-grained, uncached on the device): an upload would be a copy
// kernel on the histogram stream, which waits for CUs behind the frame's
// persistent ray cast. Two such buffers alternate; one is rewritten only
// after the histogram that read it has finished.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "lvr_internal.h"

namespace lvr
{
namespace
{
constexpr int kMaxLdsBins = 32768;

struct HistJob
{
    int32_t firstBrick; // index into the slot array
    int32_t nBricks;
    int32_t shift;      // 3 * (depth - level - 1)
    int32_t binLo;      // first bin of this job's part
    int32_t negate;     // 1: subtract (brick left the set)
    int32_t pad[3];
};

template <typename T>
__device__ inline int binOf(T v, float r0, float perBin, int bins)
{
    return int(v);
}
template <>
__device__ inline int binOf<float>(float v, float r0, float perBin, int bins)
{
    const float b = (v - r0) / perBin;
    int i = b >= 0.f ? int(b) : 0;
    return i >= bins ? bins - 1 : i;
}

// voxel k (0..15) of a z slice of a micro-tile held as 16 * sizeof(T) bytes
template <typename T>
__device__ inline T sliceVoxel(const uint4* w, int k);
template <>
__device__ inline uint8_t sliceVoxel<uint8_t>(const uint4* w, int k)
{
    const uint32_t d = (&w[0].x)[k >> 2];
    return uint8_t(d >> (8 * (k & 3)));
}
template <>
__device__ inline uint16_t sliceVoxel<uint16_t>(const uint4* w, int k)
{
    const uint32_t d = (&w[0].x)[k >> 1];
    return uint16_t(d >> (16 * (k & 1)));
}
template <>
__device__ inline float sliceVoxel<float>(const uint4* w, int k)
{
    return __uint_as_float((&w[0].x)[k]);
}

// bits i (0..3) of the voxels 4 * tc + i that lie in [lo, hi)
__device__ inline uint32_t interiorBits(int tc, int lo, int hi)
{
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        m |= uint32_t(4 * tc + i >= lo && 4 * tc + i < hi) << i;
    return m;
}

template <typename T>
__global__ __launch_bounds__(1024) void histKernel(
    const HistJob* __restrict__ jobs, const int32_t* __restrict__ slots,
    const uint8_t* __restrict__ pool, size_t slotStride, int M, int ov, uint32_t tileStride,
    int binsPerJob, int bins, float r0, float perBin, BrickLayout L,
    unsigned long long* __restrict__ out)
{
    extern __shared__ uint32_t sHist[];
    const HistJob job = jobs[blockIdx.x];
    for (int i = threadIdx.x; i < binsPerJob; i += blockDim.x)
        sHist[i] = 0;
    __syncthreads();
    const uint32_t T0 = L.tiles[0], T1 = L.tiles[1];
    const uint32_t nT = T0 * T1 * L.tiles[2];
    constexpr int kWords = int(sizeof(T)); // 16-byte words per z slice
    for (int b = 0; b < job.nBricks; ++b)
    {
        const uint4* tiles = reinterpret_cast<const uint4*>(
            pool + size_t(slots[job.firstBrick + b]) * slotStride);
        for (uint32_t t = threadIdx.x; t < nT; t += blockDim.x)
        {
            const uint32_t tile = uint32_t((uint64_t(t) * tileStride) % nT);
            const uint32_t row = tile / T0;
            const uint32_t xm = interiorBits(int(tile - row * T0), ov, M - ov);
            const uint32_t ym = interiorBits(int(row % T1), ov, M - ov);
            const uint32_t zm = interiorBits(int(row / T1), ov, M - ov);
            if (!xm || !ym || !zm)
                continue;
            // the 16 voxels (y, x) of a slice that are interior in x and y
            uint32_t yx = 0;
#pragma unroll
            for (int yi = 0; yi < 4; ++yi)
                yx |= ((ym >> yi) & 1u) ? xm << (4 * yi) : 0u;
            uint4 w[4 * kWords];
            const uint4* src = tiles + size_t(tile) * (4 * kWords);
#pragma unroll
            for (int i = 0; i < 4 * kWords; ++i)
                w[i] = src[i];
#pragma unroll
            for (int zi = 0; zi < 4; ++zi)
            {
                if (!((zm >> zi) & 1u))
                    continue;
#pragma unroll
                for (int k = 0; k < 16; ++k)
                {
                    const T v = sliceVoxel<T>(w + zi * kWords, k);
                    const int bin = binOf<T>(v, r0, perBin, bins) - job.binLo;
                    if (((yx >> k) & 1u) && uint32_t(bin) < uint32_t(binsPerJob))
                        atomicAdd(&sHist[bin], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < binsPerJob; i += blockDim.x)
    {
        const uint32_t c = sHist[i];
        if (c)
        {
            const unsigned long long v = (unsigned long long)c << job.shift;
            atomicAdd(&out[job.binLo + i], job.negate ? 0ull - v : v);
        }
    }
}

__global__ void clearKernel(unsigned long long* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0;
}
} // namespace
} // namespace lvr

extern "C" {

int lvr_histogram(lvr_ctx* ctx, const uint64_t* ids, int n, int bins)
{
    using namespace lvr;
    if (!ctx || n < 0 || (n > 0 && !ids))
        return LVR_EINVAL;
    const int dt = ctx->vi.data_type;
    int expected = 0;
    if (dt == LVR_DT_UINT8 || dt == LVR_DT_FLOAT)
        expected = 256;
    else if (dt == LVR_DT_UINT16)
        expected = 65536;
    else
        return fail(ctx, LVR_EINVAL, "histogram supports u8, u16 and f32 volumes");
    if (bins != expected)
        return fail(ctx, LVR_EINVAL, "bin count does not match the data type");
    if (ctx->vi.overlap[0] != ctx->vi.overlap[1] ||
        ctx->vi.overlap[0] != ctx->vi.overlap[2] ||
        ctx->vi.max_block[0] != ctx->vi.max_block[1] ||
        ctx->vi.max_block[0] != ctx->vi.max_block[2])
        return fail(ctx, LVR_EINVAL, "histogram needs cubic bricks");
    if (dt == LVR_DT_FLOAT && !(ctx->tfSet && ctx->range1 > ctx->range0))
        return fail(ctx, LVR_EINVAL, "float histogram needs the TF range");
    const int depth = int(ctx->vi.depth);
    for (int i = 0; i < n; ++i)
        if (int(decodeNode(ids[i]).level) >= depth)
            return fail(ctx, LVR_EINVAL, "node id outside the octree");
    int rc = ensureDevice(ctx);
    if (rc)
        return rc;

    // resident bricks of the new set (missing bricks are skipped, as the
    // histogram cache fails to load them), as a sorted multiset
    std::vector<uint64_t> next;
    next.reserve(size_t(n));
    for (int i = 0; i < n; ++i)
        if (slotOf(ctx, ids[i]) >= 0)
            next.push_back(ids[i]);
    std::sort(next.begin(), next.end());

    const float r0 = dt == LVR_DT_FLOAT ? ctx->range0 : 0.f;
    const float r1 = dt == LVR_DT_FLOAT ? ctx->range1 : 0.f;
    bool delta = ctx->histValid && ctx->histLast == bins &&
                 ctx->histGeneration == ctx->poolGeneration &&
                 ctx->histRange0 == r0 && ctx->histRange1 == r1;
    std::vector<uint64_t> added, removed;
    if (delta)
    {
        std::set_difference(next.begin(), next.end(), ctx->histSet.begin(),
                            ctx->histSet.end(), std::back_inserter(added));
        std::set_difference(ctx->histSet.begin(), ctx->histSet.end(), next.begin(),
                            next.end(), std::back_inserter(removed));
        delta = added.size() + removed.size() < next.size();
    }
    if (!delta)
    {
        added = next;
        removed.clear();
    }

    // bricks per workgroup: about four workgroups per CU in all
    const int halves = bins > kMaxLdsBins ? bins / kMaxLdsBins : 1;
    const int binsPerJob = bins / halves;
    const size_t parts = (added.size() + removed.size()) * size_t(halves);
    const size_t target = size_t(4) * size_t(std::max(ctx->numCUs, 1));
    const int perJob = int(std::max<size_t>(1, (parts + target - 1) / target));
    std::vector<int32_t> slots;
    std::vector<HistJob> jobs;
    for (int sign = 0; sign < 2; ++sign)
    {
        const std::vector<uint64_t>& src = sign ? removed : added;
        std::vector<std::vector<int32_t>> byLevel(static_cast<size_t>(depth));
        for (uint64_t id : src)
            byLevel[decodeNode(id).level].push_back(slotOf(ctx, id));
        for (int l = 0; l < depth; ++l)
        {
            const auto& v = byLevel[size_t(l)];
            for (size_t b = 0; b < v.size(); b += size_t(perJob))
            {
                const int first = int(slots.size());
                const int cnt = int(std::min<size_t>(size_t(perJob), v.size() - b));
                slots.insert(slots.end(), v.begin() + b, v.begin() + b + cnt);
                for (int h = 0; h < halves; ++h)
                    jobs.push_back({first, cnt, 3 * (depth - l - 1), h * binsPerJob,
                                    sign, {0, 0, 0}});
            }
        }
    }

    if (bins > ctx->histBins)
    {
        LVR_HIP(ctx, hipStreamSynchronize(ctx->histStream));
        (void)hipFree(ctx->d_hist);
        ctx->d_hist = nullptr;
        ctx->histValid = false;
        LVR_HIP(ctx, hipMalloc((void**)&ctx->d_hist, sizeof(unsigned long long) * bins));
        ctx->histBins = bins;
    }
    ctx->histLast = bins;
    // timed in the statistics of the last rendered frame
    lvr_ctx::FrameRecord* rec =
        ctx->timing && ctx->frameCount > 0
            ? &ctx->frames[(ctx->frameCount - 1) % lvr_ctx::kFrameSlots]
            : nullptr;
    if (rec && !rec->timed)
        rec = nullptr;
    if (rec)
        LVR_HIP(ctx, hipEventRecord(rec->ev[3], ctx->histStream));
    if (!delta)
        hipLaunchKernelGGL(clearKernel, dim3((bins + 255) / 256), dim3(256), 0,
                           ctx->histStream, ctx->d_hist, bins);
    if (!jobs.empty())
    {
        // job records, then the slot array, in the next pinned buffer
        const int r = ctx->histRingNext;
        ctx->histRingNext ^= 1;
        if (!ctx->histRingDone[r])
            LVR_HIP(ctx, hipEventCreateWithFlags(&ctx->histRingDone[r], hipEventDisableTiming));
        else
            LVR_HIP(ctx, hipEventSynchronize(ctx->histRingDone[r]));
        const size_t jobBytes = sizeof(HistJob) * jobs.size();
        const size_t need = jobBytes + sizeof(int32_t) * slots.size();
        if (need > ctx->histRingBytes[r])
        {
            if (ctx->h_histRing[r])
                ctx->retiredHost.push_back(ctx->h_histRing[r]);
            ctx->h_histRing[r] = nullptr;
            const size_t cap = std::max(need + need / 4 + 4096, 2 * ctx->histRingBytes[r]);
            ctx->histRingBytes[r] = 0;
            LVR_HIP(ctx, hipHostMalloc((void**)&ctx->h_histRing[r], cap,
                                       hipHostMallocMapped | hipHostMallocCoherent));
            ctx->histRingBytes[r] = cap;
        }
        std::memcpy(ctx->h_histRing[r], jobs.data(), jobBytes);
        std::memcpy(ctx->h_histRing[r] + jobBytes, slots.data(), sizeof(int32_t) * slots.size());
        uint8_t* dRing = nullptr;
        LVR_HIP(ctx, hipHostGetDevicePointer((void**)&dRing, ctx->h_histRing[r], 0));
        const HistJob* dj = reinterpret_cast<const HistJob*>(dRing);
        const int32_t* ds = reinterpret_cast<const int32_t*>(dRing + jobBytes);

        const int M = int(ctx->vi.max_block[0]);
        const int ov = int(ctx->vi.overlap[0]);
        const size_t lds = sizeof(uint32_t) * binsPerJob;
        const float perBin = (r1 - r0) / float(bins - 1);
        // a stride near 0.618 of the tile count and coprime to it
        const uint32_t nT = ctx->layout.tiles[0] * ctx->layout.tiles[1] * ctx->layout.tiles[2];
        uint32_t tileStride = std::max<uint32_t>(1, uint32_t(double(nT) * 0.6180339887));
        while (std::gcd(tileStride, nT) != 1)
            ++tileStride;
        const dim3 grid(unsigned(jobs.size())), block(1024);
        auto launch = [&](auto kernel) {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
            hipLaunchKernelGGL(kernel, grid, block, lds, ctx->histStream, dj, ds, ctx->d_pool,
                               ctx->slotStride, M, ov, tileStride, binsPerJob, bins, r0, perBin,
                               ctx->layout, ctx->d_hist);
        };
        if (dt == LVR_DT_UINT8)
            launch(&histKernel<uint8_t>);
        else if (dt == LVR_DT_UINT16)
            launch(&histKernel<uint16_t>);
        else
            launch(&histKernel<float>);
        LVR_HIP(ctx, hipGetLastError());
        LVR_HIP(ctx, hipEventRecord(ctx->histRingDone[r], ctx->histStream));
    }
    if (rec)
    {
        LVR_HIP(ctx, hipEventRecord(rec->ev[4], ctx->histStream));
        rec->histTimed = true;
    }
    LVR_HIP(ctx, hipGetLastError());
    LVR_HIP(ctx, hipEventRecord(ctx->histDone, ctx->histStream));
    ctx->histSet.swap(next);
    ctx->histValid = true;
    ctx->histGeneration = ctx->poolGeneration;
    ctx->histRange0 = r0;
    ctx->histRange1 = r1;
    ctx->histDeltaBricks = added.size() + removed.size();
    return LVR_OK;
}

int lvr_histogram_read(lvr_ctx* ctx, uint64_t* out, int bins)
{
    if (!ctx || !out || bins != ctx->histLast || !ctx->d_hist)
        return LVR_EINVAL;
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    LVR_HIP(ctx, hipMemcpyAsync(out, ctx->d_hist, sizeof(uint64_t) * bins,
                                hipMemcpyDeviceToHost, ctx->histStream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->histStream));
    return LVR_OK;
}
}
