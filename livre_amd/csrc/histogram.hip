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
// of one level and one half of the bin range, counts into an LDS histogram
// of 32-bit counters (at most 32768 bins = 128 KiB, within gfx950's 160 KiB),
// and flushes the non-zero bins to the u64 result with one atomic each,
// shifted by 3 * (depth - level - 1) and negated for leaving bricks.
#include <algorithm>
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
    int32_t binLo;      // first bin of this job's half
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

template <typename T>
__global__ __launch_bounds__(1024) void histKernel(
    const HistJob* __restrict__ jobs, const int32_t* __restrict__ slots,
    const uint8_t* __restrict__ pool, size_t slotStride, int M, int ov,
    int binsPerJob, int bins, float r0, float perBin, BrickLayout L,
    unsigned long long* __restrict__ out)
{
    extern __shared__ uint32_t sHist[];
    const HistJob job = jobs[blockIdx.x];
    for (int i = threadIdx.x; i < binsPerJob; i += blockDim.x)
        sHist[i] = 0;
    __syncthreads();
    const int B = M - 2 * ov;
    const int n = B * B * B;
    for (int b = 0; b < job.nBricks; ++b)
    {
        const T* vox = reinterpret_cast<const T*>(
            pool + size_t(slots[job.firstBrick + b]) * slotStride);
        for (int i = threadIdx.x; i < n; i += blockDim.x)
        {
            const int x = i % B;
            const int y = (i / B) % B;
            const int z = i / (B * B);
            const T v = vox[tiledIndex(L, uint32_t(x + ov), uint32_t(y + ov), uint32_t(z + ov))];
            const int bin = binOf<T>(v, r0, perBin, bins) - job.binLo;
            if (bin >= 0 && bin < binsPerJob)
                atomicAdd(&sHist[bin], 1u);
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

    std::vector<int32_t> slots;
    std::vector<HistJob> jobs;
    const int halves = bins > kMaxLdsBins ? bins / kMaxLdsBins : 1;
    const int binsPerJob = bins / halves;
    const int perJob = 4; // bricks per workgroup
    for (int sign = 0; sign < 2; ++sign)
    {
        const std::vector<uint64_t>& src = sign ? removed : added;
        std::vector<std::vector<int32_t>> byLevel(static_cast<size_t>(depth));
        for (uint64_t id : src)
            byLevel[decodeNode(id).level].push_back(slotOf(ctx, id));
        for (int l = 0; l < depth; ++l)
        {
            const auto& v = byLevel[size_t(l)];
            for (size_t b = 0; b < v.size(); b += perJob)
            {
                const int first = int(slots.size());
                const int cnt = int(std::min<size_t>(perJob, v.size() - b));
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
    if (ctx->timing)
        LVR_HIP(ctx, hipEventRecord(ctx->ev[3], ctx->histStream));
    if (!delta)
        hipLaunchKernelGGL(clearKernel, dim3((bins + 255) / 256), dim3(256), 0,
                           ctx->histStream, ctx->d_hist, bins);
    if (!jobs.empty())
    {
        const size_t slotWords = slots.size();
        const size_t jobWords = jobs.size() * sizeof(HistJob) / sizeof(int32_t);
        if (slotWords > size_t(ctx->histCapacity) || jobWords > size_t(ctx->histJobCapacity))
        {
            LVR_HIP(ctx, hipStreamSynchronize(ctx->histStream));
            (void)hipFree(ctx->d_histSlots);
            (void)hipFree(ctx->d_histScale);
            ctx->d_histSlots = nullptr;
            ctx->d_histScale = nullptr;
            const size_t cs = slotWords + slotWords / 4 + 64;
            const size_t cj = jobWords + jobWords / 4 + 64;
            LVR_HIP(ctx, hipMalloc((void**)&ctx->d_histSlots, sizeof(int32_t) * cs));
            LVR_HIP(ctx, hipMalloc((void**)&ctx->d_histScale, sizeof(int32_t) * cj));
            ctx->histCapacity = int(cs);
            ctx->histJobCapacity = int(cj);
        }
        LVR_HIP(ctx, hipStreamSynchronize(ctx->histStream));
        LVR_HIP(ctx, hipMemcpyAsync(ctx->d_histSlots, slots.data(),
                                    sizeof(int32_t) * slots.size(), hipMemcpyHostToDevice,
                                    ctx->histStream));
        LVR_HIP(ctx, hipMemcpyAsync(ctx->d_histScale, jobs.data(),
                                    sizeof(HistJob) * jobs.size(), hipMemcpyHostToDevice,
                                    ctx->histStream));
        // the host vectors go out of scope at the end of this call
        LVR_HIP(ctx, hipStreamSynchronize(ctx->histStream));
        const int M = int(ctx->vi.max_block[0]);
        const int ov = int(ctx->vi.overlap[0]);
        const size_t lds = sizeof(uint32_t) * binsPerJob;
        const float perBin = (r1 - r0) / float(bins - 1);
        const HistJob* dj = reinterpret_cast<const HistJob*>(ctx->d_histScale);
        const dim3 grid(unsigned(jobs.size())), block(1024);
        if (dt == LVR_DT_UINT8)
        {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&histKernel<uint8_t>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
            hipLaunchKernelGGL(histKernel<uint8_t>, grid, block, lds, ctx->histStream, dj,
                               ctx->d_histSlots, ctx->d_pool, ctx->slotStride, M, ov,
                               binsPerJob, bins, r0, perBin, ctx->layout, ctx->d_hist);
        }
        else if (dt == LVR_DT_UINT16)
        {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&histKernel<uint16_t>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
            hipLaunchKernelGGL(histKernel<uint16_t>, grid, block, lds, ctx->histStream, dj,
                               ctx->d_histSlots, ctx->d_pool, ctx->slotStride, M, ov,
                               binsPerJob, bins, r0, perBin, ctx->layout, ctx->d_hist);
        }
        else
        {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&histKernel<float>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
            hipLaunchKernelGGL(histKernel<float>, grid, block, lds, ctx->histStream, dj,
                               ctx->d_histSlots, ctx->d_pool, ctx->slotStride, M, ov,
                               binsPerJob, bins, r0, perBin, ctx->layout, ctx->d_hist);
        }
    }
    if (ctx->timing)
    {
        LVR_HIP(ctx, hipEventRecord(ctx->ev[4], ctx->histStream));
        ctx->histTimed = true;
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
