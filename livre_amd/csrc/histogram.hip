// Per-frame histogram of the rendered bricks (HistogramObject.cpp:140-224,
// HistogramFilter.cpp:78-149): interior voxels of each brick are counted and
// scaled by 8^(depth - level - 1), the number of finest voxels a voxel of
// that level stands for. Integer arithmetic throughout, so the result is
// exact and independent of the order of the atomics.
//
// Bricks are grouped by level on the host; a workgroup takes a run of bricks
// of one level and one half of the bin range, counts into an LDS histogram
// of 32-bit counters (at most 32768 bins = 128 KiB, within gfx950's 160 KiB),
// and flushes the non-zero bins to the u64 result with one atomic each,
// shifted by 3 * (depth - level - 1).
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
    int binsPerJob, int bins, float r0, float perBin,
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
            const T v = vox[((z + ov) * M + (y + ov)) * M + (x + ov)];
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
            atomicAdd(&out[job.binLo + i], (unsigned long long)c << job.shift);
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
    int rc = ensureDevice(ctx);
    if (rc)
        return rc;
    const int depth = int(ctx->vi.depth);
    // group resident bricks by level (missing bricks are skipped, as the
    // histogram cache fails to load them)
    std::vector<std::vector<int32_t>> byLevel(static_cast<size_t>(depth));
    for (int i = 0; i < n; ++i)
    {
        const NodeKey k = decodeNode(ids[i]);
        if (int(k.level) >= depth)
            return fail(ctx, LVR_EINVAL, "node id outside the octree");
        const int32_t s = slotOf(ctx, ids[i]);
        if (s >= 0)
            byLevel[k.level].push_back(s);
    }
    std::vector<int32_t> slots;
    std::vector<HistJob> jobs;
    const int halves = bins > kMaxLdsBins ? bins / kMaxLdsBins : 1;
    const int binsPerJob = bins / halves;
    const int perJob = 4; // bricks per workgroup
    for (int l = 0; l < depth; ++l)
    {
        const auto& v = byLevel[size_t(l)];
        for (size_t b = 0; b < v.size(); b += perJob)
        {
            const int first = int(slots.size());
            const int cnt = int(std::min<size_t>(perJob, v.size() - b));
            slots.insert(slots.end(), v.begin() + b, v.begin() + b + cnt);
            for (int h = 0; h < halves; ++h)
                jobs.push_back({first, cnt, 3 * (depth - l - 1), h * binsPerJob});
        }
    }
    if (bins > ctx->histBins)
    {
        LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        (void)hipFree(ctx->d_hist);
        ctx->d_hist = nullptr;
        LVR_HIP(ctx, hipMalloc((void**)&ctx->d_hist, sizeof(unsigned long long) * bins));
        ctx->histBins = bins;
    }
    ctx->histLast = bins;
    if (ctx->timing)
        LVR_HIP(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
    hipLaunchKernelGGL(clearKernel, dim3((bins + 255) / 256), dim3(256), 0,
                       ctx->stream, ctx->d_hist, bins);
    if (!jobs.empty())
    {
        const size_t need = std::max(slots.size(), jobs.size() * sizeof(HistJob) / 4 + 1);
        if (int(need) > ctx->histCapacity)
        {
            LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
            (void)hipFree(ctx->d_histSlots);
            (void)hipFree(ctx->d_histScale);
            ctx->d_histSlots = nullptr;
            ctx->d_histScale = nullptr;
            const size_t cap = need + need / 4;
            LVR_HIP(ctx, hipMalloc((void**)&ctx->d_histSlots, sizeof(int32_t) * cap));
            LVR_HIP(ctx, hipMalloc((void**)&ctx->d_histScale, sizeof(int32_t) * cap));
            ctx->histCapacity = int(cap);
        }
        LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        LVR_HIP(ctx, hipMemcpy(ctx->d_histSlots, slots.data(),
                               sizeof(int32_t) * slots.size(), hipMemcpyHostToDevice));
        LVR_HIP(ctx, hipMemcpy(ctx->d_histScale, jobs.data(),
                               sizeof(HistJob) * jobs.size(), hipMemcpyHostToDevice));
        const int M = int(ctx->vi.max_block[0]);
        const int ov = int(ctx->vi.overlap[0]);
        const size_t lds = sizeof(uint32_t) * binsPerJob;
        float r0 = ctx->range0, r1 = ctx->range1;
        if (!ctx->tfSet)
        {
            r0 = 0.f;
            r1 = 0.f;
        }
        const float perBin = (r1 - r0) / float(bins - 1);
        const HistJob* dj = reinterpret_cast<const HistJob*>(ctx->d_histScale);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&histKernel<uint8_t>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&histKernel<uint16_t>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&histKernel<float>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (dt == LVR_DT_UINT8)
            hipLaunchKernelGGL(histKernel<uint8_t>, dim3(jobs.size()), dim3(1024), lds,
                               ctx->stream, dj, ctx->d_histSlots, ctx->d_pool,
                               ctx->slotStride, M, ov, binsPerJob, bins, r0, perBin,
                               ctx->d_hist);
        else if (dt == LVR_DT_UINT16)
            hipLaunchKernelGGL(histKernel<uint16_t>, dim3(jobs.size()), dim3(1024), lds,
                               ctx->stream, dj, ctx->d_histSlots, ctx->d_pool,
                               ctx->slotStride, M, ov, binsPerJob, bins, r0, perBin,
                               ctx->d_hist);
        else
        {
            if (!(r1 > r0))
                return fail(ctx, LVR_EINVAL, "float histogram needs the TF range");
            hipLaunchKernelGGL(histKernel<float>, dim3(jobs.size()), dim3(1024), lds,
                               ctx->stream, dj, ctx->d_histSlots, ctx->d_pool,
                               ctx->slotStride, M, ov, binsPerJob, bins, r0, perBin,
                               ctx->d_hist);
        }
    }
    if (ctx->timing)
    {
        LVR_HIP(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
        ctx->histTimed = true;
    }
    LVR_HIP(ctx, hipGetLastError());
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
                                hipMemcpyDeviceToHost, ctx->stream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LVR_OK;
}
}
