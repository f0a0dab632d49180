// Per-frame histogram of the rendered bricks (HistogramObject.cpp:140-224,
// HistogramFilter.cpp:78-149): interior voxels of each brick are counted and
// scaled by 8^(depth - level - 1), the number of finest voxels a voxel of
// that level stands for. Integer arithmetic throughout, so the result is
// exact and independent of the order of the atomics.
//
// The reference caches one histogram per brick and sums the cached
// histograms of the rendered set every frame. Here each pool slot has a cache
// row of u32 counts, binned once for the brick it holds (histogramCached),
// and the device keeps the running sum of the last set: a new frame adds the
// rows of the bricks that entered the set and subtracts those that left it
// (two's-complement u64 adds, exact modulo 2^64, and the true result is never
// negative), and sums afresh when that is cheaper or a leaving brick's row is
// gone. When the rows would take more than half the pool's bytes, bricks are
// binned straight into the running sum instead (entering and leaving bricks
// every frame; from scratch when any brick of the pool changed).
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

// TO_CACHE = false: counts of the job's bricks, scaled and signed, added to
// the u64 result `out`. TO_CACHE = true (one brick per job): the brick's
// unscaled counts of the job's bin part stored into its row of the per-slot
// cache (row = slot, `bins` u32 per row), zeros included.
template <typename T, bool TO_CACHE>
__global__ __launch_bounds__(1024) void histKernel(
    const HistJob* __restrict__ jobs, const int32_t* __restrict__ slots,
    const uint8_t* __restrict__ pool, size_t slotStride, int M, int ov, uint32_t tileStride,
    int binsPerJob, int bins, float r0, float perBin, BrickLayout L,
    unsigned long long* __restrict__ out, uint32_t* __restrict__ cache)
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
    if (TO_CACHE)
    {
        uint32_t* row = cache + size_t(slots[job.firstBrick]) * size_t(bins) + job.binLo;
        for (int i = threadIdx.x; i < binsPerJob; i += blockDim.x)
            row[i] = sHist[i];
        return;
    }
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

// A term of the cached sum: the row of `slot`, scaled by 2^shift, added or
// (negate) subtracted.
struct SumTerm
{
    int32_t slot;
    int32_t shift;
    int32_t negate;
    int32_t pad;
};

// out[bin] += sum over a chunk of terms of +-(cache[slot][bin] << shift):
// blockIdx.x picks 4 * blockDim.x consecutive bins (one 16-byte load per
// term and thread, coalesced over the row), blockIdx.y a chunk of terms;
// the chunk's sum goes to `out` with one atomic per non-zero bin. Integer
// arithmetic modulo 2^64: exact in any order.
constexpr int kSumTermsPerBlock = 32;
__global__ __launch_bounds__(256) void histSumKernel(const SumTerm* __restrict__ terms, int nTerms,
                                                     const uint32_t* __restrict__ cache, int bins,
                                                     unsigned long long* __restrict__ out)
{
    const int b0 = (int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x)) * 4;
    if (b0 >= bins)
        return;
    const int t0 = int(blockIdx.y) * kSumTermsPerBlock;
    const int t1 = min(nTerms, t0 + kSumTermsPerBlock);
    unsigned long long acc[4] = {0ull, 0ull, 0ull, 0ull};
    for (int t = t0; t < t1; ++t)
    {
        const SumTerm term = terms[t];
        const uint4 c =
            *reinterpret_cast<const uint4*>(cache + size_t(term.slot) * size_t(bins) + b0);
        const uint32_t v[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            const unsigned long long x = (unsigned long long)v[k] << term.shift;
            acc[k] += term.negate ? 0ull - x : x;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (acc[k])
            atomicAdd(&out[b0 + k], acc[k]);
}

__global__ void clearKernel(unsigned long long* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0;
}
} // namespace

// The per-slot cache is used when its rows (bins x u32 per pool slot) take at
// most half the pool's bytes: 1 KiB per slot for 8-bit and float volumes,
// 256 KiB per slot for 16-bit ones.
static bool cacheFits(const lvr_ctx* ctx, int bins)
{
    const uint64_t rows = uint64_t(ctx->nSlots) * uint64_t(bins) * sizeof(uint32_t);
    return rows <= uint64_t(ctx->nSlots) * uint64_t(ctx->slotStride) / 2;
}

// Bins per job part and a tile stride near 0.618 of the tile count, coprime
// to it (consecutive threads take tiles far apart in the brick).
static void jobShape(const lvr_ctx* ctx, int bins, int* halves, uint32_t* tileStride)
{
    *halves = bins > kMaxLdsBins ? bins / kMaxLdsBins : 1;
    const uint32_t nT = ctx->layout.tiles[0] * ctx->layout.tiles[1] * ctx->layout.tiles[2];
    uint32_t st = std::max<uint32_t>(1, uint32_t(double(nT) * 0.6180339887));
    while (std::gcd(st, nT) != 1)
        ++st;
    *tileStride = st;
}

// The cache rows of every pool slot, allocated on first use.
static int ensureHistCache(lvr_ctx* ctx, int bins)
{
    if (ctx->d_histCache && ctx->histCacheBins == bins)
        return LVR_OK;
    LVR_HIP(ctx, hipStreamSynchronize(ctx->histStream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->uploadStream));
    (void)hipFree(ctx->d_histCache);
    ctx->d_histCache = nullptr;
    ctx->histCacheBins = 0;
    LVR_HIP(ctx, hipMalloc((void**)&ctx->d_histCache,
                           size_t(ctx->nSlots) * size_t(bins) * sizeof(uint32_t)));
    ctx->histCacheBins = bins;
    ctx->histCachedId.assign(size_t(ctx->nSlots), ~0ull);
    ctx->histRowVersion.assign(size_t(ctx->nSlots), 0u);
    ctx->histValid = false;
    return LVR_OK;
}

// The next pinned, device-mapped job buffer of at least `need` bytes, once
// the kernels that read it last have finished; *dev is its device address.
static int histRing(lvr_ctx* ctx, size_t need, uint8_t** host, uint8_t** dev, int* which)
{
    const int r = ctx->histRingNext;
    ctx->histRingNext ^= 1;
    if (!ctx->histRingDone[r])
        LVR_HIP(ctx, hipEventCreateWithFlags(&ctx->histRingDone[r], hipEventDisableTiming));
    else
        LVR_HIP(ctx, hipEventSynchronize(ctx->histRingDone[r]));
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
    *host = ctx->h_histRing[r];
    LVR_HIP(ctx, hipHostGetDevicePointer((void**)dev, ctx->h_histRing[r], 0));
    *which = r;
    return LVR_OK;
}

// HistogramObject per brick and HistogramFilter's sum (HistogramObject.cpp:
// 140-224, HistogramFilter.cpp:78-149): each brick of the set is binned once
// into its slot's cache row (unscaled counts of its interior voxels), the
// first time it is counted after entering that slot; the frame's histogram is
// the sum of the rows scaled by 8^(depth - level - 1). The running sum of the
// previous set is updated by the rows of the bricks that entered (+) and
// left (-) when every leaving brick's row is still its own and that is the
// cheaper sum; otherwise it is summed afresh.
static int histogramCached(lvr_ctx* ctx, std::vector<uint64_t>& next, int bins, float r0,
                           float r1)
{
    const int dt = ctx->vi.data_type;
    const int depth = int(ctx->vi.depth);
    int rc0 = ensureHistCache(ctx, bins);
    if (rc0)
        return rc0;
    if (dt == LVR_DT_FLOAT && (r0 != ctx->histCacheR0 || r1 != ctx->histCacheR1))
    {
        // float rows depend on the bin range (the TF range)
        std::fill(ctx->histCachedId.begin(), ctx->histCachedId.end(), ~0ull);
        ctx->histValid = false;
    }
    ctx->histCacheR0 = r0;
    ctx->histCacheR1 = r1;
    if (bins > ctx->histBins)
    {
        LVR_HIP(ctx, hipStreamSynchronize(ctx->histStream));
        (void)hipFree(ctx->d_hist);
        ctx->d_hist = nullptr;
        ctx->histValid = false;
        LVR_HIP(ctx, hipMalloc((void**)&ctx->d_hist, sizeof(unsigned long long) * bins));
        ctx->histBins = bins;
    }

    std::vector<int32_t> nextSlots(next.size());
    for (size_t i = 0; i < next.size(); ++i)
        nextSlots[i] = slotOf(ctx, next[i]);
    // entering and leaving bricks, with the slot each was (or will be)
    // counted from
    std::vector<std::pair<uint64_t, int32_t>> added, removed;
    bool delta = ctx->histValid && ctx->histLast == bins &&
                 ctx->histSetSlots.size() == ctx->histSet.size() &&
                 ctx->histSetVersions.size() == ctx->histSet.size();
    // a row still holds the counts the running sum took from it
    auto same = [&](size_t j) {
        const size_t slot = size_t(ctx->histSetSlots[j]);
        return ctx->histCachedId[slot] == ctx->histSet[j] &&
               ctx->histRowVersion[slot] == ctx->histSetVersions[j];
    };
    std::vector<size_t> removedAt;
    if (delta)
    {
        size_t i = 0, j = 0;
        const std::vector<uint64_t>& prev = ctx->histSet;
        while (i < next.size() || j < prev.size())
        {
            if (j == prev.size() || (i < next.size() && next[i] < prev[j]))
            {
                added.emplace_back(next[i], nextSlots[i]);
                ++i;
            }
            else if (i == next.size() || prev[j] < next[i])
            {
                removed.emplace_back(prev[j], ctx->histSetSlots[j]);
                removedAt.push_back(j);
                ++j;
            }
            else
            {
                // the same brick in both sets: unchanged unless it moved slot;
                // a brick rewritten since has lost the row its counts in the
                // running sum came from, so the sum starts afresh
                if (nextSlots[i] != ctx->histSetSlots[j])
                {
                    added.emplace_back(next[i], nextSlots[i]);
                    removed.emplace_back(prev[j], ctx->histSetSlots[j]);
                    removedAt.push_back(j);
                }
                else if (!same(j))
                    delta = false;
                ++i;
                ++j;
            }
        }
        for (size_t j : removedAt)
            delta = delta && same(j);
        delta = delta && added.size() + removed.size() < next.size();
    }
    if (!delta)
    {
        added.clear();
        removed.clear();
        for (size_t i = 0; i < next.size(); ++i)
            added.emplace_back(next[i], nextSlots[i]);
    }

    // rows to bin: bricks counted now whose slot does not hold their counts
    int halves = 1;
    uint32_t tileStride = 1;
    jobShape(ctx, bins, &halves, &tileStride);
    const int binsPerJob = bins / halves;
    std::vector<int32_t> binSlots;
    std::vector<HistJob> jobs;
    for (const auto& a : added)
        if (ctx->histCachedId[size_t(a.second)] != a.first)
        {
            ctx->histCachedId[size_t(a.second)] = a.first;
            ++ctx->histRowVersion[size_t(a.second)];
            const int first = int(binSlots.size());
            binSlots.push_back(a.second);
            for (int h = 0; h < halves; ++h)
                jobs.push_back({first, 1, 0, h * binsPerJob, 0, {0, 0, 0}});
        }
    std::vector<SumTerm> terms;
    terms.reserve(added.size() + removed.size());
    for (const auto& a : added)
        terms.push_back({a.second, 3 * (depth - int(decodeNode(a.first).level) - 1), 0, 0});
    for (const auto& r : removed)
        terms.push_back({r.second, 3 * (depth - int(decodeNode(r.first).level) - 1), 1, 0});

    ctx->histLast = bins;
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
    if (!jobs.empty() || !terms.empty())
    {
        const size_t jobBytes = sizeof(HistJob) * jobs.size();
        const size_t slotBytes = (sizeof(int32_t) * binSlots.size() + 15) & ~size_t(15);
        const size_t termBytes = sizeof(SumTerm) * terms.size();
        uint8_t *h = nullptr, *d = nullptr;
        int r = 0;
        int rc = histRing(ctx, jobBytes + slotBytes + termBytes, &h, &d, &r);
        if (rc)
            return rc;
        std::memcpy(h, jobs.data(), jobBytes);
        std::memcpy(h + jobBytes, binSlots.data(), sizeof(int32_t) * binSlots.size());
        std::memcpy(h + jobBytes + slotBytes, terms.data(), termBytes);
        if (!jobs.empty())
        {
            const int M = int(ctx->vi.max_block[0]);
            const int ov = int(ctx->vi.overlap[0]);
            const size_t lds = sizeof(uint32_t) * binsPerJob;
            const float perBin = (r1 - r0) / float(bins - 1);
            const HistJob* dj = reinterpret_cast<const HistJob*>(d);
            const int32_t* ds = reinterpret_cast<const int32_t*>(d + jobBytes);
            auto launch = [&](auto kernel) {
                (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
                hipLaunchKernelGGL(kernel, dim3(unsigned(jobs.size())), dim3(1024), lds,
                                   ctx->histStream, dj, ds, ctx->d_pool, ctx->slotStride, M, ov,
                                   tileStride, binsPerJob, bins, r0, perBin, ctx->layout,
                                   ctx->d_hist, ctx->d_histCache);
            };
            if (dt == LVR_DT_UINT8)
                launch(&histKernel<uint8_t, true>);
            else if (dt == LVR_DT_UINT16)
                launch(&histKernel<uint16_t, true>);
            else
                launch(&histKernel<float, true>);
            LVR_HIP(ctx, hipGetLastError());
        }
        if (!terms.empty())
        {
            const int threads = std::min(256, bins / 4);
            const dim3 grid(unsigned((bins / 4 + threads - 1) / threads),
                            unsigned((terms.size() + kSumTermsPerBlock - 1) / kSumTermsPerBlock));
            hipLaunchKernelGGL(histSumKernel, grid, dim3(unsigned(threads)), 0, ctx->histStream,
                               reinterpret_cast<const SumTerm*>(d + jobBytes + slotBytes),
                               int(terms.size()), ctx->d_histCache, bins, ctx->d_hist);
            LVR_HIP(ctx, hipGetLastError());
        }
        LVR_HIP(ctx, hipEventRecord(ctx->histRingDone[r], ctx->histStream));
    }
    if (rec)
    {
        LVR_HIP(ctx, hipEventRecord(rec->ev[4], ctx->histStream));
        rec->histTimed = true;
    }
    LVR_HIP(ctx, hipEventRecord(ctx->histDone, ctx->histStream));
    std::vector<uint32_t> versions(nextSlots.size());
    for (size_t i = 0; i < nextSlots.size(); ++i)
        versions[i] = ctx->histRowVersion[size_t(nextSlots[i])];
    ctx->histSet.swap(next);
    ctx->histSetSlots.swap(nextSlots);
    ctx->histSetVersions.swap(versions);
    ctx->histValid = true;
    ctx->histGeneration = ctx->poolGeneration;
    ctx->histRange0 = r0;
    ctx->histRange1 = r1;
    ctx->histDeltaBricks = delta ? added.size() + removed.size() : next.size();
    ctx->histBinned = binSlots.size();
    return LVR_OK;
}
int cacheBrickHistograms(lvr_ctx* ctx, const uint64_t* ids, const int32_t* slots, int n,
                         hipStream_t st)
{
    const int dt = ctx->vi.data_type;
    // float rows depend on the TF range, known only when a frame is counted
    if (n <= 0 || (dt != LVR_DT_UINT8 && dt != LVR_DT_UINT16))
        return LVR_OK;
    const int bins = dt == LVR_DT_UINT8 ? 256 : 65536;
    const lvr_volume_info& vi = ctx->vi;
    if (!cacheFits(ctx, bins) || vi.overlap[0] != vi.overlap[1] ||
        vi.overlap[0] != vi.overlap[2] || vi.max_block[0] != vi.max_block[1] ||
        vi.max_block[0] != vi.max_block[2])
        return LVR_OK;
    int rc = ensureHistCache(ctx, bins);
    if (rc)
        return rc;
    int halves = 1;
    uint32_t tileStride = 1;
    jobShape(ctx, bins, &halves, &tileStride);
    const int binsPerJob = bins / halves;
    std::vector<HistJob> jobs;
    jobs.reserve(size_t(n) * size_t(halves));
    for (int i = 0; i < n; ++i)
    {
        ctx->histCachedId[size_t(slots[i])] = ids[i];
        ++ctx->histRowVersion[size_t(slots[i])];
        for (int h = 0; h < halves; ++h)
            jobs.push_back({i, 1, 0, h * binsPerJob, 0, {0, 0, 0}});
    }
    const size_t jobBytes = sizeof(HistJob) * jobs.size();
    uint8_t *h = nullptr, *d = nullptr;
    int r = 0;
    rc = histRing(ctx, jobBytes + sizeof(int32_t) * size_t(n), &h, &d, &r);
    if (rc)
        return rc;
    std::memcpy(h, jobs.data(), jobBytes);
    std::memcpy(h + jobBytes, slots, sizeof(int32_t) * size_t(n));
    const size_t lds = sizeof(uint32_t) * binsPerJob;
    auto launch = [&](auto kernel) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        hipLaunchKernelGGL(kernel, dim3(unsigned(jobs.size())), dim3(1024), lds, st,
                           reinterpret_cast<const HistJob*>(d),
                           reinterpret_cast<const int32_t*>(d + jobBytes), ctx->d_pool,
                           ctx->slotStride, int(vi.max_block[0]), int(vi.overlap[0]), tileStride,
                           binsPerJob, bins, 0.f, 1.f, ctx->layout, ctx->d_hist,
                           ctx->d_histCache);
    };
    if (dt == LVR_DT_UINT8)
        launch(&histKernel<uint8_t, true>);
    else
        launch(&histKernel<uint16_t, true>);
    LVR_HIP(ctx, hipGetLastError());
    LVR_HIP(ctx, hipEventRecord(ctx->histRingDone[r], st));
    return LVR_OK;
}
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
    if (cacheFits(ctx, bins))
        return histogramCached(ctx, next, bins, r0, r1);
    ctx->histBinned = 0;
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
                               ctx->layout, ctx->d_hist, nullptr);
        };
        if (dt == LVR_DT_UINT8)
            launch(&histKernel<uint8_t, false>);
        else if (dt == LVR_DT_UINT16)
            launch(&histKernel<uint16_t, false>);
        else
            launch(&histKernel<float, false>);
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
    ctx->histSetSlots.clear(); // counted from the pool, not from cached rows
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
