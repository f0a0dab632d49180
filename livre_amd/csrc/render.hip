// Ray casting over the ordered brick list on gfx950: one frame (lvr_render)
// or a batch of frames of different volumes, viewports and transfer
// functions in one persistent launch (lvr_batch_*).
//
// The reference draws the back faces of every brick in front-to-back order,
// one draw call plus a full memory barrier per brick, and each fragment runs
// fragRayCast.glsl over its brick into an RGBA32F image
// (RayCastRenderer.cpp:467-571, fragRayCast.glsl:113-207). A pixel therefore
// composites the bricks its ray crosses in list order. Here the same per-pixel
// result is produced by:
//   1. the host projecting each brick to a conservative range of screen tiles;
//   2. binCount / scan / binFill building, per tile, the list of brick
//      indices in global order (one wavefront per tile, ballot compaction
//      keeps the order without sorting), for all frames of a batch at once;
//   3. a persistent ray-cast kernel: wavefronts pull 8x8 ray tiles from an
//      atomic work queue, stage the transfer function in LDS, read brick
//      records with scalar loads, composite in FP32 registers, exit early
//      per ray and per wavefront (ballot) and store RGBA8 directly (no
//      RGBA32F round trip, fragTexCopy.glsl:11-15).
// The per-sample arithmetic follows the shader expression by expression
// (raycast_core.h): sample positions, texel indices, TF indices and the
// early-exit opacity are rounded exactly as written; the linear filter of the
// TF and the colour blend use explicit FMAs, which moves colours by an ulp.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "raycast_core.h"

namespace lvr
{
namespace
{
// One wavefront per global tile: count the bricks of its frame whose tile
// range covers the tile.
// Binning in two levels, both keeping the front-to-back brick order that
// the shader's draw order gives (RayCastRenderer.cpp:550-571): first every
// super-tile of kSuperTile x kSuperTile ray tiles scans the frame's bricks
// in order and keeps those whose screen rectangle meets it; then every tile
// scans only its super-tile's list. A wavefront tests 64 bricks per step and
// compacts the hits with a ballot, so the lists come out ordered.
constexpr int kSuperTile = 8;

__device__ inline int findBySuper(const FrameUniforms* __restrict__ reqs, int nReq, int gsuper)
{
    int lo = 0, hi = nReq - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (reqs[mid].superBase <= gsuper)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// FILL = false: count into out[gsuper]; FILL = true: write the ordered list
// at offsets[gsuper].
template <bool FILL>
__global__ __launch_bounds__(256) void superBinKernel(const FrameUniforms* __restrict__ reqs,
                                                      int nReq, const int4* __restrict__ ranges,
                                                      int totalSupers,
                                                      const int32_t* __restrict__ offsets,
                                                      int32_t* __restrict__ out)
{
    const int gsuper = __builtin_amdgcn_readfirstlane(
        int(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave));
    if (gsuper >= totalSupers)
        return;
    const int r = __builtin_amdgcn_readfirstlane(findBySuper(reqs, nReq, gsuper));
    const FrameUniforms& U = reqs[r];
    const int lane = threadIdx.x & (kWave - 1);
    const int st = gsuper - U.superBase;
    const int x0 = (st % U.superX) * kSuperTile, y0 = (st / U.superX) * kSuperTile;
    const int x1 = x0 + kSuperTile - 1, y1 = y0 + kSuperTile - 1;
    const uint64_t below = (uint64_t(1) << lane) - 1;
    int pos = FILL ? offsets[gsuper] : 0;
    for (int base = 0; base < U.nBricks; base += kWave)
    {
        const int i = base + lane;
        bool hit = false;
        if (i < U.nBricks)
        {
            const int4 q = ranges[U.brickBase + i];
            // empty ranges (x > z: brick off screen or not resident) never
            // hit; the host sizes the lists by the same rule
            hit = q.x <= q.z && q.y <= q.w && q.x <= x1 && q.z >= x0 && q.y <= y1 && q.w >= y0;
        }
        const uint64_t mask = __ballot(hit);
        if (FILL && hit)
            out[pos + __popcll(mask & below)] = U.brickBase + i;
        pos += __popcll(mask);
    }
    if (!FILL && lane == 0)
        out[gsuper] = pos;
}

// Second level: FILL = false counts into out[gtile], FILL = true writes the
// tile's ordered list (global brick indices) at offsets[gtile].
template <bool FILL>
__global__ __launch_bounds__(256) void tileBinKernel(const FrameUniforms* __restrict__ reqs,
                                                     int nReq, const int4* __restrict__ ranges,
                                                     int totalTiles,
                                                     const int32_t* __restrict__ superOffsets,
                                                     const int32_t* __restrict__ superCounts,
                                                     const int32_t* __restrict__ superLists,
                                                     const int32_t* __restrict__ offsets,
                                                     int32_t* __restrict__ out)
{
    const int gtile = __builtin_amdgcn_readfirstlane(
        int(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave));
    if (gtile >= totalTiles)
        return;
    const int r = __builtin_amdgcn_readfirstlane(findByTile(reqs, nReq, gtile));
    const FrameUniforms& U = reqs[r];
    const int lane = threadIdx.x & (kWave - 1);
    const int tile = gtile - U.tileBase;
    const int tx = tile % U.tilesX, ty = tile / U.tilesX;
    const int gsuper = U.superBase + (ty / kSuperTile) * U.superX + tx / kSuperTile;
    const int beg = superOffsets[gsuper], n = superCounts[gsuper];
    const uint64_t below = (uint64_t(1) << lane) - 1;
    int pos = FILL ? offsets[gtile] : 0;
    for (int base = 0; base < n; base += kWave)
    {
        const int k = base + lane;
        bool hit = false;
        int b = 0;
        if (k < n)
        {
            b = superLists[beg + k];
            const int4 q = ranges[b];
            hit = tx >= q.x && tx <= q.z && ty >= q.y && ty <= q.w;
        }
        const uint64_t mask = __ballot(hit);
        if (FILL && hit)
            out[pos + __popcll(mask & below)] = b;
        pos += __popcll(mask);
    }
    if (!FILL && lane == 0)
        out[gtile] = pos;
}

// Exclusive scan of the tile counts, one workgroup of 1024 threads. Also
// clears the batch's counters (composite steps, work queue, per-frame steps)
// for the kernels that follow on the stream.
__global__ __launch_bounds__(1024) void scanKernel(const int32_t* __restrict__ counts, int n,
                                                   int32_t* __restrict__ offsets,
                                                   unsigned long long* __restrict__ counters,
                                                   int nCounters)
{
    // one workgroup: exclusive scan in passes of 4096 counts (4 per thread,
    // read and written coalesced), wavefront shuffles, then the 16
    // wavefront totals through LDS; a running carry joins the passes
    __shared__ int32_t waveTotal[16];
    for (int i = threadIdx.x; i < nCounters; i += blockDim.x)
        counters[i] = 0;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int carry = 0;
    for (int base = 0; base < n; base += 4096)
    {
        const int i0 = base + 4 * t;
        int v[4];
        if (i0 + 3 < n)
        {
            const int4 q = *reinterpret_cast<const int4*>(counts + i0);
            v[0] = q.x;
            v[1] = q.y;
            v[2] = q.z;
            v[3] = q.w;
        }
        else
            for (int k = 0; k < 4; ++k)
                v[k] = i0 + k < n ? counts[i0 + k] : 0;
        const int mine = v[0] + v[1] + v[2] + v[3];
        int inc = mine;
        for (int d = 1; d < 64; d <<= 1)
        {
            const int o = __shfl_up(inc, d, 64);
            if (lane >= d)
                inc += o;
        }
        if (lane == 63)
            waveTotal[wave] = inc;
        __syncthreads();
        int before = 0, total = 0;
        for (int w = 0; w < 16; ++w)
        {
            const int x = waveTotal[w];
            before += w < wave ? x : 0;
            total += x;
        }
        int run = carry + before + inc - mine;
        for (int k = 0; k < 4; ++k)
        {
            if (i0 + k < n)
                offsets[i0 + k] = run;
            run += v[k];
        }
        carry += total;
        __syncthreads();
    }
}

// Both binning levels with their scans, on stream st. The second scan also
// clears nCounters counters for the kernels that follow.
static void enqueueBinning(hipStream_t st, const FrameUniforms* reqs, int nReq,
                           const int4* ranges, int supers, int tiles, int32_t* sCount,
                           int32_t* sOffset, int32_t* sList, int32_t* tCount, int32_t* tOffset,
                           int32_t* tList, unsigned long long* counters, int nCounters)
{
    const int sBlocks = (supers + 3) / 4, tBlocks = (tiles + 3) / 4;
    hipLaunchKernelGGL(superBinKernel<false>, dim3(sBlocks), dim3(256), 0, st, reqs, nReq,
                       ranges, supers, nullptr, sCount);
    hipLaunchKernelGGL(scanKernel, dim3(1), dim3(1024), 0, st, sCount, supers, sOffset, counters,
                       0);
    hipLaunchKernelGGL(superBinKernel<true>, dim3(sBlocks), dim3(256), 0, st, reqs, nReq, ranges,
                       supers, sOffset, sList);
    hipLaunchKernelGGL(tileBinKernel<false>, dim3(tBlocks), dim3(256), 0, st, reqs, nReq, ranges,
                       tiles, sOffset, sCount, sList, nullptr, tCount);
    hipLaunchKernelGGL(scanKernel, dim3(1), dim3(1024), 0, st, tCount, tiles, tOffset, counters,
                       nCounters);
    hipLaunchKernelGGL(tileBinKernel<true>, dim3(tBlocks), dim3(256), 0, st, reqs, nReq, ranges,
                       tiles, sOffset, sCount, sList, tOffset, tList);
}

// The order of a frame's work super-tiles (kSuper x kSuper ray tiles) for
// its work queue: the heaviest quarter first, then the rest, each part in
// screen order. The weight of a super-tile is the (tile, brick) pairs binned
// to its tiles; a persistent launch then ends on light tiles, and tiles
// marched at the same time stay neighbours that share voxel lines in L2. The
// order changes when a tile is marched, never what it renders.
// One workgroup: a 1024-bucket weight histogram gives the threshold, then a
// stable partition by a block-wide scan.
__global__ __launch_bounds__(1024) void superOrderKernel(const int32_t* __restrict__ tCount,
                                                         int tilesX, int tilesY,
                                                         int32_t* __restrict__ order)
{
    constexpr int kBuckets = 1024;
    __shared__ int32_t hist[kBuckets];
    __shared__ int32_t waveSums[1024 / kWave];
    __shared__ int32_t shared[2];
    const int superX = (tilesX + kSuper - 1) / kSuper;
    const int nSupers = superX * ((tilesY + kSuper - 1) / kSuper);
    const int t = int(threadIdx.x), lane = t & (kWave - 1), wave = t / kWave;
    auto weight = [&](int sIdx) {
        const int tx0 = (sIdx % superX) * kSuper, ty0 = (sIdx / superX) * kSuper;
        int w = 0;
        for (int ty = ty0; ty < min(ty0 + kSuper, tilesY); ++ty)
            for (int tx = tx0; tx < min(tx0 + kSuper, tilesX); ++tx)
                w += tCount[ty * tilesX + tx];
        return min(w, kBuckets - 1);
    };
    hist[t] = 0;
    __syncthreads();
    for (int sIdx = t; sIdx < nSupers; sIdx += 1024)
        atomicAdd(&hist[weight(sIdx)], 1);
    __syncthreads();
    if (t == 0)
    {
        // the lowest weight whose super-tiles and all heavier ones make up
        // at least a quarter
        int acc = 0, th = kBuckets - 1;
        for (; th > 0; --th)
        {
            acc += hist[th];
            if (4 * acc >= nSupers)
                break;
        }
        shared[0] = th;
        shared[1] = 0; // heavy super-tiles placed so far
    }
    __syncthreads();
    const int th = shared[0];
    // heavy ones first, in screen order: a block-wide exclusive scan per
    // pass of 1024 super-tiles, carried across passes
    for (int pass = 0; pass < 2; ++pass)
    {
        int base = pass == 0 ? 0 : shared[1];
        for (int s0 = 0; s0 < nSupers; s0 += 1024)
        {
            const int sIdx = s0 + t;
            const bool heavy = sIdx < nSupers && weight(sIdx) >= th;
            const bool mine = sIdx < nSupers && (pass == 0 ? heavy : !heavy);
            const unsigned long long m = __ballot(mine);
            const int before = __popcll(m & ((1ull << lane) - 1ull));
            if (lane == 0)
                waveSums[wave] = __popcll(m);
            __syncthreads();
            int off = 0;
            for (int w = 0; w < wave; ++w)
                off += waveSums[w];
            int total = 0;
            for (int w = 0; w < 1024 / kWave; ++w)
                total += waveSums[w];
            if (mine)
                order[base + off + before] = sIdx;
            base += total;
            __syncthreads();
        }
        if (t == 0 && pass == 0)
            shared[1] = base;
        __syncthreads();
    }
}

__device__ inline int findByBrick(const FrameUniforms* __restrict__ reqs, int nReq, int brick)
{
    int lo = 0, hi = nReq - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (reqs[mid].brickBase <= brick)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Brick records of every request from the keys the host sent (one thread
// per brick; requests are sorted by brickBase).
__global__ __launch_bounds__(256) void brickRecordsKernel(const FrameUniforms* __restrict__ reqs,
                                                          int nReq,
                                                          const BrickKey* __restrict__ keys,
                                                          int nBricks,
                                                          BrickEntry* __restrict__ out)
{
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= nBricks)
        return;
    const FrameUniforms& U = reqs[findByBrick(reqs, nReq, i)];
    out[i] = makeBrickEntry(U, keys[i]);
}

static void enqueueRecords(hipStream_t st, const FrameUniforms* reqs, int nReq,
                           const BrickKey* keys, int nBricks, BrickEntry* out)
{
    if (nBricks > 0)
        hipLaunchKernelGGL(brickRecordsKernel, dim3((nBricks + 255) / 256), dim3(256), 0, st,
                           reqs, nReq, keys, nBricks, out);
}

__device__ inline int nextItem(unsigned long long* counter, int lane)
{
    unsigned long long got = 0;
    if (lane == 0)
        got = atomicAdd(counter, 1ull);
    return __builtin_amdgcn_readfirstlane(int(__shfl(got, 0, kWave)));
}

// Persistent single-frame kernel: each workgroup (4 to 16 wavefronts) stages
// the whole transfer function in LDS once (up to 8192 FP32 entries), then
// every wavefront pulls tiles from the frame's atomic counter until none are
// left. Tiles differ by orders of magnitude in work (empty sky vs rays
// through the whole volume), so pulling them dynamically keeps every CU busy
// to the end. Up to 8 wavefronts per SIMD fit in <= 64 VGPRs; fewer keep
// fewer voxel lines in flight per CU, which the 32 KB L1 holds better.
// The windowed march (castTileWindow) keeps two bricks' segments per lane and
// is held to 80 VGPRs, 6 wavefronts per SIMD: a workgroup of 2 * 5
// wavefronts spreads 3, 3, 2, 2 over the SIMDs, so two of them need room
// for 6 on a SIMD (at 81 VGPRs, allocated as 88, only one fits per CU and
// the kernel runs 25% slower). The GL_LINEAR march keeps two samples' eight
// texels in flight and may take up to 128 VGPRs: its workgroups of 2 * w
// wavefronts put at most 4 on a SIMD for every w up to 8. The others fit 8.
template <bool LINEAR, int D, bool CLS>
constexpr int kRaycastWavesPerEu = LINEAR ? 4 : ((!CLS && D == 2) ? 6 : 8);

template <typename T, bool LINEAR, int D, bool CLS>
__global__ __launch_bounds__(1024) __attribute__((
    amdgpu_waves_per_eu(kRaycastWavesPerEu<LINEAR, D, CLS>))) void raycastKernel(
    const FrameUniforms U, const BrickEntry* __restrict__ bricks,
    const int32_t* __restrict__ offsets, const int32_t* __restrict__ counts,
    const int32_t* __restrict__ lists, unsigned long long* __restrict__ counters)
{
    // LDS: the transfer function, or with CLS the 8-bit classified table
    // (16-bit tables stay in global memory)
    extern __shared__ float4 sTf[];
    const float4* table = sTf;
    // the marches read the slot offsets of their texels from per-axis
    // tables placed after the transfer function or the 8-bit classified
    // table (rayLds)
    constexpr bool kTab = true;
    uint32_t* offTab = reinterpret_cast<uint32_t*>(
        sTf + (!CLS ? U.nTf + 2 : (sizeof(T) == 1 ? 256 : 0)));
    if (!CLS)
    {
        for (int i = threadIdx.x; i < U.nTf + 2; i += blockDim.x)
            sTf[i] = U.tf[i];
    }
    else if (sizeof(T) == 1)
    {
        for (int i = threadIdx.x; i < 256; i += blockDim.x)
            sTf[i] = U.cls[i];
    }
    else
        table = U.cls;
    if (kTab)
        offsetTables<T>(U, offTab, int(threadIdx.x), int(blockDim.x));
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    uint32_t nsamp = 0;
    for (;;)
    {
        const int item = nextItem(&counters[1], lane);
        if (item >= U.nItems)
            break;
        const int tile = itemToTile(U, item);
        if (tile < 0)
            continue;
        if (!LINEAR && !CLS)
            castNearestTile<T, D, kTab>(U, tile, lane, bricks, offsets, counts, lists, table,
                                        nsamp, offTab);
        else if (LINEAR && D == 2)
            castTileWindow<T, kTab, true>(U, tile, lane, bricks, offsets, counts, lists, table,
                                          offTab, nsamp);
        else
            castTile<T, LINEAR, D, CLS, kTab>(U, tile, lane, bricks, offsets, counts, lists,
                                              table, nsamp, offTab);
    }
    const unsigned long long s = waveSum(nsamp);
    if (lane == 0 && s)
        atomicAdd(&counters[0], s);
}

// Dynamic LDS of a single-frame ray cast: the transfer function with its
// edge copies (or, with a classified table, the 8-bit table; a 16-bit one
// stays in global memory), then the per-axis slot-offset tables of the
// GL_NEAREST marches.
inline size_t offsetTablesLds(const FrameUniforms& U)
{
    return sizeof(uint32_t) * size_t(U.texDim[0] + U.texDim[1] + U.texDim[2]);
}
inline size_t rayLds(const FrameUniforms& U)
{
    return size_t(U.nTf + 2) * sizeof(float4) + offsetTablesLds(U);
}

// LDS entries per wavefront for a batched request's transfer function: a
// 256-entry table with its two edge copies.
constexpr int kBatchTf = 258;

// The batch kernel's voxel-type branches (8-bit, 16-bit, float) and the
// fetch depths it is built with (a depth of 1 runs as 2: the same samples).
inline int batchTypeClass(int dataType)
{
    return dataType == LVR_DT_UINT8 ? 0 : (dataType == LVR_DT_UINT16 ? 1 : 2);
}
inline int batchDepth(int depth) { return depth >= 3 ? 3 : 2; }


// One tile of a batched request, brick by brick: its transfer function (or
// 8-bit classified table) from this wavefront's LDS slice or, when it did not
// fit there, from global memory. The two are separate instantiations so that
// LDS reads stay ds_read. (Per-wavefront slot-offset tables, as the
// single-frame kernel stages per workgroup, made this kernel 4% slower on the
// batched uneven-request workload.)
template <typename T, int D, bool CLS>
__device__ __forceinline__ void batchCast(const FrameUniforms& U, int tile, int lane,
                                          const BrickEntry* __restrict__ bricks,
                                          const int32_t* __restrict__ offsets,
                                          const int32_t* __restrict__ counts,
                                          const int32_t* __restrict__ lists, const float4* myTf,
                                          bool lds, uint32_t& nsamp)
{
    if (lds)
        castTile<T, false, D, CLS>(U, tile, lane, bricks, offsets, counts, lists, myTf, nsamp);
    else
        castTile<T, false, D, CLS>(U, tile, lane, bricks, offsets, counts, lists, U.tf, nsamp);
}

// Persistent batch kernel: the work items of all frames of a batch form one
// queue. A wavefront looks up the frame of each item in the request table,
// restages that frame's transfer function in its own LDS slice when the
// frame changes (tables above kBatchTf entries are read from global memory),
// and dispatches on the frame's voxel type, each type with its own number of
// fetches in flight (D8, D16, DF: the autotuner's choice for that type).
// Composite steps are counted per frame. A workgroup holds 2 * wavesPerSimd
// (4 to 16) wavefronts.
// The batch kernel marches brick by brick (castTile) at every depth.
template <int D8, int D16, int DF>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void batchKernel(
    const FrameUniforms* __restrict__ reqs, int nReq, int totalItems,
    const BrickEntry* __restrict__ bricks, const int32_t* __restrict__ offsets,
    const int32_t* __restrict__ counts, const int32_t* __restrict__ lists,
    unsigned long long* __restrict__ counters)
{
    __shared__ float4 sTf[kSuper * kSuper][kBatchTf];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / kWave);
    float4* myTf = sTf[wave];
    int cur = -1;
    uint32_t nsamp = 0;
    for (;;)
    {
        const int item = nextItem(&counters[1], lane);
        if (item >= totalItems)
            break;
        const int r = __builtin_amdgcn_readfirstlane(findByItem(reqs, nReq, item));
        const FrameUniforms& U = reqs[r];
        if (r != cur)
        {
            if (cur >= 0)
            {
                const unsigned long long s = waveSum(nsamp);
                if (lane == 0 && s)
                {
                    atomicAdd(&counters[0], s);
                    atomicAdd(&counters[2 + cur], s);
                }
                nsamp = 0;
            }
            cur = r;
            if (U.dataType == LVR_DT_UINT8 && U.classify == 1)
            {
                // 8-bit request: its classified table, built from the
                // request's TF into this wavefront's slice
                const Classify c = makeClassify(U);
                for (int i = lane; i < 256; i += kWave)
                    myTf[i] = c(U.tf, float(i));
            }
            else if (U.nTf + 2 <= kBatchTf)
            {
                // this wavefront's slice only: ordered by the wave barrier
                for (int i = lane; i < U.nTf + 2; i += kWave)
                    myTf[i] = U.tf[i];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        const int tile = itemToTile(U, item - U.itemBase);
        if (tile < 0)
            continue;
        const bool lds = U.nTf + 2 <= kBatchTf;
        switch (U.dataType)
        {
        case LVR_DT_UINT8:
            if (U.classify == 1)
                batchCast<uint8_t, D8, true>(U, tile, lane, bricks, offsets, counts, lists, myTf,
                                             true, nsamp);
            else
                batchCast<uint8_t, D8, false>(U, tile, lane, bricks, offsets, counts, lists,
                                              myTf, lds, nsamp);
            break;
        case LVR_DT_UINT16:
            batchCast<uint16_t, D16, false>(U, tile, lane, bricks, offsets, counts, lists, myTf,
                                            lds, nsamp);
            break;
        default:
            batchCast<float, DF, false>(U, tile, lane, bricks, offsets, counts, lists, myTf, lds,
                                        nsamp);
            break;
        }
    }
    if (cur >= 0)
    {
        const unsigned long long s = waveSum(nsamp);
        if (lane == 0 && s)
        {
            atomicAdd(&counters[0], s);
            atomicAdd(&counters[2 + cur], s);
        }
    }
}

// Sort-last compositing (eq::Compositor::blendFrames in depth order,
// Channel.cpp:464-603): front-to-back "over" of n premultiplied RGBA32F
// layers, front first, then fragTexCopy's RGBA8 rounding. One thread per
// pixel, each layer read once, coalesced: HBM-bound.
constexpr int kMaxCompositeLayers = 16;
struct LayerSet
{
    const float4* p[kMaxCompositeLayers];
};

__global__ __launch_bounds__(256) void compositeKernel(LayerSet L, int n, int npix,
                                                       uint32_t* __restrict__ out)
{
    for (int i = int(blockIdx.x * blockDim.x + threadIdx.x); i < npix;
         i += int(gridDim.x * blockDim.x))
    {
        float r = 0.f, g = 0.f, b = 0.f, a = 0.f;
        for (int k0 = 0; k0 < n; k0 += 8)
        {
            // eight layers' loads in flight before their blends; each layer
            // is read once, so the loads bypass the caches' retention
            typedef float vec4 __attribute__((ext_vector_type(4)));
            vec4 v4[8];
#pragma unroll
            for (int j = 0; j < 8; ++j)
                v4[j] = k0 + j < n ? __builtin_nontemporal_load(
                                         reinterpret_cast<const vec4*>(L.p[k0 + j]) + i)
                                   : vec4{0.f, 0.f, 0.f, 0.f};
            float4 c[8];
#pragma unroll
            for (int j = 0; j < 8; ++j)
                c[j] = make_float4(v4[j].x, v4[j].y, v4[j].z, v4[j].w);
#pragma unroll
            for (int j = 0; j < 8; ++j)
            {
                const float rem = 1.0f - a;
                r = __builtin_fmaf(c[j].x, rem, r);
                g = __builtin_fmaf(c[j].y, rem, g);
                b = __builtin_fmaf(c[j].z, rem, b);
                a = __builtin_fmaf(c[j].w, rem, a);
            }
        }
        const float v[4] = {r, g, b, a};
        uint32_t packed = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            packed |= uint32_t(floorf(fminf(1.0f, fmaxf(0.0f, v[q])) * 255.0f + 0.5f)) << (8 * q);
        out[i] = packed;
    }
}

// Classified table of an 8- or 16-bit volume: entry v = the classification
// of raw value v, computed by the same code as the in-loop path.
__global__ __launch_bounds__(256) void classifyKernel(Classify c, const float4* __restrict__ tf,
                                                      int nValues, float4* __restrict__ out)
{
    const int v = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (v < nValues)
        out[v] = c(tf, float(v));
}

// The classification constants of a frame, as makeClassify forms them on
// the device.
inline Classify makeClassifyHost(const FrameUniforms& U)
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

template <typename T, bool LINEAR, int D, bool CLS>
void launchRaycastAs(int blocks, size_t lds, hipStream_t st, const FrameUniforms& U,
                     const BrickEntry* bricks, const int32_t* offsets, const int32_t* counts,
                     const int32_t* lists, unsigned long long* counters)
{
    // two workgroups per CU, each with wavesPerSimd * 2 wavefronts (spread
    // over the CU's 4 SIMDs)
    const int threads = U.wavesPerSimd * 2 * kWave;
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&raycastKernel<T, LINEAR, D, CLS>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    hipLaunchKernelGGL((raycastKernel<T, LINEAR, D, CLS>), dim3(blocks), dim3(threads), lds, st,
                       U, bricks, offsets, counts, lists, counters);
}

template <typename T, bool CLS>
void launchNearest(int blocks, size_t lds, hipStream_t st, const FrameUniforms& U,
                   const BrickEntry* bricks, const int32_t* offsets, const int32_t* counts,
                   const int32_t* lists, unsigned long long* counters)
{
    if (U.depth == 1)
        launchRaycastAs<T, false, 1, CLS>(blocks, lds, st, U, bricks, offsets, counts, lists,
                                          counters);
    else if (U.depth == 3)
        launchRaycastAs<T, false, 3, CLS>(blocks, lds, st, U, bricks, offsets, counts, lists,
                                          counters);
    else
        launchRaycastAs<T, false, 2, CLS>(blocks, lds, st, U, bricks, offsets, counts, lists,
                                          counters);
}

template <typename T>
void launchRaycast(int blocks, size_t lds, hipStream_t st, const FrameUniforms& U,
                   const BrickEntry* bricks, const int32_t* offsets, const int32_t* counts,
                   const int32_t* lists, unsigned long long* counters)
{
    // GL_LINEAR: through the two-brick window at the default depth of 2,
    // brick by brick at 1
    if (U.linear && U.depth == 1)
        launchRaycastAs<T, true, 1, false>(blocks, lds, st, U, bricks, offsets, counts, lists,
                                           counters);
    else if (U.linear)
        launchRaycastAs<T, true, 2, false>(blocks, lds, st, U, bricks, offsets, counts, lists,
                                           counters);
    else
        launchNearest<T, false>(blocks, lds, st, U, bricks, offsets, counts, lists, counters);
}
} // namespace

// ------------------------------------------------ no-allocation frames
// lvr_render_noalloc keeps every buffer its kernels use beyond the inputs and
// the output in one static device array of fixed capacity: the frame's
// uniforms, the brick keys and records, the screen ranges, the two levels of
// tile binning, the counters and the classified table. Capacity per call:
constexpr int kNaBricks = 32768;                  // bricks of the ordered list
constexpr int kNaTiles = 262144;                  // ray tiles (4096 x 4096 at 8 x 8)
constexpr int kNaSupers = 8192;                   // binning super-tiles
constexpr uint64_t kNaTileEntries = 4ull << 20;   // (tile, brick) pairs
constexpr uint64_t kNaSuperEntries = 1ull << 20;  // (super-tile, brick) pairs
constexpr int kNaCls = 65536;                     // classified-table entries
constexpr size_t naUp(size_t x)
{
    return (x + 255) & ~size_t(255);
}
constexpr size_t kNaOffU = 0;
constexpr size_t kNaOffCounters = kNaOffU + naUp(sizeof(FrameUniforms));
constexpr size_t kNaOffKeys = kNaOffCounters + 256;
constexpr size_t kNaOffBricks = kNaOffKeys + naUp(sizeof(BrickKey) * kNaBricks);
constexpr size_t kNaOffRanges = kNaOffBricks + naUp(sizeof(BrickEntry) * kNaBricks);
constexpr size_t kNaOffTileCount = kNaOffRanges + naUp(sizeof(int4) * kNaBricks);
constexpr size_t kNaOffTileOffset = kNaOffTileCount + naUp(4u * kNaTiles);
constexpr size_t kNaOffSuperCount = kNaOffTileOffset + naUp(4u * kNaTiles);
constexpr size_t kNaOffSuperOffset = kNaOffSuperCount + naUp(4u * kNaSupers);
constexpr size_t kNaOffTileList = kNaOffSuperOffset + naUp(4u * kNaSupers);
constexpr size_t kNaOffSuperList = kNaOffTileList + naUp(4u * kNaTileEntries);
constexpr size_t kNaOffCls = kNaOffSuperList + naUp(4u * kNaSuperEntries);
constexpr size_t kNaScratchBytes = kNaOffCls + naUp(sizeof(float4) * kNaCls);
// pinned host staging per context: uniforms, keys, ranges
constexpr size_t kNaStagingBytes =
    sizeof(FrameUniforms) + (sizeof(BrickKey) + sizeof(int4)) * size_t(kNaBricks);

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

// Host side of one frame: uniforms (RayCastRenderer::Impl::onFrameStart,
// RayCastRenderer.cpp:213-386), per-brick records and conservative tile
// ranges of the ordered bricks (appended to `bricks` / `ranges`). Bricks
// that are not resident are skipped, as the reference skips invalid
// textures (RayCastRenderer.cpp:517-522).
static int prepareFrame(lvr_ctx* ctx, const lvr_frustum* f, const int viewport[4],
                        const uint64_t* ids, int n, const lvr_render_params* params,
                        int nTf, float tf0, float tf1, FrameUniforms* U,
                        std::vector<BrickKey>& bricks, std::vector<int4>& ranges,
                        uint64_t* listTotal, uint64_t* superTotal, uint64_t* onScreen)
{
    if (!f || !viewport || !params || n < 0 || (n > 0 && !ids))
        return fail(ctx, LVR_EINVAL, "invalid arguments");
    const int W = viewport[2], H = viewport[3];
    if (W <= 0 || H <= 0 || W > 16384 || H > 16384)
        return fail(ctx, LVR_EINVAL, "invalid viewport");
    if (params->n_samples_per_ray < 0 || params->n_clip_planes < 0 ||
        params->n_clip_planes > 6)
        return fail(ctx, LVR_EINVAL, "invalid render parameters");
    // 0 samples per ray: the reference's rule from the finest rendered level
    int32_t nSamples = 0;
    if (lvr_samples_per_ray(&ctx->vi, ids, n, params->n_samples_per_ray, &nSamples) != LVR_OK)
        return fail(ctx, LVR_EINVAL, "node id outside the octree");
    // the march's termination bound (kMaxSegment) holds up to 2^20
    if (nSamples > kMaxSamplesPerRay)
        return fail(ctx, LVR_EINVAL, "more than 1048576 samples per ray");
    int tw = params->tile_w, th = params->tile_h;
    if (tw == 0 && th == 0)
    {
        tw = 8;
        th = 8;
    }
    if (tw <= 0 || th <= 0 || tw * th != kWave)
        return fail(ctx, LVR_EINVAL, "ray tile must hold 64 rays");
    const int depth = params->samples_per_thread == 0 ? 2 : params->samples_per_thread;
    if (depth < 1 || depth > 3)
        return fail(ctx, LVR_EINVAL, "samples per thread must be 1, 2 or 3");
    if (params->classify < 0 || params->classify > 2)
        return fail(ctx, LVR_EINVAL, "classify must be 0, 1 or 2");
    const int wps = params->waves_per_simd;
    if (wps != 0 && (wps < 2 || wps > 8))
        return fail(ctx, LVR_EINVAL, "waves per SIMD must be 2 to 8");

    std::memset(U, 0, sizeof *U);
    std::memcpy(U->invProj, f->inv_proj, sizeof U->invProj);
    std::memcpy(U->invMV, f->inv_mv, sizeof U->invMV);
    for (int i = 0; i < 3; ++i)
        U->eye[i] = f->eye[i];
    U->nearPlane = f->near_plane;
    U->farPlane = f->far_plane;
    U->width = W;
    U->height = H;
    const lvr_volume_info& vi = ctx->vi;
    for (int i = 0; i < 3; ++i)
    {
        const float half = vi.world_size[i] / 2.0f;
        U->gmin[i] = -half;
        U->gmax[i] = half;
        // TextureObject::Impl::initialize (TextureObject.cpp:76-92)
        const float ov = float(vi.overlap[i]) / float(vi.max_block[i]);
        const float sz = float(vi.max_block[i] - 2u * vi.overlap[i]) / float(vi.max_block[i]);
        U->texMin[i] = ov;
        U->texMax[i] = ov + sz;
        U->texDim[i] = int32_t(vi.max_block[i]);
    }
    U->nSamples = nSamples;
    U->maxSamples = params->max_samples_per_ray;
    for (int i = 0; i < 3; ++i)
    {
        // the same FP32 operations the shader performs per fragment
        U->texExt[i] = U->texMax[i] - U->texMin[i];
        U->texDimF[i] = float(U->texDim[i]);
        U->texLastF[i] = float(U->texDim[i] - 1);
    }
    U->stepSize = 1.0f / float(U->nSamples);
    U->alphaCorrection = float(U->maxSamples) / float(U->nSamples);
    U->exitA = params->early_exit;
    float r0, r1;
    dataRange(vi.data_type, tf0, tf1, &r0, &r1);
    U->mult = 1.0f / (r1 - r0);
    U->add = -r0 / (r1 - r0);
    U->nTf = nTf;
    U->nTfF = float(nTf);
    U->tileW = tw;
    U->tileH = th;
    U->tilesX = (W + tw - 1) / tw;
    U->tilesY = (H + th - 1) / th;
    U->superX = (U->tilesX + kSuperTile - 1) / kSuperTile;
    U->superY = (U->tilesY + kSuperTile - 1) / kSuperTile;
    U->nBricks = n;
    U->linear = params->linear_filter ? 1 : 0;
    U->depth = depth;
    // automatic: a 256-entry table sits in LDS beside the march; a 65536-entry
    // one is a dependent global gather per sample, slower than computing
    U->classify = params->classify != 0 ? params->classify
                                        : (ctx->vi.data_type == LVR_DT_UINT8 ? 1 : 2);
    // the two-brick window (NEAREST, two fetches per ray, classification
    // computed per sample) holds 80 VGPRs, room for six wavefronts per SIMD;
    // five measured fastest on the 2048^3 uint16 orbit. The other variants
    // are built for 64 VGPRs and run best at four.
    const bool window = !U->linear && depth == 2 &&
                        !(U->classify == 1 && (ctx->vi.data_type == LVR_DT_UINT8 ||
                                               ctx->vi.data_type == LVR_DT_UINT16));
    U->wavesPerSimd = wps != 0 ? wps : (window ? 5 : 4);
    U->nClip = params->n_clip_planes;
    std::memcpy(U->clip, params->clip_planes, sizeof U->clip);
    U->slotStride = ctx->slotStride;
    U->layout = ctx->layout;
    {
        const uint32_t bpv = uint32_t(ctx->voxelBytes);
        const uint32_t T0 = ctx->layout.tiles[0], T1 = ctx->layout.tiles[1];
        U->tileK[0] = bpv * 60u;
        U->tileK[1] = bpv * (64u * T0 - 16u);
        U->tileK[2] = bpv * (64u * T0 * T1 - 64u);
    }
    U->dataType = vi.data_type;
    U->pool = ctx->d_pool;
    U->pageTable = ctx->pt.d_slots;
    for (int a = 0; a < 3; ++a)
    {
        U->worldSize[a] = vi.world_size[a];
        U->rootBlocks[a] = vi.root_blocks[a];
    }
    U->nItems = ((U->tilesX + kSuper - 1) / kSuper) * ((U->tilesY + kSuper - 1) / kSuper) *
                kSuper * kSuper;
    U->brickBase = int32_t(bricks.size());

    double mvp[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
        {
            double s = 0;
            for (int k = 0; k < 4; ++k)
                s += double(f->proj[k * 4 + i]) * double(f->mv[j * 4 + k]);
            mvp[j * 4 + i] = s;
        }
    for (int b = 0; b < n; ++b)
    {
        const NodeKey key = decodeNode(ids[b]);
        if (key.level >= vi.depth)
            return fail(ctx, LVR_EINVAL, "node id outside the octree");
        // the key goes to the device, which builds the brick's record
        // (brickRecordsKernel) and reads its slot from the page table at
        // draw time; the host keeps the screen range and the residency test
        BrickKey e;
        e.id = ids[b];
        e.pad = 0;
        uint64_t page = 0;
        e.page = pageIndex(ctx, ids[b], &page) ? int32_t(page) : -1;
        int4 r = make_int4(1, 1, 0, 0); // empty range
        float bmin[3], bmax[3];
        nodeWorldBox(vi, key, bmin, bmax);
        if (e.page >= 0 && slotAndTouch(ctx, ids[b]) >= 0)
        {
            // clip coordinates of the 8 corners: the min corner plus the
            // projected edge vectors
            double base[4], edge[3][4];
            for (int i = 0; i < 4; ++i)
            {
                base[i] = mvp[i] * bmin[0] + mvp[4 + i] * bmin[1] + mvp[8 + i] * bmin[2] +
                          mvp[12 + i];
                for (int a = 0; a < 3; ++a)
                    edge[a][i] = mvp[4 * a + i] * (double(bmax[a]) - double(bmin[a]));
            }
            double mnx = 1e300, mny = 1e300, mxx = -1e300, mxy = -1e300;
            bool full = false;
            for (int c = 0; c < 8 && !full; ++c)
            {
                double clip[4];
                for (int i = 0; i < 4; ++i)
                    clip[i] = base[i] + ((c & 1) ? edge[0][i] : 0.0) +
                              ((c & 2) ? edge[1][i] : 0.0) + ((c & 4) ? edge[2][i] : 0.0);
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
            bool visible = true;
            int x0 = 0, y0 = 0, x1 = W - 1, y1 = H - 1;
            if (!full)
            {
                if (mxx < -2.0 || mxy < -2.0 || mnx > W + 2.0 || mny > H + 2.0)
                    visible = false; // entirely off screen
                x0 = int(std::max(0.0, std::floor(mnx) - 2));
                y0 = int(std::max(0.0, std::floor(mny) - 2));
                x1 = int(std::min(double(W - 1), std::ceil(mxx) + 2));
                y1 = int(std::min(double(H - 1), std::ceil(mxy) + 2));
            }
            if (visible)
            {
                r = make_int4(x0 / tw, y0 / th, x1 / tw, y1 / th);
                *listTotal += uint64_t(r.z - r.x + 1) * uint64_t(r.w - r.y + 1);
                *superTotal += uint64_t(r.z / kSuperTile - r.x / kSuperTile + 1) *
                               uint64_t(r.w / kSuperTile - r.y / kSuperTile + 1);
                ++*onScreen;
            }
        }
        bricks.push_back(e);
        ranges.push_back(r);
    }
    if (*listTotal > (1ull << 31))
        return fail(ctx, LVR_ENOMEM, "tile lists exceed 2^31 entries");
    return LVR_OK;
}

// Per-frame device buffers grow without waiting for the frames in flight:
// the old buffer is retired (frames queued before still read it) and freed
// at the next point where every stream has drained (lvr_ctx_sync, destroy).
// Growth at least doubles, so a camera path settles after a few frames.
template <typename T>
static int growBuffer(lvr_ctx* ctx, T** ptr, size_t* cap, size_t need)
{
    if (need <= *cap && *ptr)
        return LVR_OK;
    if (ctx->retired.size() >= 64)
    {
        int rc = lvr_ctx_sync(ctx);
        if (rc)
            return rc;
    }
    size_t n = std::max<size_t>(need, 1);
    n = std::max(n + n / 4, 2 * *cap);
    T* fresh = nullptr;
    LVR_HIP(ctx, hipMalloc((void**)&fresh, n * sizeof(T)));
    if (*ptr)
        ctx->retired.push_back(*ptr);
    *ptr = fresh;
    *cap = n;
    return LVR_OK;
}
} // namespace lvr

// The scratch of lvr_render_noalloc (about 26 MB, held by every device the
// library runs on until the process ends). In host code this name is only a
// stand-in for the symbol: kernels get the device address that
// lvr::noallocSetup takes with hipGetSymbolAddress on each device.
static __device__ __align__(256) uint8_t scratch[lvr::kNaScratchBytes];

namespace lvr
{
int noallocSetup(lvr_ctx* ctx)
{
    // the device address of `scratch`, once per device
    static std::mutex lock;
    static std::unordered_map<int, uint8_t*> perDevice;
    {
        std::lock_guard<std::mutex> g(lock);
        auto it = perDevice.find(ctx->device);
        if (it == perDevice.end())
        {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return hipFail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol,
                               "hipGetSymbolAddress(scratch)");
            it = perDevice.emplace(ctx->device, static_cast<uint8_t*>(p)).first;
        }
        ctx->d_scratch = it->second;
    }
    // host side of a no-allocation frame: its keys and ranges go into
    // vectors of fixed capacity and through a fixed pinned staging buffer
    ctx->hostBricks.reserve(size_t(kNaBricks));
    ctx->hostRanges.reserve(size_t(kNaBricks));
    LVR_HIP(ctx, hipHostMalloc((void**)&ctx->h_noalloc, kNaStagingBytes, hipHostMallocDefault));
    LVR_HIP(ctx, hipEventCreateWithFlags(&ctx->noallocCopied, hipEventDisableTiming));
    return LVR_OK;
}
} // namespace lvr

// --------------------------------------------------------------- batches
struct lvr_batch
{
    int device = 0;
    int numCUs = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::string error;
    lvr_ctx* errCtx = nullptr; // error sink for the shared helpers

    std::vector<lvr::FrameUniforms> reqs;
    bool wpsDefault = false; // the first request left waves_per_simd to the library
    std::vector<lvr::BrickKey> bricks;
    std::vector<int4> ranges;
    std::vector<float4> tf;
    std::vector<size_t> tfOffset, frameOffset;
    uint64_t listTotal = 0, superTotal = 0, onScreen = 0, pixels = 0;

    // device buffers
    lvr::FrameUniforms* d_reqs = nullptr;
    size_t reqCap = 0;
    lvr::BrickEntry* d_bricks = nullptr;
    size_t brickCap = 0;
    lvr::BrickKey* d_keys = nullptr;
    size_t keyCap = 0;
    int4* d_ranges = nullptr;
    size_t rangeCap = 0;
    float4* d_tf = nullptr;
    size_t tfCap = 0;
    uint32_t* d_frames = nullptr;
    size_t frameCap = 0;
    int32_t* d_counts = nullptr;
    int32_t* d_offsets = nullptr;
    size_t tileCap = 0;
    int32_t* d_lists = nullptr;
    size_t listCap = 0;
    int32_t* d_scounts = nullptr;
    int32_t* d_soffsets = nullptr;
    size_t superCap = 0;
    int32_t* d_slists = nullptr;
    size_t slistCap = 0;
    unsigned long long* d_counters = nullptr;
    size_t counterCap = 0;
    std::vector<unsigned long long> h_counters;
    lvr::PinnedRing staging; // host tables on their way to d_*
    bool ran = false;

    // graph of the last launch sequence and the shape it was captured for
    bool useGraph = true;
    int soloMinTiles = 0;  // lvr_batch_set_solo_tiles
    std::vector<int> solo; // requests of the last run launched on their own
    // streams beside `stream` on which solo launches run at once, so that
    // one launch's tail overlaps the next (created on first use; used while
    // no CUs are reserved), the event that starts them after the binning,
    // and the ones that join them back
    static constexpr int kSoloStreams = 3;
    hipStream_t side[kSoloStreams] = {};
    hipEvent_t sideDone[kSoloStreams] = {};
    hipEvent_t prepDone = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int64_t> graphKey;
    uint64_t graphLaunches = 0, graphCaptures = 0;

    // contexts whose pools the current requests read, and the event that
    // marks the end of the last launch
    std::vector<lvr_ctx*> ctxs;
    hipEvent_t done = nullptr;
    float kernelMs = 0.f;

    // per-request output frames in caller memory (null: the batch's own)
    std::vector<uint32_t*> outputs;
    // counters land in pinned host memory at the end of each launch
    unsigned long long* h_pinned = nullptr;
    size_t pinnedCap = 0;
    // CUs kept free of the batch kernel (stream CU mask)
    int reservedCUs = 0;
};

// The batch no longer reads these pools (its launch finished or it is going
// away): take it off their reader lists.
static void detachContexts(lvr_batch* b)
{
    for (lvr_ctx* c : b->ctxs)
        c->readers.erase(std::remove(c->readers.begin(), c->readers.end(), b), c->readers.end());
    b->ctxs.clear();
}

namespace lvr
{
hipEvent_t batchDoneEvent(lvr_batch* b)
{
    return b->done;
}
void batchForgetCtx(lvr_batch* b, lvr_ctx* ctx)
{
    b->ctxs.erase(std::remove(b->ctxs.begin(), b->ctxs.end(), ctx), b->ctxs.end());
}
hipStream_t batchStream(lvr_batch* b)
{
    return b->stream;
}
uint32_t* batchFrame(lvr_batch* b, int request)
{
    return b->d_frames + b->frameOffset[size_t(request)];
}
} // namespace lvr

namespace lvr
{
static int bfail(lvr_batch* b, int code, const std::string& msg)
{
    b->error = msg;
    return code;
}

#define LVR_BHIP(b, call)                                                                  \
    do                                                                                     \
    {                                                                                      \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess)                                                              \
            return lvr::bfail((b), LVR_EHIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

template <typename T>
static int bgrow(lvr_batch* b, T** ptr, size_t* cap, size_t need, bool* grew)
{
    if (need <= *cap && *ptr)
        return LVR_OK;
    LVR_BHIP(b, hipStreamSynchronize(b->stream));
    (void)hipFree(*ptr);
    *ptr = nullptr;
    size_t n = std::max<size_t>(need, 1);
    n = n + n / 4;
    LVR_BHIP(b, hipMalloc((void**)ptr, n * sizeof(T)));
    *cap = n;
    *grew = true;
    return LVR_OK;
}
} // namespace lvr

extern "C" {

// The classified table of the context for the frame's TF, data range and
// sampling rate, recomputed on the render stream when any of them changed.
static int ensureClassified(lvr_ctx* ctx, const lvr::FrameUniforms& U, int nValues)
{
    using namespace lvr;
    uint32_t mult, add;
    std::memcpy(&mult, &U.mult, 4);
    std::memcpy(&add, &U.add, 4);
    const uint64_t key[4] = {ctx->tfGeneration,
                             (uint64_t(uint32_t(U.nSamples)) << 32) | uint32_t(U.maxSamples),
                             (uint64_t(mult) << 32) | add,
                             (uint64_t(uint32_t(U.nTf)) << 32) | uint32_t(nValues)};
    if (ctx->d_cls && std::memcmp(key, ctx->clsKey, sizeof key) == 0)
        return LVR_OK;
    if (!ctx->d_cls)
        LVR_HIP(ctx, hipMalloc((void**)&ctx->d_cls, sizeof(float4) * 65536));
    const Classify c = makeClassifyHost(U);
    hipLaunchKernelGGL(classifyKernel, dim3((nValues + 255) / 256), dim3(256), 0, ctx->stream, c,
                       U.tf, nValues, ctx->d_cls);
    LVR_HIP(ctx, hipGetLastError());
    std::memcpy(ctx->clsKey, key, sizeof key);
    return LVR_OK;
}

// End of a frame on the render stream: ray-cast end event, the composite
// step count into the frame's pinned slot (copied before the next frame's
// binning clears the counter), and the frame's completion event.
static int finishFrame(lvr_ctx* ctx, lvr_ctx::FrameRecord& rec)
{
    const int slot = int(ctx->frameCount % lvr_ctx::kFrameSlots);
    if (rec.timed)
        LVR_HIP(ctx, hipEventRecord(rec.ev[2], ctx->stream));
    LVR_HIP(ctx, hipMemcpyAsync(&ctx->h_frameSamples[slot], ctx->curCounters, sizeof(uint64_t),
                                hipMemcpyDeviceToHost, ctx->stream));
    LVR_HIP(ctx, hipEventRecord(rec.done, ctx->stream));
    rec.armed = true;
    ++ctx->frameCount;
    return LVR_OK;
}

// A frame into the RGBA8 framebuffer (layer < 0) or, as a sort-last
// partial, into the context's RGBA32F layer `layer`.
static int renderFrame(lvr_ctx* ctx, const lvr_frustum* f, const int viewport[4],
                       const uint64_t* ordered_ids, int n, const lvr_render_params* params,
                       int layer)
{
    using namespace lvr;
    if (!ctx)
        return LVR_EINVAL;
    const auto hostStart = std::chrono::steady_clock::now();
    if (!ctx->tfSet)
        return fail(ctx, LVR_EINVAL, "no transfer function set");
    int rc = ensureDevice(ctx);
    if (rc)
        return rc;
    FrameUniforms U;
    std::vector<BrickKey>& bricks = ctx->hostBricks;
    std::vector<int4>& ranges = ctx->hostRanges;
    bricks.clear();
    ranges.clear();
    uint64_t listTotal = 0, superTotal = 0, onScreen = 0;
    rc = prepareFrame(ctx, f, viewport, ordered_ids, n, params, ctx->nTf, ctx->range0,
                      ctx->range1, &U, bricks, ranges, &listTotal, &superTotal, &onScreen);
    if (rc)
        return rc;
    const int W = U.width, H = U.height;
    const int nTiles = U.tilesX * U.tilesY;

    // device buffers
    if (layer >= 0)
    {
        rc = growBuffer(ctx, &ctx->d_layers[layer], &ctx->layerCap[layer], size_t(W) * H);
        if (rc)
            return rc;
        ctx->layerW[layer] = W;
        ctx->layerH[layer] = H;
        U.layer = ctx->d_layers[layer];
    }
    else
    {
        rc = growBuffer(ctx, &ctx->d_frame, &ctx->fbCapacity, size_t(W) * H);
        if (rc)
            return rc;
        ctx->fbW = W;
        ctx->fbH = H;
        U.frame = ctx->d_frame;
    }
    U.tf = reinterpret_cast<const float4*>(ctx->d_tf);
    lvr_ctx::RecordSet& rs = ctx->recSets[ctx->recNext];
    if (!ctx->recStream)
        LVR_HIP(ctx, hipStreamCreateWithFlags(&ctx->recStream, hipStreamNonBlocking));
    if (!rs.ready)
    {
        LVR_HIP(ctx, hipEventCreateWithFlags(&rs.ready, hipEventDisableTiming));
        LVR_HIP(ctx, hipEventCreateWithFlags(&rs.free, hipEventDisableTiming));
    }
    rc = growBuffer(ctx, &rs.bricks, &rs.brickCap, std::max<size_t>(1, size_t(n)));
    if (rc)
        return rc;
    rc = growBuffer(ctx, &rs.keys, &rs.keyCap, std::max<size_t>(1, size_t(n)));
    if (rc)
        return rc;
    rc = growBuffer(ctx, &rs.ranges, &rs.rangeCap, std::max<size_t>(1, size_t(n)));
    if (rc)
        return rc;
    rc = growBuffer(ctx, &rs.req, &rs.reqCap, 1);
    if (rc)
        return rc;
    const int nSupers = U.superX * U.superY;
    if ((rc = growBuffer(ctx, &rs.tileCount, &rs.tileCountCap, size_t(nTiles))) ||
        (rc = growBuffer(ctx, &rs.tileOffset, &rs.tileOffsetCap, size_t(nTiles))) ||
        (rc = growBuffer(ctx, &rs.tileList, &rs.tileListCap,
                         size_t(std::max<uint64_t>(listTotal, 1)))) ||
        (rc = growBuffer(ctx, &rs.superCount, &rs.superCountCap, size_t(nSupers))) ||
        (rc = growBuffer(ctx, &rs.superOffset, &rs.superOffsetCap, size_t(nSupers))) ||
        (rc = growBuffer(ctx, &rs.superList, &rs.superListCap,
                         size_t(std::max<uint64_t>(superTotal, 1)))) ||
        (rc = growBuffer(ctx, &rs.counters, &rs.countersCap, 4)) ||
        (rc = growBuffer(ctx, &rs.superOrder, &rs.superOrderCap,
                         size_t(U.nItems / (kSuper * kSuper)))))
        return rc;
    U.superOrder = rs.superOrder;
    ctx->recNext ^= 1;
    ctx->d_bricks = rs.bricks;
    ctx->d_ranges = rs.ranges;
    ctx->d_req = rs.req;
    ctx->d_tileCount = rs.tileCount;
    ctx->d_tileOffset = rs.tileOffset;
    ctx->d_tileList = rs.tileList;
    ctx->d_superCount = rs.superCount;
    ctx->d_superOffset = rs.superOffset;
    ctx->d_superList = rs.superList;
    ctx->curCounters = rs.counters;

    // uniforms, brick keys and ranges through one pinned staging buffer,
    // copied on recStream (while the previous frame still ray-casts) once
    // the ray cast that last read this record set is done
    {
        const size_t bBytes = sizeof(BrickKey) * size_t(n), rBytes = sizeof(int4) * size_t(n);
        uint8_t* h = nullptr;
        int slot = 0;
        LVR_HIP(ctx, ctx->frameStaging.acquire(sizeof(FrameUniforms) + bBytes + rBytes, &h,
                                               &slot));
        std::memcpy(h, &U, sizeof(FrameUniforms));
        std::memcpy(h + sizeof(FrameUniforms), bricks.data(), bBytes);
        std::memcpy(h + sizeof(FrameUniforms) + bBytes, ranges.data(), rBytes);
        if (rs.used)
            LVR_HIP(ctx, hipStreamWaitEvent(ctx->recStream, rs.free, 0));
        LVR_HIP(ctx, hipMemcpyAsync(ctx->d_req, h, sizeof(FrameUniforms),
                                    hipMemcpyHostToDevice, ctx->recStream));
        if (n > 0)
        {
            LVR_HIP(ctx, hipMemcpyAsync(rs.keys, h + sizeof(FrameUniforms), bBytes,
                                        hipMemcpyHostToDevice, ctx->recStream));
            LVR_HIP(ctx, hipMemcpyAsync(ctx->d_ranges, h + sizeof(FrameUniforms) + bBytes,
                                        rBytes, hipMemcpyHostToDevice, ctx->recStream));
        }
        LVR_HIP(ctx, ctx->frameStaging.release(slot, ctx->recStream));
    }
    // this frame's statistics slot; its previous frame (kFrameSlots ago) has
    // long completed, but its events are only reused once it has
    lvr_ctx::FrameRecord& rec = ctx->frames[ctx->frameCount % lvr_ctx::kFrameSlots];
    if (rec.armed)
        LVR_HIP(ctx, hipEventSynchronize(rec.done));
    rec.armed = false;
    rec.bricks = onScreen;
    rec.entries = listTotal;
    rec.timed = ctx->timing;
    rec.histTimed = false;
    if (rec.timed)
        LVR_HIP(ctx, hipEventRecord(rec.ev[0], ctx->recStream));
    // page-table changes since the last frame, then the records that read it
    rc = applyPageTable(ctx, ctx->recStream);
    if (rc)
        return rc;
    enqueueRecords(ctx->recStream, ctx->d_req, 1, rs.keys, n, ctx->d_bricks);
    enqueueBinning(ctx->recStream, ctx->d_req, 1, ctx->d_ranges, nSupers, nTiles,
                   ctx->d_superCount, ctx->d_superOffset, ctx->d_superList, ctx->d_tileCount,
                   ctx->d_tileOffset, ctx->d_tileList, ctx->curCounters, 2);
    hipLaunchKernelGGL(superOrderKernel, dim3(1), dim3(1024), 0, ctx->recStream,
                       ctx->d_tileCount, U.tilesX, U.tilesY, rs.superOrder);
    LVR_HIP(ctx, hipGetLastError());
    if (rec.timed)
        LVR_HIP(ctx, hipEventRecord(rec.ev[5], ctx->recStream));
    LVR_HIP(ctx, hipEventRecord(rs.ready, ctx->recStream));
    LVR_HIP(ctx, hipStreamWaitEvent(ctx->stream, rs.ready, 0));
    if (rec.timed)
        LVR_HIP(ctx, hipEventRecord(rec.ev[1], ctx->stream));
    // persistent: two workgroups per CU are resident (the LDS budget of a
    // 4096-entry transfer function); more would only find the queue empty
    const int perBlock = U.wavesPerSimd * 2;
    const int blocks =
        std::max(1, std::min((U.nItems + perBlock - 1) / perBlock, 2 * ctx->numCUs));
    size_t lds = rayLds(U);
    const int dt = ctx->vi.data_type;
    if (!U.linear && U.classify == 1 && (dt == LVR_DT_UINT8 || dt == LVR_DT_UINT16))
    {
        // Integer volumes sampled NEAREST: classification depends on the raw
        // value only, so it is tabulated once per (TF, range, rate) and the
        // march reads one table entry per sample.
        const int nValues = dt == LVR_DT_UINT8 ? 256 : 65536;
        rc = ensureClassified(ctx, U, nValues);
        if (rc)
            return rc;
        U.cls = ctx->d_cls;
        lds = (dt == LVR_DT_UINT8 ? 256 * sizeof(float4) : 0) + offsetTablesLds(U);
        if (dt == LVR_DT_UINT8)
            launchNearest<uint8_t, true>(blocks, lds, ctx->stream, U, ctx->d_bricks,
                                         ctx->d_tileOffset, ctx->d_tileCount, ctx->d_tileList,
                                         ctx->curCounters);
        else
            launchNearest<uint16_t, true>(blocks, lds, ctx->stream, U, ctx->d_bricks,
                                          ctx->d_tileOffset, ctx->d_tileCount, ctx->d_tileList,
                                          ctx->curCounters);
        LVR_HIP(ctx, hipGetLastError());
        rec.hostMs = std::chrono::duration<float, std::milli>(
                         std::chrono::steady_clock::now() - hostStart).count();
        if ((rc = finishFrame(ctx, rec)))
            return rc;
        LVR_HIP(ctx, hipEventRecord(rs.free, ctx->stream));
        rs.used = true;
        return LVR_OK;
    }
    switch (dt)
    {
    case LVR_DT_UINT8:
        launchRaycast<uint8_t>(blocks, lds, ctx->stream, U, ctx->d_bricks, ctx->d_tileOffset,
                               ctx->d_tileCount, ctx->d_tileList, ctx->curCounters);
        break;
    case LVR_DT_UINT16:
        launchRaycast<uint16_t>(blocks, lds, ctx->stream, U, ctx->d_bricks, ctx->d_tileOffset,
                                ctx->d_tileCount, ctx->d_tileList, ctx->curCounters);
        break;
    case LVR_DT_UINT32:
        launchRaycast<uint32_t>(blocks, lds, ctx->stream, U, ctx->d_bricks, ctx->d_tileOffset,
                                ctx->d_tileCount, ctx->d_tileList, ctx->curCounters);
        break;
    case LVR_DT_INT8:
        launchRaycast<int8_t>(blocks, lds, ctx->stream, U, ctx->d_bricks, ctx->d_tileOffset,
                              ctx->d_tileCount, ctx->d_tileList, ctx->curCounters);
        break;
    case LVR_DT_INT16:
        launchRaycast<int16_t>(blocks, lds, ctx->stream, U, ctx->d_bricks, ctx->d_tileOffset,
                               ctx->d_tileCount, ctx->d_tileList, ctx->curCounters);
        break;
    case LVR_DT_INT32:
        launchRaycast<int32_t>(blocks, lds, ctx->stream, U, ctx->d_bricks, ctx->d_tileOffset,
                               ctx->d_tileCount, ctx->d_tileList, ctx->curCounters);
        break;
    default:
        launchRaycast<float>(blocks, lds, ctx->stream, U, ctx->d_bricks, ctx->d_tileOffset,
                             ctx->d_tileCount, ctx->d_tileList, ctx->curCounters);
        break;
    }
    LVR_HIP(ctx, hipGetLastError());
    rec.hostMs = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() -
                                                          hostStart)
                     .count();
    if ((rc = finishFrame(ctx, rec)))
        return rc;
    LVR_HIP(ctx, hipEventRecord(rs.free, ctx->stream));
    rs.used = true;
    return LVR_OK;
}

int lvr_render(lvr_ctx* ctx, const lvr_frustum* f, const int viewport[4],
               const uint64_t* ordered_ids, int n, const lvr_render_params* params)
{
    return renderFrame(ctx, f, viewport, ordered_ids, n, params, -1);
}

int lvr_render_noalloc(lvr_ctx* ctx, const lvr_frustum* f, const int viewport[4],
                       const uint64_t* ordered_ids, int n, const lvr_render_params* params,
                       void* out_rgba8, uint64_t* out_samples)
{
    using namespace lvr;
    if (!ctx)
        return LVR_EINVAL;
    if (!out_rgba8 || reinterpret_cast<uintptr_t>(out_rgba8) % 4u ||
        reinterpret_cast<uintptr_t>(out_samples) % 8u)
        return fail(ctx, LVR_EINVAL, "output must be 4-byte (frame) and 8-byte (samples) aligned");
    if (!ctx->tfSet)
        return fail(ctx, LVR_EINVAL, "no transfer function set");
    if (!ctx->d_scratch || !ctx->h_noalloc)
        return fail(ctx, LVR_EHIP, "no-allocation scratch unavailable on this device");
    if (n > kNaBricks)
        return fail(ctx, LVR_ERANGE, "no-allocation render: more than 32768 bricks");
    int rc = ensureDevice(ctx);
    if (rc)
        return rc;
    FrameUniforms U;
    std::vector<BrickKey>& keys = ctx->hostBricks; // capacity reserved at set-up
    std::vector<int4>& ranges = ctx->hostRanges;
    keys.clear();
    ranges.clear();
    uint64_t listTotal = 0, superTotal = 0, onScreen = 0;
    rc = prepareFrame(ctx, f, viewport, ordered_ids, n, params, ctx->nTf, ctx->range0,
                      ctx->range1, &U, keys, ranges, &listTotal, &superTotal, &onScreen);
    if (rc)
        return rc;
    const int nTiles = U.tilesX * U.tilesY, nSupers = U.superX * U.superY;
    if (nTiles > kNaTiles || nSupers > kNaSupers)
        return fail(ctx, LVR_ERANGE, "no-allocation render: viewport above 262144 ray tiles");
    if (listTotal > kNaTileEntries || superTotal > kNaSuperEntries)
        return fail(ctx, LVR_ERANGE,
                    "no-allocation render: more than 4M (tile, brick) pairs to bin");
    uint8_t* d = ctx->d_scratch;
    FrameUniforms* dU = reinterpret_cast<FrameUniforms*>(d + kNaOffU);
    unsigned long long* dCounters = reinterpret_cast<unsigned long long*>(d + kNaOffCounters);
    BrickKey* dKeys = reinterpret_cast<BrickKey*>(d + kNaOffKeys);
    BrickEntry* dBricks = reinterpret_cast<BrickEntry*>(d + kNaOffBricks);
    int4* dRanges = reinterpret_cast<int4*>(d + kNaOffRanges);
    int32_t* dTileCount = reinterpret_cast<int32_t*>(d + kNaOffTileCount);
    int32_t* dTileOffset = reinterpret_cast<int32_t*>(d + kNaOffTileOffset);
    int32_t* dSuperCount = reinterpret_cast<int32_t*>(d + kNaOffSuperCount);
    int32_t* dSuperOffset = reinterpret_cast<int32_t*>(d + kNaOffSuperOffset);
    int32_t* dTileList = reinterpret_cast<int32_t*>(d + kNaOffTileList);
    int32_t* dSuperList = reinterpret_cast<int32_t*>(d + kNaOffSuperList);
    float4* dCls = reinterpret_cast<float4*>(d + kNaOffCls);

    U.frame = static_cast<uint32_t*>(out_rgba8);
    U.layer = nullptr;
    U.tf = reinterpret_cast<const float4*>(ctx->d_tf);
    const int dt = ctx->vi.data_type;
    const bool table = !U.linear && U.classify == 1 && (dt == LVR_DT_UINT8 || dt == LVR_DT_UINT16);
    if (table)
        U.cls = dCls;

    // uniforms, keys and ranges through the fixed pinned staging, once the
    // previous call's copies have read it; every stage runs on the render
    // stream in order
    if (ctx->noallocArmed)
        LVR_HIP(ctx, hipEventSynchronize(ctx->noallocCopied));
    ctx->noallocArmed = false;
    const size_t kB = sizeof(BrickKey) * size_t(n), rB = sizeof(int4) * size_t(n);
    uint8_t* h = ctx->h_noalloc;
    std::memcpy(h, &U, sizeof(FrameUniforms));
    std::memcpy(h + sizeof(FrameUniforms), keys.data(), kB);
    std::memcpy(h + sizeof(FrameUniforms) + kB, ranges.data(), rB);
    LVR_HIP(ctx, hipMemcpyAsync(dU, h, sizeof(FrameUniforms), hipMemcpyHostToDevice,
                                ctx->stream));
    if (n > 0)
    {
        LVR_HIP(ctx, hipMemcpyAsync(dKeys, h + sizeof(FrameUniforms), kB,
                                    hipMemcpyHostToDevice, ctx->stream));
        LVR_HIP(ctx, hipMemcpyAsync(dRanges, h + sizeof(FrameUniforms) + kB, rB,
                                    hipMemcpyHostToDevice, ctx->stream));
    }
    LVR_HIP(ctx, hipEventRecord(ctx->noallocCopied, ctx->stream));
    ctx->noallocArmed = true;
    rc = applyPageTable(ctx, ctx->stream);
    if (rc)
        return rc;
    enqueueRecords(ctx->stream, dU, 1, dKeys, n, dBricks);
    // the second scan clears the samples and work-queue counters
    enqueueBinning(ctx->stream, dU, 1, dRanges, nSupers, nTiles, dSuperCount, dSuperOffset,
                   dSuperList, dTileCount, dTileOffset, dTileList, dCounters, 2);
    LVR_HIP(ctx, hipGetLastError());
    const int perBlock = U.wavesPerSimd * 2;
    const int blocks =
        std::max(1, std::min((U.nItems + perBlock - 1) / perBlock, 2 * ctx->numCUs));
    size_t lds = rayLds(U);
    if (table)
    {
        const int nValues = dt == LVR_DT_UINT8 ? 256 : 65536;
        Classify c = makeClassifyHost(U);
        hipLaunchKernelGGL(classifyKernel, dim3((nValues + 255) / 256), dim3(256), 0,
                           ctx->stream, c, U.tf, nValues, dCls);
        lds = (dt == LVR_DT_UINT8 ? 256 * sizeof(float4) : 0) + offsetTablesLds(U);
        if (dt == LVR_DT_UINT8)
            launchNearest<uint8_t, true>(blocks, lds, ctx->stream, U, dBricks, dTileOffset,
                                         dTileCount, dTileList, dCounters);
        else
            launchNearest<uint16_t, true>(blocks, lds, ctx->stream, U, dBricks, dTileOffset,
                                          dTileCount, dTileList, dCounters);
    }
    else
        switch (dt)
        {
        case LVR_DT_UINT8:
            launchRaycast<uint8_t>(blocks, lds, ctx->stream, U, dBricks, dTileOffset,
                                   dTileCount, dTileList, dCounters);
            break;
        case LVR_DT_UINT16:
            launchRaycast<uint16_t>(blocks, lds, ctx->stream, U, dBricks, dTileOffset,
                                    dTileCount, dTileList, dCounters);
            break;
        case LVR_DT_UINT32:
            launchRaycast<uint32_t>(blocks, lds, ctx->stream, U, dBricks, dTileOffset,
                                    dTileCount, dTileList, dCounters);
            break;
        case LVR_DT_INT8:
            launchRaycast<int8_t>(blocks, lds, ctx->stream, U, dBricks, dTileOffset, dTileCount,
                                  dTileList, dCounters);
            break;
        case LVR_DT_INT16:
            launchRaycast<int16_t>(blocks, lds, ctx->stream, U, dBricks, dTileOffset,
                                   dTileCount, dTileList, dCounters);
            break;
        case LVR_DT_INT32:
            launchRaycast<int32_t>(blocks, lds, ctx->stream, U, dBricks, dTileOffset,
                                   dTileCount, dTileList, dCounters);
            break;
        default:
            launchRaycast<float>(blocks, lds, ctx->stream, U, dBricks, dTileOffset, dTileCount,
                                 dTileList, dCounters);
            break;
        }
    LVR_HIP(ctx, hipGetLastError());
    if (out_samples)
        LVR_HIP(ctx, hipMemcpyAsync(out_samples, dCounters, sizeof(uint64_t),
                                    hipMemcpyDeviceToDevice, ctx->stream));
    return LVR_OK;
}

int lvr_render_partial(lvr_ctx* ctx, const lvr_frustum* f, const int viewport[4],
                       const uint64_t* ordered_ids, int n, const lvr_render_params* params,
                       int layer)
{
    if (!ctx || layer < 0 || layer >= lvr_ctx::kMaxLayers)
        return LVR_EINVAL;
    return renderFrame(ctx, f, viewport, ordered_ids, n, params, layer);
}

int lvr_partial_device_ptr(lvr_ctx* ctx, int layer, void** ptr)
{
    if (!ctx || !ptr || layer < 0 || layer >= lvr_ctx::kMaxLayers)
        return LVR_EINVAL;
    *ptr = ctx->d_layers[layer];
    return LVR_OK;
}

int lvr_partial_read(lvr_ctx* ctx, int layer, float* out_rgba32f, size_t bytes, int* width,
                     int* height)
{
    if (!ctx || layer < 0 || layer >= lvr_ctx::kMaxLayers)
        return LVR_EINVAL;
    if (!ctx->d_layers[layer] || ctx->layerW[layer] <= 0)
        return lvr::fail(ctx, LVR_EINVAL, "layer never rendered");
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    const int W = ctx->layerW[layer], H = ctx->layerH[layer];
    if (width)
        *width = W;
    if (height)
        *height = H;
    if (!out_rgba32f)
        return LVR_OK;
    const size_t need = size_t(W) * size_t(H) * sizeof(float4);
    if (bytes < need)
        return lvr::fail(ctx, LVR_ERANGE, "layer buffer too small");
    LVR_HIP(ctx, hipMemcpyAsync(out_rgba32f, ctx->d_layers[layer], need, hipMemcpyDeviceToHost,
                                ctx->stream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LVR_OK;
}

int lvr_composite_layers(lvr_ctx* ctx, const void* const* layers, int n, int width, int height,
                         int reps, float* kernel_ms)
{
    using namespace lvr;
    if (!ctx || !layers || n < 1 || n > kMaxCompositeLayers || width < 1 || height < 1 ||
        width > 16384 || height > 16384 || reps < 1 || reps > 1000)
        return LVR_EINVAL;
    int rc = ensureDevice(ctx);
    if (rc)
        return rc;
    LayerSet L = {};
    for (int i = 0; i < n; ++i)
    {
        if (!layers[i])
            return fail(ctx, LVR_EINVAL, "composite: null layer");
        // the context's own layers are checked against the viewport they were
        // rendered at; a buffer a layer has outgrown is no longer valid
        for (int k = 0; k < lvr_ctx::kMaxLayers; ++k)
            if (layers[i] == ctx->d_layers[k] &&
                (ctx->layerW[k] != width || ctx->layerH[k] != height))
                return fail(ctx, LVR_EINVAL,
                            "composite: layer " + std::to_string(k) + " holds " +
                                std::to_string(ctx->layerW[k]) + " x " +
                                std::to_string(ctx->layerH[k]) + " pixels");
        for (void* p : ctx->retired)
            if (layers[i] == p)
                return fail(ctx, LVR_EINVAL, "composite: layer buffer was replaced");
        L.p[i] = static_cast<const float4*>(layers[i]);
    }
    rc = growBuffer(ctx, &ctx->d_frame, &ctx->fbCapacity, size_t(width) * height);
    if (rc)
        return rc;
    ctx->fbW = width;
    ctx->fbH = height;
    const int npix = width * height;
    const int grid = std::min((npix + 255) / 256, 16 * ctx->numCUs);
    if (kernel_ms)
    {
        if (!ctx->compositeEv[0])
        {
            LVR_HIP(ctx, hipEventCreate(&ctx->compositeEv[0]));
            LVR_HIP(ctx, hipEventCreate(&ctx->compositeEv[1]));
        }
        LVR_HIP(ctx, hipEventRecord(ctx->compositeEv[0], ctx->stream));
    }
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(compositeKernel, dim3(grid), dim3(256), 0, ctx->stream, L, n, npix,
                           ctx->d_frame);
    LVR_HIP(ctx, hipGetLastError());
    if (kernel_ms)
    {
        LVR_HIP(ctx, hipEventRecord(ctx->compositeEv[1], ctx->stream));
        LVR_HIP(ctx, hipEventSynchronize(ctx->compositeEv[1]));
        LVR_HIP(ctx, hipEventElapsedTime(kernel_ms, ctx->compositeEv[0], ctx->compositeEv[1]));
        *kernel_ms /= float(reps);
    }
    return LVR_OK;
}

int lvr_frame_read(lvr_ctx* ctx, uint8_t* out_rgba8, size_t bytes, uint64_t* samples)
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
        LVR_HIP(ctx, hipMemcpyAsync(out_rgba8, ctx->d_frame, need, hipMemcpyDeviceToHost,
                                    ctx->stream));
    }
    LVR_HIP(ctx, hipMemcpyAsync(ctx->h_counters,
                                ctx->curCounters ? ctx->curCounters : ctx->d_counters,
                                sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, ctx->stream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (samples)
        *samples = ctx->h_counters[0];
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

int lvr_set_timing(lvr_ctx* ctx, int enable)
{
    if (!ctx)
        return LVR_EINVAL;
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    if (enable && !ctx->frames[0].ev[0])
        for (auto& f : ctx->frames)
            for (auto& e : f.ev)
                LVR_HIP(ctx, hipEventCreate(&e));
    ctx->timing = enable != 0;
    return LVR_OK;
}

int lvr_frame_stats_back(lvr_ctx* ctx, int back, lvr_frame_stats* out)
{
    if (!ctx || !out || back < 0)
        return LVR_EINVAL;
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    std::memset(out, 0, sizeof *out);
    out->brick_bytes = ctx->brickBytes;
    if (ctx->frameCount == 0 && back == 0)
        return LVR_OK; // nothing rendered yet
    if (uint64_t(back) >= std::min<uint64_t>(ctx->frameCount, lvr_ctx::kFrameSlots))
        return lvr::fail(ctx, LVR_ERANGE, "frame statistics no longer kept");
    const uint64_t frame = ctx->frameCount - 1 - uint64_t(back);
    const int slot = int(frame % lvr_ctx::kFrameSlots);
    const lvr_ctx::FrameRecord& rec = ctx->frames[slot];
    LVR_HIP(ctx, hipEventSynchronize(rec.done));
    out->samples = ctx->h_frameSamples[slot];
    out->bricks = rec.bricks;
    out->tile_entries = rec.entries;
    out->host_ms = rec.hostMs;
    if (rec.timed)
    {
        LVR_HIP(ctx, hipEventElapsedTime(&out->bin_ms, rec.ev[0], rec.ev[5]));
        LVR_HIP(ctx, hipEventElapsedTime(&out->raycast_ms, rec.ev[1], rec.ev[2]));
        if (rec.histTimed)
        {
            LVR_HIP(ctx, hipEventSynchronize(rec.ev[4]));
            LVR_HIP(ctx, hipEventElapsedTime(&out->histogram_ms, rec.ev[3], rec.ev[4]));
        }
    }
    return LVR_OK;
}

int lvr_frame_stats_get(lvr_ctx* ctx, lvr_frame_stats* out)
{
    return lvr_frame_stats_back(ctx, 0, out);
}

// ---------------------------------------------------------------- batch
int lvr_batch_create(int device, lvr_batch** out)
{
    if (!out)
        return LVR_EINVAL;
    *out = nullptr;
    lvr_batch* b = new lvr_batch();
    b->device = device;
    if (hipSetDevice(device) != hipSuccess ||
        hipDeviceGetAttribute(&b->numCUs, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev[0]) != hipSuccess || hipEventCreate(&b->ev[1]) != hipSuccess ||
        hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess)
    {
        lvr_batch_destroy(b);
        return LVR_EHIP;
    }
    *out = b;
    return LVR_OK;
}

void lvr_batch_destroy(lvr_batch* b)
{
    if (!b)
        return;
    (void)hipSetDevice(b->device);
    if (b->stream)
        (void)hipStreamSynchronize(b->stream);
    detachContexts(b);
    if (b->done)
        (void)hipEventDestroy(b->done);
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    (void)hipFree(b->d_reqs);
    (void)hipFree(b->d_bricks);
    (void)hipFree(b->d_keys);
    (void)hipFree(b->d_ranges);
    (void)hipFree(b->d_tf);
    (void)hipFree(b->d_frames);
    (void)hipFree(b->d_counts);
    (void)hipFree(b->d_offsets);
    (void)hipFree(b->d_lists);
    (void)hipFree(b->d_scounts);
    (void)hipFree(b->d_soffsets);
    (void)hipFree(b->d_slists);
    (void)hipFree(b->d_counters);
    if (b->h_pinned)
        (void)hipHostFree(b->h_pinned);
    b->staging.destroy();
    for (auto& e : b->ev)
        if (e)
            (void)hipEventDestroy(e);
    for (int i = 0; i < lvr_batch::kSoloStreams; ++i)
    {
        if (b->side[i])
            (void)hipStreamDestroy(b->side[i]);
        if (b->sideDone[i])
            (void)hipEventDestroy(b->sideDone[i]);
    }
    if (b->prepDone)
        (void)hipEventDestroy(b->prepDone);
    if (b->stream)
        (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* lvr_batch_last_error(const lvr_batch* b)
{
    return b ? b->error.c_str() : "null batch";
}

int lvr_batch_set_graph(lvr_batch* b, int enable)
{
    if (!b)
        return LVR_EINVAL;
    b->useGraph = enable != 0;
    return LVR_OK;
}

int lvr_batch_set_solo_tiles(lvr_batch* b, int min_tiles)
{
    if (!b)
        return LVR_EINVAL;
    b->soloMinTiles = min_tiles;
    return LVR_OK;
}

int lvr_batch_clear(lvr_batch* b)
{
    if (!b)
        return LVR_EINVAL;
    LVR_BHIP(b, hipStreamSynchronize(b->stream));
    detachContexts(b);
    b->reqs.clear();
    b->bricks.clear();
    b->ranges.clear();
    b->tf.clear();
    b->tfOffset.clear();
    b->frameOffset.clear();
    b->outputs.clear();
    b->listTotal = b->superTotal = b->onScreen = b->pixels = 0;
    b->ran = false;
    return LVR_OK;
}

int lvr_batch_add(lvr_batch* b, lvr_ctx* ctx, const lvr_frustum* f, const int viewport[4],
                  const uint64_t* ordered_ids, int n, const lvr_render_params* params,
                  const float* tf_rgba, int n_tf, float range0, float range1, int* request)
{
    using namespace lvr;
    if (!b || !ctx || !tf_rgba || n_tf < 1 || n_tf > 8192)
        return LVR_EINVAL;
    if (ctx->device != b->device)
        return bfail(b, LVR_EINVAL, "context on another device");
    const int dt = ctx->vi.data_type;
    if (dt != LVR_DT_UINT8 && dt != LVR_DT_UINT16 && dt != LVR_DT_FLOAT)
        return bfail(b, LVR_EINVAL, "batches support u8, u16 and f32 volumes");
    if (params && params->linear_filter)
        return bfail(b, LVR_EINVAL, "batches render with nearest filtering");
    if (b->ran)
    {
        const int rc = lvr_batch_clear(b);
        if (rc)
            return rc;
    }
    FrameUniforms U;
    uint64_t listTotal = 0, superTotal = 0, onScreen = 0;
    const size_t nb = b->bricks.size();
    const int rc = prepareFrame(ctx, f, viewport, ordered_ids, n, params, n_tf, range0, range1,
                                &U, b->bricks, b->ranges, &listTotal, &superTotal, &onScreen);
    if (rc)
    {
        b->error = ctx->error;
        b->bricks.resize(nb);
        b->ranges.resize(nb);
        return rc;
    }
    for (const FrameUniforms& o : b->reqs)
        if (batchTypeClass(o.dataType) == batchTypeClass(U.dataType) &&
            batchDepth(o.depth) != batchDepth(U.depth))
        {
            b->bricks.resize(nb);
            b->ranges.resize(nb);
            return bfail(b, LVR_EINVAL,
                         "samples per thread differs from the batch's for this voxel type");
        }
    // the batch kernel runs four wavefronts per SIMD unless told otherwise,
    // whatever single-frame variant a request would take alone
    const bool wpsDefault = params->waves_per_simd == 0;
    if (wpsDefault)
        U.wavesPerSimd = 4;
    if (!b->reqs.empty() && U.wavesPerSimd != b->reqs.front().wavesPerSimd)
    {
        b->bricks.resize(nb);
        b->ranges.resize(nb);
        return bfail(b, LVR_EINVAL, "waves per SIMD differs from the batch's");
    }
    U.request = int32_t(b->reqs.size());
    if (std::find(b->ctxs.begin(), b->ctxs.end(), ctx) == b->ctxs.end())
        b->ctxs.push_back(ctx);
    // the table with a copy of its first entry before it and of its last
    // after it (the layout the classification reads)
    b->tfOffset.push_back(b->tf.size());
    for (int j = -1; j <= n_tf; ++j)
    {
        const int i = std::min(std::max(j, 0), n_tf - 1);
        b->tf.push_back(make_float4(tf_rgba[4 * i], tf_rgba[4 * i + 1], tf_rgba[4 * i + 2],
                                    tf_rgba[4 * i + 3]));
    }
    b->frameOffset.push_back(b->pixels);
    b->outputs.push_back(nullptr);
    b->pixels += size_t(U.width) * U.height;
    b->listTotal += listTotal;
    b->superTotal += superTotal;
    b->onScreen += onScreen;
    if (b->reqs.empty())
        b->wpsDefault = wpsDefault;
    b->reqs.push_back(U);
    if (request)
        *request = U.request;
    return LVR_OK;
}

int lvr_batch_run(lvr_batch* b)
{
    using namespace lvr;
    if (!b || b->reqs.empty())
        return LVR_EINVAL;
    LVR_BHIP(b, hipSetDevice(b->device));
    LVR_BHIP(b, hipStreamSynchronize(b->stream)); // host arrays are reused
    const int nReq = int(b->reqs.size());
    // Requests the single-frame kernel marches on its own: those its
    // two-brick window covers (NEAREST, two fetches in flight, classification
    // per sample) when they are alone or large enough to fill the GPU for
    // most of their launch; the batch kernel takes the others.
    const int slots = std::max(1, 2 * (b->numCUs - b->reservedCUs));
    const int wps = b->reqs.front().wavesPerSimd;
    const int soloWps = b->wpsDefault ? 5 : wps;
    const int64_t soloMin = b->soloMinTiles > 0
                                ? int64_t(b->soloMinTiles)
                                : (b->soloMinTiles < 0 ? INT64_MAX
                                                       : 4 * int64_t(slots) * 2 * soloWps);
    b->solo.clear();
    std::vector<char> isSolo(size_t(nReq), 0);
    for (int r = 0; r < nReq; ++r)
    {
        const FrameUniforms& U = b->reqs[size_t(r)];
        const bool window = !U.linear && U.depth == 2 &&
                            !(U.classify == 1 && (U.dataType == LVR_DT_UINT8 ||
                                                  U.dataType == LVR_DT_UINT16));
        if (window && (nReq == 1 || int64_t(U.tilesX) * U.tilesY >= soloMin))
        {
            isSolo[size_t(r)] = 1;
            b->solo.push_back(r);
        }
    }
    const int nSolo = int(b->solo.size());
    // solo launches on `stream` and up to kSoloStreams more, the largest
    // first, each to the stream with the fewest tiles so far
    const int nSide = b->reservedCUs == 0
                          ? std::min(lvr_batch::kSoloStreams, std::max(0, nSolo - 1))
                          : 0;
    for (int i = 0; i < nSide; ++i)
    {
        if (!b->side[i])
            LVR_BHIP(b, hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking));
        if (!b->sideDone[i])
            LVR_BHIP(b, hipEventCreateWithFlags(&b->sideDone[i], hipEventDisableTiming));
    }
    if (nSide > 0 && !b->prepDone)
        LVR_BHIP(b, hipEventCreateWithFlags(&b->prepDone, hipEventDisableTiming));
    std::vector<int> soloOrder(static_cast<size_t>(nSolo)), soloLane(static_cast<size_t>(nSolo), 0);
    for (int j = 0; j < nSolo; ++j)
        soloOrder[size_t(j)] = j;
    auto tilesOf = [&](int j) {
        const FrameUniforms& U = b->reqs[size_t(b->solo[size_t(j)])];
        return int64_t(U.tilesX) * U.tilesY;
    };
    std::stable_sort(soloOrder.begin(), soloOrder.end(),
                     [&](int x, int y) { return tilesOf(x) > tilesOf(y); });
    {
        std::vector<int64_t> load(size_t(nSide + 1), 0);
        for (int j : soloOrder)
        {
            const int lane = int(std::min_element(load.begin(), load.end()) - load.begin());
            soloLane[size_t(j)] = lane;
            load[size_t(lane)] += tilesOf(j);
        }
    }
    // counters: total and queue of the batch kernel, one per request, then
    // samples and queue of each solo launch
    const int nCounters = nReq + 2 + 2 * nSolo;
    bool grew = false;
    int rc = LVR_OK;
    if ((rc = bgrow(b, &b->d_reqs, &b->reqCap, size_t(nReq), &grew)) ||
        (rc = bgrow(b, &b->d_bricks, &b->brickCap, std::max<size_t>(1, b->bricks.size()), &grew)) ||
        (rc = bgrow(b, &b->d_keys, &b->keyCap, std::max<size_t>(1, b->bricks.size()), &grew)) ||
        (rc = bgrow(b, &b->d_ranges, &b->rangeCap, std::max<size_t>(1, b->ranges.size()), &grew)) ||
        (rc = bgrow(b, &b->d_tf, &b->tfCap, b->tf.size(), &grew)) ||
        (rc = bgrow(b, &b->d_frames, &b->frameCap, b->pixels, &grew)) ||
        (rc = bgrow(b, &b->d_lists, &b->listCap, std::max<uint64_t>(1, b->listTotal), &grew)) ||
        (rc = bgrow(b, &b->d_slists, &b->slistCap, std::max<uint64_t>(1, b->superTotal),
                    &grew)) ||
        (rc = bgrow(b, &b->d_counters, &b->counterCap, size_t(nCounters), &grew)))
        return rc;
    if (b->pinnedCap < size_t(nReq) + 2)
    {
        if (b->h_pinned)
            (void)hipHostFree(b->h_pinned);
        b->h_pinned = nullptr;
        b->pinnedCap = 0;
        LVR_BHIP(b, hipHostMalloc((void**)&b->h_pinned,
                                  sizeof(unsigned long long) * (size_t(nReq) + 2),
                                  hipHostMallocDefault));
        b->pinnedCap = size_t(nReq) + 2;
    }
    // request table: positions in the batch-wide arrays. A solo request's
    // work items are not in the batch kernel's queue: its item range is
    // empty, and the search by item (the last request whose range starts at
    // or before the item) passes over it.
    int tiles = 0, items = 0, supers = 0;
    for (int r = 0; r < nReq; ++r)
    {
        FrameUniforms& U = b->reqs[size_t(r)];
        U.tileBase = tiles;
        U.itemBase = items;
        U.superBase = supers;
        tiles += U.tilesX * U.tilesY;
        supers += U.superX * U.superY;
        if (!isSolo[size_t(r)])
            items += U.nItems;
        U.tf = b->d_tf + b->tfOffset[size_t(r)];
        U.frame = b->outputs[size_t(r)] ? b->outputs[size_t(r)]
                                        : b->d_frames + b->frameOffset[size_t(r)];
    }
    bool superGrew = false;
    if ((rc = bgrow(b, &b->d_counts, &b->tileCap, size_t(tiles), &grew)) ||
        (rc = bgrow(b, &b->d_scounts, &b->superCap, size_t(supers), &superGrew)))
        return rc;
    if (superGrew || !b->d_soffsets)
    {
        (void)hipFree(b->d_soffsets);
        b->d_soffsets = nullptr;
        LVR_BHIP(b, hipMalloc((void**)&b->d_soffsets, sizeof(int32_t) * b->superCap));
        grew = true;
    }
    if (b->tileCap > 0 && !b->d_offsets)
        grew = true;
    if (grew || !b->d_offsets)
    {
        (void)hipFree(b->d_offsets);
        b->d_offsets = nullptr;
        LVR_BHIP(b, hipMalloc((void**)&b->d_offsets, sizeof(int32_t) * b->tileCap));
    }

    // the launch sequence depends on these sizes and buffers only
    int depthOf[3] = {2, 2, 2};
    for (const FrameUniforms& r : b->reqs)
        depthOf[batchTypeClass(r.dataType)] = batchDepth(r.depth);
    const int depthCode = (depthOf[0] == 3 ? 4 : 0) + (depthOf[1] == 3 ? 2 : 0) +
                          (depthOf[2] == 3 ? 1 : 0);
    std::vector<int64_t> key = {nReq,
                                depthCode,
                                wps,
                                int64_t(b->bricks.size()),
                                int64_t(b->tf.size()),
                                tiles,
                                supers,
                                items,
                                int64_t(reinterpret_cast<intptr_t>(b->d_reqs)),
                                int64_t(reinterpret_cast<intptr_t>(b->d_bricks)),
                                int64_t(reinterpret_cast<intptr_t>(b->d_keys)),
                                int64_t(reinterpret_cast<intptr_t>(b->d_ranges)),
                                int64_t(reinterpret_cast<intptr_t>(b->d_tf)),
                                int64_t(reinterpret_cast<intptr_t>(b->d_counts)),
                                int64_t(reinterpret_cast<intptr_t>(b->d_offsets)),
                                int64_t(reinterpret_cast<intptr_t>(b->d_lists)),
                                int64_t(reinterpret_cast<intptr_t>(b->d_scounts)),
                                int64_t(reinterpret_cast<intptr_t>(b->d_soffsets)),
                                int64_t(reinterpret_cast<intptr_t>(b->d_slists)),
                                int64_t(reinterpret_cast<intptr_t>(b->d_counters))};

    // inputs: the host tables go through one pinned staging buffer and are
    // copied before the launch sequence
    {
        const size_t qB = sizeof(FrameUniforms) * size_t(nReq);
        const size_t bB = sizeof(BrickKey) * b->bricks.size();
        const size_t rB = sizeof(int4) * b->ranges.size();
        const size_t tB = sizeof(float4) * b->tf.size();
        uint8_t* h = nullptr;
        int slot = 0;
        LVR_BHIP(b, b->staging.acquire(qB + bB + rB + tB, &h, &slot));
        std::memcpy(h, b->reqs.data(), qB);
        std::memcpy(h + qB, b->bricks.data(), bB);
        std::memcpy(h + qB + bB, b->ranges.data(), rB);
        std::memcpy(h + qB + bB + rB, b->tf.data(), tB);
        LVR_BHIP(b, hipMemcpyAsync(b->d_reqs, h, qB, hipMemcpyHostToDevice, b->stream));
        if (bB)
        {
            LVR_BHIP(b, hipMemcpyAsync(b->d_keys, h + qB, bB, hipMemcpyHostToDevice,
                                       b->stream));
            LVR_BHIP(b, hipMemcpyAsync(b->d_ranges, h + qB + bB, rB, hipMemcpyHostToDevice,
                                       b->stream));
        }
        if (tB)
            LVR_BHIP(b, hipMemcpyAsync(b->d_tf, h + qB + bB + rB, tB, hipMemcpyHostToDevice,
                                       b->stream));
        LVR_BHIP(b, b->staging.release(slot, b->stream));
    }

    // the pools' page-table changes reach the device before the record pass
    // of the launch reads the tables
    for (lvr_ctx* c : b->ctxs)
        if ((rc = applyPageTable(c, b->stream)))
            return bfail(b, rc, c->error);

    auto enqueue = [&](hipStream_t st) -> hipError_t {
        enqueueRecords(st, b->d_reqs, nReq, b->d_keys, int(b->bricks.size()), b->d_bricks);
        enqueueBinning(st, b->d_reqs, nReq, b->d_ranges, supers, tiles, b->d_scounts,
                       b->d_soffsets, b->d_slists, b->d_counts, b->d_offsets, b->d_lists,
                       b->d_counters, nCounters);
        // the side streams start after the binning (beside the shared launch)
        if (nSide > 0)
        {
            hipError_t e = hipEventRecord(b->prepDone, st);
            for (int i = 0; i < nSide && e == hipSuccess; ++i)
                e = hipStreamWaitEvent(b->side[i], b->prepDone, 0);
            if (e != hipSuccess)
                return e;
        }
        // two persistent workgroups of wps * 2 wavefronts per CU the stream
        // may use
        if (items > 0)
        {
            const int blocks = std::max(1, std::min((items + 2 * wps - 1) / (2 * wps), slots));
            const dim3 grid(blocks), block(2 * wps * kWave);
#define LVR_BATCH(a, c, d)                                                                   \
    hipLaunchKernelGGL((batchKernel<a, c, d>), grid, block, 0, st, b->d_reqs, nReq, items,      \
                       b->d_bricks, b->d_offsets, b->d_counts, b->d_lists, b->d_counters)
            switch (depthCode)
            {
            case 0: LVR_BATCH(2, 2, 2); break;
            case 1: LVR_BATCH(2, 2, 3); break;
            case 2: LVR_BATCH(2, 3, 2); break;
            case 3: LVR_BATCH(2, 3, 3); break;
            case 4: LVR_BATCH(3, 2, 2); break;
            case 5: LVR_BATCH(3, 2, 3); break;
            case 6: LVR_BATCH(3, 3, 2); break;
            default: LVR_BATCH(3, 3, 3); break;
            }
#undef LVR_BATCH
        }
        // the solo requests through the single-frame kernel (uniforms by
        // value), each with its own sample counter and tile queue, on the
        // streams they were dealt to; the side streams rejoin `stream` at
        // the end
        for (int j : soloOrder)
        {
            const int r = b->solo[size_t(j)];
            const hipStream_t sj = soloLane[size_t(j)] == 0 ? st : b->side[soloLane[size_t(j)] - 1];
            FrameUniforms Us = b->reqs[size_t(r)];
            Us.wavesPerSimd = soloWps;
            unsigned long long* c = b->d_counters + nReq + 2 + 2 * j;
            const int sblocks =
                std::max(1, std::min((Us.nItems + 2 * soloWps - 1) / (2 * soloWps), slots));
            const size_t lds = rayLds(Us);
            switch (Us.dataType)
            {
#define LVR_SOLO(T)                                                                         \
    launchRaycastAs<T, false, 2, false>(sblocks, lds, sj, Us, b->d_bricks, b->d_offsets,     \
                                        b->d_counts, b->d_lists, c)
            case LVR_DT_UINT8: LVR_SOLO(uint8_t); break;
            case LVR_DT_UINT16: LVR_SOLO(uint16_t); break;
            case LVR_DT_UINT32: LVR_SOLO(uint32_t); break;
            case LVR_DT_INT8: LVR_SOLO(int8_t); break;
            case LVR_DT_INT16: LVR_SOLO(int16_t); break;
            case LVR_DT_INT32: LVR_SOLO(int32_t); break;
            default: LVR_SOLO(float); break;
            }
#undef LVR_SOLO
            // the request's composite steps where the batch keeps them
            const hipError_t e = hipMemcpyAsync(b->d_counters + 2 + r, c,
                                                sizeof(unsigned long long),
                                                hipMemcpyDeviceToDevice, sj);
            if (e != hipSuccess)
                return e;
        }
        for (int i = 0; i < nSide; ++i)
        {
            hipError_t e = hipEventRecord(b->sideDone[i], b->side[i]);
            if (e == hipSuccess)
                e = hipStreamWaitEvent(st, b->sideDone[i], 0);
            if (e != hipSuccess)
                return e;
        }
        return hipGetLastError();
    };

    LVR_BHIP(b, hipEventRecord(b->ev[0], b->stream));
    if (b->useGraph && nSolo == 0)
    {
        if (!b->exec || key != b->graphKey)
        {
            if (b->exec)
            {
                (void)hipGraphExecDestroy(b->exec);
                b->exec = nullptr;
            }
            hipGraph_t graph = nullptr;
            LVR_BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t e = enqueue(b->stream);
            const hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
            if (e != hipSuccess || e2 != hipSuccess)
            {
                if (graph)
                    (void)hipGraphDestroy(graph);
                return bfail(b, LVR_EHIP, std::string("graph capture: ") +
                                              hipGetErrorString(e != hipSuccess ? e : e2));
            }
            const hipError_t e3 = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e3 != hipSuccess)
                return bfail(b, LVR_EHIP, std::string("graph instantiate: ") +
                                              hipGetErrorString(e3));
            b->graphKey = key;
            ++b->graphCaptures;
        }
        LVR_BHIP(b, hipGraphLaunch(b->exec, b->stream));
        ++b->graphLaunches;
    }
    else
        LVR_BHIP(b, enqueue(b->stream));
    LVR_BHIP(b, hipEventRecord(b->ev[1], b->stream));
    LVR_BHIP(b, hipMemcpyAsync(b->h_pinned, b->d_counters,
                               sizeof(unsigned long long) * (size_t(nReq) + 2),
                               hipMemcpyDeviceToHost, b->stream));
    LVR_BHIP(b, hipEventRecord(b->done, b->stream));
    for (lvr_ctx* c : b->ctxs)
        if (std::find(c->readers.begin(), c->readers.end(), b) == c->readers.end())
            c->readers.push_back(b);
    b->ran = true;
    return LVR_OK;
}

int lvr_batch_set_output(lvr_batch* b, int request, void* device_rgba8)
{
    if (!b || request < 0 || size_t(request) >= b->reqs.size() || b->ran)
        return LVR_EINVAL;
    if (reinterpret_cast<uintptr_t>(device_rgba8) % 4u)
        return lvr::bfail(b, LVR_EINVAL, "output frame must be 4-byte aligned");
    b->outputs[size_t(request)] = static_cast<uint32_t*>(device_rgba8);
    return LVR_OK;
}

int lvr_batch_reserve_cus(lvr_batch* b, int n)
{
    if (!b || n < 0 || n >= b->numCUs)
        return LVR_EINVAL;
    LVR_BHIP(b, hipSetDevice(b->device));
    LVR_BHIP(b, hipStreamSynchronize(b->stream));
    hipStream_t st = nullptr;
    if (n == 0)
        LVR_BHIP(b, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    else
    {
        // n CUs spread evenly over the CU index space stay off the mask
        std::vector<uint32_t> mask(size_t((b->numCUs + 31) / 32), 0u);
        for (int c = 0; c < b->numCUs; ++c)
            mask[size_t(c / 32)] |= 1u << (c % 32);
        for (int j = 0; j < n; ++j)
        {
            const int c = int((int64_t(j) * b->numCUs) / n);
            mask[size_t(c / 32)] &= ~(1u << (c % 32));
        }
        LVR_BHIP(b, hipExtStreamCreateWithCUMask(&st, uint32_t(mask.size()), mask.data()));
    }
    if (b->exec)
    {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
        b->graphKey.clear();
    }
    (void)hipStreamDestroy(b->stream);
    b->stream = st;
    b->reservedCUs = n;
    return LVR_OK;
}

int lvr_batch_query(lvr_batch* b)
{
    if (!b)
        return LVR_EINVAL;
    const hipError_t e = hipStreamQuery(b->stream);
    if (e == hipSuccess)
        return 1;
    if (e == hipErrorNotReady)
        return 0;
    return lvr::bfail(b, LVR_EHIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int lvr_batch_wait(lvr_batch* b, lvr_batch_stats* stats)
{
    if (!b)
        return LVR_EINVAL;
    LVR_BHIP(b, hipSetDevice(b->device));
    LVR_BHIP(b, hipStreamSynchronize(b->stream));
    if (!b->ran)
        return lvr::bfail(b, LVR_EINVAL, "batch has not run");
    b->h_counters.assign(b->h_pinned, b->h_pinned + b->reqs.size() + 2);
    LVR_BHIP(b, hipEventElapsedTime(&b->kernelMs, b->ev[0], b->ev[1]));
    if (stats)
    {
        stats->requests = b->reqs.size();
        // the batch kernel's total, and the solo launches' counts
        uint64_t samples = b->h_counters[0];
        for (int r : b->solo)
            samples += b->h_counters[2 + size_t(r)];
        stats->samples = samples;
        stats->bricks = b->onScreen;
        stats->tile_entries = b->listTotal;
        stats->gpu_ms = b->kernelMs;
        stats->graph_captures = b->graphCaptures;
        stats->graph_launches = b->graphLaunches;
        stats->solo_requests = b->solo.size();
    }
    return LVR_OK;
}

int lvr_batch_read(lvr_batch* b, int request, uint8_t* out_rgba8, size_t bytes,
                   uint64_t* samples)
{
    if (!b || request < 0 || size_t(request) >= b->reqs.size())
        return LVR_EINVAL;
    if (b->h_counters.size() != b->reqs.size() + 2)
        return lvr::bfail(b, LVR_EINVAL, "call lvr_batch_wait first");
    const lvr::FrameUniforms& U = b->reqs[size_t(request)];
    const size_t need = size_t(U.width) * U.height * 4;
    if (out_rgba8)
    {
        if (bytes < need)
            return lvr::bfail(b, LVR_ERANGE, "frame buffer too small");
        const uint32_t* src = b->outputs[size_t(request)]
                                  ? b->outputs[size_t(request)]
                                  : b->d_frames + b->frameOffset[size_t(request)];
        LVR_BHIP(b, hipMemcpyAsync(out_rgba8, src, need, hipMemcpyDeviceToHost, b->stream));
        LVR_BHIP(b, hipStreamSynchronize(b->stream));
    }
    if (samples)
        *samples = b->h_counters[2 + size_t(request)];
    return LVR_OK;
}
}
