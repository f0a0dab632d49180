// Screen-space-error brick selection on the device.
//
// The reference walks the octree depth first on the CPU and asks
// SelectVisibles::visit for each node (DFSTraversal.cpp:29-76,
// SelectVisibles.cpp:45-107): culled nodes stop the descent, nodes whose
// projected voxel size is within the screen-space error are emitted, the
// others are refined. The decision for a node depends only on the node, so
// the same set is found level by level: one workgroup holds the current
// level's frontier, evaluates every node in parallel with the reference's
// FP32 expressions (contraction off, same operation order), appends visible
// nodes to the output and refined nodes' eight children to the next level.
// The depth-first order the renderer's sort starts from is restored on the
// host with a pre-order key (root index, then the child index x*4 + y*2 + z
// of each level, NodeId.cpp:92-113), which sorts like the recursion visits.
#include <algorithm>
#include <vector>

#include "lvr_internal.h"
#include "radix.h"

namespace lvr
{
namespace
{
struct SelectUniforms
{
    lvr_volume_info vi;
    float planes[6][4];
    float eye[3], dir[3];
    float nearPlane, top, bottom;
    float sse;
    uint32_t height, minLod, maxLod;
    int32_t capFrontier, capOut;
    int32_t nClip;
    float clip[6][4]; // normal xyz, d: the kept half-space is n . x + d >= 0
};

__device__ inline float planeDot(const float* p, const float* v)
{
    return p[0] * v[0] + p[1] * v[1] + p[2] * v[2] + p[3] * 1.0f;
}

__device__ inline void nearFar(const float* mn, const float* mx, const float* p,
                               float* vmin, float* vmax)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const bool pos = p[i] >= 0.f;
        vmin[i] = pos ? mn[i] : mx[i];
        vmax[i] = pos ? mx[i] : mn[i];
    }
}

// 0: culled, 1: visible, 2: refine
__device__ int visitNode(const SelectUniforms& U, uint64_t id)
{
    const NodeKey k = decodeNode(id);
    float mn[3], mx[3];
    nodeWorldBox(U.vi, k, mn, mx);
    for (int p = 0; p < 6; ++p)
    {
        float vmin[3], vmax[3];
        nearFar(mn, mx, U.planes[p], vmin, vmax);
        if (planeDot(U.planes[p], vmax) < 0.f)
            return 0;
    }
    // _clipPlanes.isOutside(worldBox) (SelectVisibles.cpp:69): the box lies
    // wholly in the half-space the shader clips away (fragRayCast.glsl:152-
    // 167 keeps n . x + d >= 0), tested at its corner furthest along n
    for (int p = 0; p < U.nClip; ++p)
    {
        float vmin[3], vmax[3];
        nearFar(mn, mx, U.clip[p], vmin, vmax);
        if (planeDot(U.clip[p], vmax) < 0.f)
            return 0;
    }
    const float* np = U.planes[4];
    float vmin[3], vmax[3];
    nearFar(mn, mx, np, vmin, vmax);
    if (planeDot(np, vmin) < 0.f || planeDot(np, vmax) < 0.f)
    {
#pragma unroll
        for (int i = 0; i < 3; ++i)
            vmin[i] = U.eye[i] - U.dir[i] * U.nearPlane;
    }
    float wspv = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const uint32_t block = U.vi.max_block[i] - 2u * U.vi.overlap[i];
        const float v = (mx[i] - mn[i]) / float(block);
        wspv = i == 0 ? v : fminf(wspv, v);
    }
    // SelectVisibles::Impl::isLODVisible
    const float worldSpacePerPixel = (U.top - U.bottom) / float(U.height);
    const float pixelPerVoxel = wspv / worldSpacePerPixel;
    const float distance = fabsf(planeDot(np, vmin));
    const float n = U.nearPlane;
    const float ppvd = pixelPerVoxel * n / (n + distance);
    bool visible = ppvd <= U.sse;
    visible = (visible && k.level >= U.minLod) || k.level == U.maxLod ||
              k.level == U.vi.depth - 1;
    return visible ? 1 : 2;
}

__global__ __launch_bounds__(1024) void selectKernel(
    const SelectUniforms U, uint64_t* __restrict__ bufA, uint64_t* __restrict__ bufB,
    uint64_t* __restrict__ out, int32_t* __restrict__ counters, uint64_t* __restrict__ hostOut,
    int32_t* __restrict__ hostCounters)
{
    __shared__ int nCur, nNext, nOut, overflow;
    const uint32_t rx = U.vi.root_blocks[0], ry = U.vi.root_blocks[1],
                   rz = U.vi.root_blocks[2];
    const int nRoots = int(rx * ry * rz);
    if (threadIdx.x == 0)
    {
        nCur = nRoots;
        nNext = 0;
        nOut = 0;
        overflow = nRoots > U.capFrontier;
    }
    __syncthreads();
    if (overflow)
    {
        if (threadIdx.x == 0)
        {
            counters[0] = hostCounters[0] = 0;
            counters[1] = hostCounters[1] = 1;
        }
        return;
    }
    for (int i = threadIdx.x; i < nRoots; i += blockDim.x)
    {
        const uint32_t x = uint32_t(i) / (ry * rz);
        const uint32_t y = (uint32_t(i) / rz) % ry;
        const uint32_t z = uint32_t(i) % rz;
        bufA[i] = encodeNode(0, x, y, z, 0);
    }
    uint64_t* cur = bufA;
    uint64_t* next = bufB;
    for (uint32_t level = 0; level < U.vi.depth; ++level)
    {
        __syncthreads();
        const int count = nCur;
        for (int i = threadIdx.x; i < count; i += blockDim.x)
        {
            const uint64_t id = cur[i];
            const int r = visitNode(U, id);
            if (r == 1)
            {
                const int k = atomicAdd(&nOut, 1);
                if (k < U.capOut)
                    out[k] = id;
                else
                    overflow = 1;
            }
            else if (r == 2 && level + 1 < U.vi.depth)
            {
                const int k = atomicAdd(&nNext, 8);
                if (k + 8 <= U.capFrontier)
                {
                    const NodeKey p = decodeNode(id);
                    for (int c = 0; c < 8; ++c)
                        next[k + c] = encodeNode(p.level + 1, 2 * p.x + (c >> 2),
                                                 2 * p.y + ((c >> 1) & 1),
                                                 2 * p.z + (c & 1), p.t);
                }
                else
                    overflow = 1;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            nCur = min(nNext, U.capFrontier);
            nNext = 0;
        }
        uint64_t* t = cur;
        cur = next;
        next = t;
    }
    __syncthreads();
    // the result straight into the caller's pinned host buffer
    const int n = min(nOut, U.capOut);
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        hostOut[i] = out[i];
    if (threadIdx.x == 0)
    {
        counters[0] = hostCounters[0] = n;
        counters[1] = hostCounters[1] = overflow;
    }
}

// pre-order key of a node: root index, then 3 bits per level
uint64_t dfsKey(const lvr_volume_info& vi, uint64_t id)
{
    const NodeKey k = decodeNode(id);
    const uint32_t rx = k.x >> k.level, ry = k.y >> k.level, rz = k.z >> k.level;
    uint64_t key = (uint64_t(rx) * vi.root_blocks[1] + ry) * vi.root_blocks[2] + rz;
    for (uint32_t l = 1; l < vi.depth; ++l)
    {
        uint64_t c = 0;
        if (l <= k.level)
        {
            const uint32_t s = k.level - l;
            c = (((k.x >> s) & 1u) << 2) | (((k.y >> s) & 1u) << 1) | ((k.z >> s) & 1u);
        }
        key = key * 8 + c;
    }
    return key;
}
} // namespace
} // namespace lvr

extern "C" {

int lvr_select_visibles_ex_async(lvr_ctx* ctx, const lvr_frustum* f,
                                 const lvr_select_params* sp)
{
    using namespace lvr;
    if (!ctx || !f || !sp || sp->window_height == 0 || sp->n_clip_planes < 0 ||
        sp->n_clip_planes > 6 || !(sp->range0 >= 0.f) || !(sp->range1 <= 1.f) ||
        !(sp->range0 <= sp->range1))
        return LVR_EINVAL;
    const uint32_t window_height = sp->window_height, min_lod = sp->min_lod,
                   max_lod = sp->max_lod;
    const float sse = sp->sse;
    int rc = ensureDevice(ctx);
    if (rc)
        return rc;
    if (ctx->selPending)
        return fail(ctx, LVR_EINVAL, "a selection is already pending");
    // frontier capacity: the largest level of the octree
    size_t maxLevel = 1;
    for (uint32_t l = 0; l < ctx->vi.depth; ++l)
        maxLevel = std::max<size_t>(maxLevel, size_t(ctx->pt.levelOffset[l + 1] -
                                                     ctx->pt.levelOffset[l]));
    const size_t capNeed = maxLevel;
    if (capNeed > ctx->selCapacity)
    {
        LVR_HIP(ctx, hipStreamSynchronize(ctx->selStream));
        (void)hipFree(ctx->d_selA);
        (void)hipFree(ctx->d_selB);
        (void)hipFree(ctx->d_selOut);
        ctx->d_selA = ctx->d_selB = ctx->d_selOut = nullptr;
        if (ctx->h_sel)
            ctx->retiredHost.push_back(ctx->h_sel);
        ctx->h_sel = nullptr;
        ctx->selCapacity = 0;
        LVR_HIP(ctx, hipMalloc((void**)&ctx->d_selA, sizeof(uint64_t) * capNeed));
        LVR_HIP(ctx, hipMalloc((void**)&ctx->d_selB, sizeof(uint64_t) * capNeed));
        LVR_HIP(ctx, hipMalloc((void**)&ctx->d_selOut, sizeof(uint64_t) * capNeed));
        // counters, then the ids; written by the kernel over the bus
        LVR_HIP(ctx, hipHostMalloc((void**)&ctx->h_sel, sizeof(uint64_t) * (capNeed + 1),
                                   hipHostMallocMapped | hipHostMallocCoherent));
        ctx->selCapacity = capNeed;
    }
    if (!ctx->d_selCount)
        LVR_HIP(ctx, hipMalloc((void**)&ctx->d_selCount, sizeof(int32_t) * 2));
    if (!ctx->selDone)
        LVR_HIP(ctx, hipEventCreateWithFlags(&ctx->selDone, hipEventDisableTiming));
    SelectUniforms U;
    U.vi = ctx->vi;
    for (int p = 0; p < 6; ++p)
        for (int j = 0; j < 4; ++j)
            U.planes[p][j] = f->planes[p][j];
    for (int i = 0; i < 3; ++i)
    {
        U.eye[i] = f->eye[i];
        U.dir[i] = f->view_dir[i];
    }
    U.nearPlane = f->near_plane;
    U.top = f->top;
    U.bottom = f->bottom;
    U.sse = sse;
    U.height = window_height;
    U.minLod = min_lod;
    U.maxLod = max_lod;
    U.capFrontier = int(std::min<size_t>(capNeed, 1u << 30));
    U.capOut = U.capFrontier;
    U.nClip = sp->n_clip_planes;
    for (int p = 0; p < 6; ++p)
        for (int j = 0; j < 4; ++j)
            U.clip[p][j] = p < U.nClip ? sp->clip_planes[4 * p + j] : 0.f;
    ctx->selRange[0] = sp->range0;
    ctx->selRange[1] = sp->range1;
    uint8_t* dHost = nullptr;
    LVR_HIP(ctx, hipHostGetDevicePointer((void**)&dHost, ctx->h_sel, 0));
    hipLaunchKernelGGL(selectKernel, dim3(1), dim3(1024), 0, ctx->selStream, U,
                       ctx->d_selA, ctx->d_selB, ctx->d_selOut, ctx->d_selCount,
                       reinterpret_cast<uint64_t*>(dHost + sizeof(uint64_t)),
                       reinterpret_cast<int32_t*>(dHost));
    LVR_HIP(ctx, hipGetLastError());
    LVR_HIP(ctx, hipEventRecord(ctx->selDone, ctx->selStream));
    ctx->selPending = true;
    return LVR_OK;
}

int lvr_select_visibles_wait(lvr_ctx* ctx, uint64_t* out, int cap, int* n_out)
{
    using namespace lvr;
    if (!ctx || !n_out || cap < 0 || (cap > 0 && !out))
        return LVR_EINVAL;
    if (!ctx->selPending)
        return fail(ctx, LVR_EINVAL, "no selection pending");
    ctx->selPending = false;
    LVR_HIP(ctx, hipEventSynchronize(ctx->selDone));
    const int32_t* counts = reinterpret_cast<const int32_t*>(ctx->h_sel);
    if (counts[1])
        return fail(ctx, LVR_ENOMEM, "selection frontier overflow");
    const int n = counts[0];
    // SelectVisibles::Impl::visitPost (SelectVisibles.cpp:109-131): the
    // sort-last share [range0 * N, range1 * N) of the DFS-ordered visibles,
    // the products formed in float and truncated as the reference does
    const size_t start = size_t(ctx->selRange[0] * float(n));
    const size_t end = std::min(size_t(ctx->selRange[1] * float(n)), size_t(n));
    const int kept = end > start ? int(end - start) : 0;
    *n_out = kept;
    if (kept > cap)
        return fail(ctx, LVR_ERANGE, "visible set larger than the output buffer");
    const uint64_t* ids = reinterpret_cast<const uint64_t*>(ctx->h_sel + sizeof(uint64_t));
    std::vector<std::pair<uint64_t, uint64_t>> keyed(static_cast<size_t>(n)), tmp;
    for (int i = 0; i < n; ++i)
        keyed[size_t(i)] = {dfsKey(ctx->vi, ids[i]), ids[i]};
    // (key, id) order: the radix pass orders by key; equal keys (an ancestor
    // beside its first-child descendants) are ordered by id as well
    radixSortByKey(keyed, tmp);
    if (hasEqualKeys(keyed))
        std::sort(keyed.begin(), keyed.end());
    for (int i = 0; i < kept; ++i)
        out[i] = keyed[start + size_t(i)].second;
    return LVR_OK;
}

int lvr_select_visibles_ex(lvr_ctx* ctx, const lvr_frustum* f, const lvr_select_params* sp,
                           uint64_t* out, int cap, int* n_out)
{
    if (!ctx || !f || !sp || !n_out || cap < 0 || (cap > 0 && !out))
        return LVR_EINVAL;
    const int rc = lvr_select_visibles_ex_async(ctx, f, sp);
    return rc ? rc : lvr_select_visibles_wait(ctx, out, cap, n_out);
}

int lvr_select_visibles_async(lvr_ctx* ctx, const lvr_frustum* f, uint32_t window_height,
                              float sse, uint32_t min_lod, uint32_t max_lod)
{
    lvr_select_params sp{};
    sp.window_height = window_height;
    sp.sse = sse;
    sp.min_lod = min_lod;
    sp.max_lod = max_lod;
    sp.range1 = 1.f;
    return lvr_select_visibles_ex_async(ctx, f, &sp);
}

int lvr_select_visibles(lvr_ctx* ctx, const lvr_frustum* f,
                        uint32_t window_height, float sse, uint32_t min_lod,
                        uint32_t max_lod, uint64_t* out, int cap, int* n_out)
{
    if (!ctx || !f || !n_out || cap < 0 || (cap > 0 && !out) || window_height == 0)
        return LVR_EINVAL;
    const int rc =
        lvr_select_visibles_async(ctx, f, window_height, sse, min_lod, max_lod);
    return rc ? rc : lvr_select_visibles_wait(ctx, out, cap, n_out);
}
}
