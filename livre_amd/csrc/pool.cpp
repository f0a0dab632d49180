// Context, HBM brick pool, device page table, brick cache policy and
// transfer-function upload.
//
// The pool is one hipMalloc arena of fixed-size slots (one slot per brick of
// maximumBlockSize^3 voxels, 256-byte aligned), the HBM counterpart of
// TexturePool's one-3D-texture-per-brick scheme (TexturePool.cpp:90-126).
// The page table maps a NodeId to its slot. The ray-cast kernels read the
// device copy when they draw a brick, as the reference looks each brick's
// texture up by NodeId at draw time (RayCastRenderer.cpp:510-522); the host
// copy is authoritative and its changes reach the device through
// applyPageTable, ordered on the render stream between launches.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "lvr_internal.h"

namespace lvr
{
int fail(lvr_ctx* ctx, int code, const std::string& msg)
{
    if (ctx)
        ctx->error = msg;
    return code;
}

int hipFail(lvr_ctx* ctx, hipError_t e, const char* what)
{
    return fail(ctx, LVR_EHIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

int ensureDevice(lvr_ctx* ctx)
{
    LVR_HIP(ctx, hipSetDevice(ctx->device));
    return LVR_OK;
}

bool pageIndex(const lvr_ctx* ctx, uint64_t id, uint64_t* idx)
{
    const NodeKey k = decodeNode(id);
    if (k.level >= ctx->vi.depth || k.t != 0)
        return false;
    const uint32_t* d = &ctx->pt.levelDim[3 * k.level];
    if (k.x >= d[0] || k.y >= d[1] || k.z >= d[2])
        return false;
    *idx = ctx->pt.levelOffset[k.level] + k.x +
           uint64_t(d[0]) * (k.y + uint64_t(d[1]) * k.z);
    return true;
}

int32_t slotOf(const lvr_ctx* ctx, uint64_t id)
{
    uint64_t idx;
    return pageIndex(ctx, id, &idx) ? ctx->pt.h_slots[idx] : -1;
}

hipError_t PinnedRing::acquire(size_t bytes, uint8_t** out, int* slot)
{
    const int r = next;
    next ^= 1;
    hipError_t e = hipSuccess;
    if (!done[r] && (e = hipEventCreateWithFlags(&done[r], hipEventDisableTiming)) != hipSuccess)
        return e;
    if (armed[r] && (e = hipEventSynchronize(done[r])) != hipSuccess)
        return e;
    armed[r] = false;
    if (bytes > cap[r])
    {
        if (buf[r])
            retired.push_back(buf[r]);
        buf[r] = nullptr;
        const size_t n = std::max(bytes + bytes / 4 + 4096, 2 * cap[r]);
        cap[r] = 0;
        if ((e = hipHostMalloc((void**)&buf[r], n, hipHostMallocDefault)) != hipSuccess)
            return e;
        cap[r] = n;
    }
    *out = buf[r];
    *slot = r;
    return hipSuccess;
}

hipError_t PinnedRing::release(int slot, hipStream_t st)
{
    const hipError_t e = hipEventRecord(done[slot], st);
    armed[slot] = e == hipSuccess;
    return e;
}

void PinnedRing::destroy()
{
    for (int r = 0; r < 2; ++r)
    {
        if (armed[r])
            (void)hipEventSynchronize(done[r]);
        if (buf[r])
            (void)hipHostFree(buf[r]);
        if (done[r])
            (void)hipEventDestroy(done[r]);
        buf[r] = nullptr;
        done[r] = nullptr;
        cap[r] = 0;
        armed[r] = false;
    }
    for (uint8_t* b : retired)
        (void)hipHostFree(b);
    retired.clear();
}

int32_t slotAndTouch(lvr_ctx* ctx, uint64_t id)
{
    uint64_t idx;
    if (!pageIndex(ctx, id, &idx))
        return -1;
    const int32_t slot = ctx->pt.h_slots[idx];
    if (slot >= 0 && ctx->policy == 0)
        ctx->lru.moveToBack(int32_t(idx));
    return slot;
}

void touch(lvr_ctx* ctx, uint64_t id)
{
    (void)slotAndTouch(ctx, id);
}

static int setPage(lvr_ctx* ctx, uint64_t id, int32_t slot)
{
    uint64_t idx;
    if (!pageIndex(ctx, id, &idx))
        return fail(ctx, LVR_EINVAL, "node id outside the octree");
    ctx->pt.h_slots[idx] = slot;
    if (!ctx->pt.isDirty[idx])
    {
        ctx->pt.isDirty[idx] = 1;
        ctx->pt.dirty.push_back(uint32_t(idx));
    }
    return LVR_OK;
}

// table[page] = slot for each listed pair
__global__ __launch_bounds__(256) void pageScatterKernel(const int2* __restrict__ upd, int n,
                                                         int32_t* __restrict__ table)
{
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
        table[upd[i].x] = upd[i].y;
}

int applyPageTable(lvr_ctx* ctx, hipStream_t st)
{
    PageTable& pt = ctx->pt;
    if (pt.dirty.empty())
    {
        // the last update may have run on another stream
        if (ctx->tableRecorded)
            LVR_HIP(ctx, hipStreamWaitEvent(st, ctx->tableReady, 0));
        return LVR_OK;
    }
    if (ctx->tableRecorded)
        LVR_HIP(ctx, hipStreamWaitEvent(st, ctx->tableReady, 0));
    size_t done = 0;
    while (done < pt.dirty.size())
    {
        const int cnt = int(std::min<size_t>(PageTable::kUpdateChunk, pt.dirty.size() - done));
        // the pinned pairs are rewritten only once the copy that last read
        // them has completed
        if (pt.updateArmed)
            LVR_HIP(ctx, hipEventSynchronize(pt.updateDone));
        pt.updateArmed = false;
        for (int i = 0; i < cnt; ++i)
        {
            const uint32_t idx = pt.dirty[done + size_t(i)];
            pt.h_update[i] = make_int2(int(idx), pt.h_slots[idx]);
            pt.isDirty[idx] = 0;
        }
        LVR_HIP(ctx, hipMemcpyAsync(pt.d_update, pt.h_update, sizeof(int2) * size_t(cnt),
                                    hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(pageScatterKernel, dim3((cnt + 255) / 256), dim3(256), 0, st,
                           pt.d_update, cnt, pt.d_slots);
        LVR_HIP(ctx, hipGetLastError());
        LVR_HIP(ctx, hipEventRecord(pt.updateDone, st));
        pt.updateArmed = true;
        done += size_t(cnt);
    }
    pt.dirty.clear();
    LVR_HIP(ctx, hipEventRecord(ctx->tableReady, st));
    ctx->tableRecorded = true;
    return LVR_OK;
}

int pinBrick(lvr_ctx* ctx, uint64_t id, int delta)
{
    auto it = ctx->resident.find(id);
    if (it == ctx->resident.end())
        return fail(ctx, LVR_ENOENT, "pin of a brick that is not resident");
    const int before = it->second.pinned;
    const int after = before + delta;
    if (after < 0)
        return fail(ctx, LVR_EINVAL, "unbalanced brick unpin");
    it->second.pinned = after;
    if (before == 0 && after > 0)
        ++ctx->pinnedBricks;
    else if (before > 0 && after == 0)
        --ctx->pinnedBricks;
    return LVR_OK;
}

uint64_t evictableSlots(const lvr_ctx* ctx)
{
    return ctx->freeSlots.size() + (ctx->resident.size() - ctx->pinnedBricks);
}

int waitPoolReaders(lvr_ctx* ctx, hipStream_t st)
{
    // the last histogram (its own stream) may still read any slot
    if (ctx->histDone)
    {
        if (st)
            LVR_HIP(ctx, hipStreamWaitEvent(st, ctx->histDone, 0));
        else
            LVR_HIP(ctx, hipEventSynchronize(ctx->histDone));
    }
    if (ctx->pinsGuardReaders || ctx->readers.empty())
        return LVR_OK;
    std::vector<lvr_batch*> still;
    for (lvr_batch* b : ctx->readers)
    {
        hipEvent_t e = batchDoneEvent(b);
        const hipError_t q = hipEventQuery(e);
        if (q == hipSuccess)
            continue; // finished: nothing to wait for
        if (q != hipErrorNotReady)
            return hipFail(ctx, q, "hipEventQuery(batch)");
        if (st)
            LVR_HIP(ctx, hipStreamWaitEvent(st, e, 0));
        else
            LVR_HIP(ctx, hipEventSynchronize(e));
        if (st)
            still.push_back(b);
    }
    ctx->readers.swap(still);
    return LVR_OK;
}

void* genScratch(lvr_ctx* ctx, size_t bytes)
{
    if (ctx->genCapacity >= bytes)
        return ctx->d_gen;
    // the previous scratch may still be read by queued source kernels
    if (hipStreamSynchronize(ctx->stream) != hipSuccess)
    {
        fail(ctx, LVR_EHIP, "source scratch sync");
        return nullptr;
    }
    (void)hipFree(ctx->d_gen);
    ctx->d_gen = nullptr;
    ctx->genCapacity = 0;
    const size_t cap = std::max<size_t>(bytes, size_t(1) << 20);
    if (hipMalloc(&ctx->d_gen, cap) != hipSuccess)
    {
        fail(ctx, LVR_ENOMEM, "source scratch");
        return nullptr;
    }
    ctx->genCapacity = cap;
    return ctx->d_gen;
}

// Evict the first unpinned brick in policy order (Cache.cpp:111-123 walks
// the queue in insertion order and skips referenced objects).
static int evictOne(lvr_ctx* ctx)
{
    for (int32_t i = ctx->lru.head; i >= 0; i = ctx->lru.next[size_t(i)])
    {
        const uint64_t id = ctx->lru.ids[size_t(i)];
        auto r = ctx->resident.find(id);
        if (r->second.pinned)
            continue;
        const int32_t slot = r->second.slot;
        ctx->lru.erase(i);
        ctx->resident.erase(r);
        ctx->freeSlots.push_back(slot);
        return setPage(ctx, id, -1);
    }
    return fail(ctx, LVR_ENOMEM, "brick pool full and every brick pinned");
}

// Find or allocate a slot for id; returns the slot index or a negative code.
static int acquireSlot(lvr_ctx* ctx, uint64_t id, int32_t* slot)
{
    uint64_t idx;
    if (!pageIndex(ctx, id, &idx))
        return fail(ctx, LVR_EINVAL, "node id outside the octree");
    ++ctx->poolGeneration; // the slot's contents are about to change
    auto it = ctx->resident.find(id);
    if (it != ctx->resident.end())
    {
        // the brick's voxels are about to be rewritten in place
        *slot = it->second.slot;
        if (!ctx->histCachedId.empty())
            ctx->histCachedId[size_t(*slot)] = ~0ull;
        touch(ctx, id);
        return LVR_OK;
    }
    if (ctx->freeSlots.empty())
    {
        const int rc = evictOne(ctx);
        if (rc != LVR_OK)
            return rc;
    }
    *slot = ctx->freeSlots.back();
    ctx->freeSlots.pop_back();
    if (!ctx->histCachedId.empty())
        ctx->histCachedId[size_t(*slot)] = ~0ull; // its cached counts are stale
    ctx->lru.pushBack(int32_t(idx), id);
    PoolEntry e;
    e.slot = *slot;
    e.pinned = 0;
    ctx->resident.emplace(id, e);
    return setPage(ctx, id, *slot);
}

int acquireSlotFor(lvr_ctx* ctx, uint64_t id, int32_t* slot)
{
    return acquireSlot(ctx, id, slot);
}

int releaseSlot(lvr_ctx* ctx, uint64_t id)
{
    auto it = ctx->resident.find(id);
    if (it == ctx->resident.end())
        return LVR_OK;
    if (it->second.pinned)
        --ctx->pinnedBricks;
    ctx->freeSlots.push_back(it->second.slot);
    uint64_t idx;
    if (pageIndex(ctx, id, &idx))
        ctx->lru.erase(int32_t(idx));
    ctx->resident.erase(it);
    ++ctx->poolGeneration;
    return setPage(ctx, id, -1);
}

// Linear (x fastest, getData layout) <-> tiled pool layout, one brick per
// grid row. `linear` holds the bricks back to back; slots[b] is the pool slot
// of brick b.
template <typename V>
__global__ void convertLayoutKernel(V* __restrict__ linear,
                                    const int32_t* __restrict__ slots,
                                    uint8_t* __restrict__ pool, size_t slotStride,
                                    BrickLayout L, int toTiled)
{
    const uint32_t n = L.dim[0] * L.dim[1] * L.dim[2];
    V* lin = linear + size_t(blockIdx.y) * n;
    V* tiled = reinterpret_cast<V*>(pool + size_t(slots[blockIdx.y]) * slotStride);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x)
    {
        const uint32_t x = i % L.dim[0];
        const uint32_t y = (i / L.dim[0]) % L.dim[1];
        const uint32_t z = i / (L.dim[0] * L.dim[1]);
        const uint32_t t = tiledIndex(L, x, y, z);
        if (toTiled)
            tiled[t] = lin[i];
        else
            lin[i] = tiled[t];
    }
}

int convertLayout(lvr_ctx* ctx, int nBricks, bool toTiled, hipStream_t st)
{
    const dim3 grid(64, unsigned(nBricks)), block(256);
    switch (ctx->voxelBytes)
    {
    case 1:
        hipLaunchKernelGGL(convertLayoutKernel<uint8_t>, grid, block, 0, st,
                           reinterpret_cast<uint8_t*>(ctx->d_staging), ctx->d_stagingSlots,
                           ctx->d_pool, ctx->slotStride, ctx->layout, toTiled ? 1 : 0);
        break;
    case 2:
        hipLaunchKernelGGL(convertLayoutKernel<uint16_t>, grid, block, 0, st,
                           reinterpret_cast<uint16_t*>(ctx->d_staging), ctx->d_stagingSlots,
                           ctx->d_pool, ctx->slotStride, ctx->layout, toTiled ? 1 : 0);
        break;
    default:
        hipLaunchKernelGGL(convertLayoutKernel<uint32_t>, grid, block, 0, st,
                           reinterpret_cast<uint32_t*>(ctx->d_staging), ctx->d_stagingSlots,
                           ctx->d_pool, ctx->slotStride, ctx->layout, toTiled ? 1 : 0);
        break;
    }
    LVR_HIP(ctx, hipGetLastError());
    return LVR_OK;
}

// Device buffer for `bricks` linear bricks plus their slot list.
int ensureDeviceStaging(lvr_ctx* ctx, size_t bricks)
{
    if (ctx->stagingBricks >= bricks)
        return LVR_OK;
    LVR_HIP(ctx, hipStreamSynchronize(ctx->uploadStream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(ctx->d_staging);
    (void)hipFree(ctx->d_stagingSlots);
    ctx->d_staging = nullptr;
    ctx->d_stagingSlots = nullptr;
    ctx->stagingBricks = 0;
    LVR_HIP(ctx, hipMalloc((void**)&ctx->d_staging, bricks * ctx->brickBytes));
    LVR_HIP(ctx, hipMalloc((void**)&ctx->d_stagingSlots, bricks * sizeof(int32_t)));
    ctx->stagingBricks = bricks;
    return LVR_OK;
}

static int ensureStaging(lvr_ctx* ctx, size_t bytes)
{
    if (ctx->stagingBytes >= bytes)
        return LVR_OK;
    if (ctx->h_staging)
    {
        LVR_HIP(ctx, hipStreamSynchronize(ctx->uploadStream));
        LVR_HIP(ctx, hipHostFree(ctx->h_staging));
        ctx->h_staging = nullptr;
    }
    LVR_HIP(ctx, hipHostMalloc((void**)&ctx->h_staging, bytes,
                               hipHostMallocDefault));
    ctx->stagingBytes = bytes;
    return LVR_OK;
}
} // namespace lvr

extern "C" {

int lvr_ctx_create(const lvr_volume_info* info, size_t pool_bytes, int device,
                   lvr_ctx** out)
{
    if (!info || !out)
        return LVR_EINVAL;
    *out = nullptr;
    lvr_volume_info vi = *info;
    if (lvr_fill_regular_volume_info(&vi) != LVR_OK)
        return LVR_EINVAL;
    lvr_ctx* ctx = new lvr_ctx();
    ctx->device = device;
    ctx->vi = vi;
    ctx->voxelBytes = lvr::bytesPerVoxel(vi.data_type);
    ctx->brickBytes = size_t(vi.max_block[0]) * vi.max_block[1] *
                      vi.max_block[2] * ctx->voxelBytes;
    ctx->layout = lvr::makeLayout(vi.max_block);
    ctx->slotStride =
        (size_t(lvr::layoutVoxels(ctx->layout)) * ctx->voxelBytes + 255) & ~size_t(255);
    ctx->nSlots = pool_bytes / ctx->slotStride;
    auto bail = [&](int rc) {
        std::string msg = ctx->error;
        lvr_ctx_destroy(ctx);
        (void)msg;
        return rc;
    };
    if (ctx->nSlots == 0 || ctx->nSlots > (1ull << 31))
        return bail(LVR_EINVAL);
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return bail(rc);
    if (hipDeviceGetAttribute(&ctx->numCUs, hipDeviceAttributeMultiprocessorCount,
                              device) != hipSuccess ||
        ctx->numCUs <= 0)
        return bail(LVR_EHIP);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) !=
            hipSuccess ||
        hipStreamCreateWithFlags(&ctx->uploadStream, hipStreamNonBlocking) !=
            hipSuccess ||
        hipStreamCreateWithFlags(&ctx->histStream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->selStream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->histDone, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->renderDone, hipEventDisableTiming) !=
            hipSuccess)
        return bail(LVR_EHIP);
    if (hipMalloc((void**)&ctx->d_pool, ctx->nSlots * ctx->slotStride) !=
        hipSuccess)
        return bail(LVR_ENOMEM);

    // page table: one entry per node of the full octree
    ctx->pt.levelOffset.resize(vi.depth + 1);
    ctx->pt.levelDim.resize(3 * vi.depth);
    uint64_t total = 0;
    for (uint32_t l = 0; l < vi.depth; ++l)
    {
        ctx->pt.levelOffset[l] = total;
        uint64_t cnt = 1;
        for (int a = 0; a < 3; ++a)
        {
            ctx->pt.levelDim[3 * l + a] = vi.root_blocks[a] << l;
            cnt *= ctx->pt.levelDim[3 * l + a];
        }
        total += cnt;
    }
    ctx->pt.levelOffset[vi.depth] = total;
    ctx->pt.size = total;
    if (total > uint64_t(INT32_MAX))
        return bail(LVR_EINVAL);
    ctx->pt.h_slots.assign(total, -1);
    ctx->pt.isDirty.assign(total, 0);
    ctx->lru.init(size_t(total));
    if (hipMalloc((void**)&ctx->pt.d_slots, total * sizeof(int32_t)) !=
        hipSuccess)
        return bail(LVR_ENOMEM);
    if (hipMemcpy(ctx->pt.d_slots, ctx->pt.h_slots.data(),
                  total * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess)
        return bail(LVR_EHIP);
    if (hipHostMalloc((void**)&ctx->pt.h_update, sizeof(int2) * lvr::PageTable::kUpdateChunk,
                      hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&ctx->pt.d_update, sizeof(int2) * lvr::PageTable::kUpdateChunk) !=
            hipSuccess ||
        hipEventCreateWithFlags(&ctx->pt.updateDone, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->tableReady, hipEventDisableTiming) != hipSuccess)
        return bail(LVR_ENOMEM);

    rc = lvr::noallocSetup(ctx);
    if (rc)
        return bail(rc);
    ctx->freeSlots.reserve(ctx->nSlots);
    for (int64_t s = int64_t(ctx->nSlots) - 1; s >= 0; --s)
        ctx->freeSlots.push_back(int32_t(s));
    for (auto& f : ctx->frames)
        if (hipEventCreateWithFlags(&f.done, hipEventDisableTiming) != hipSuccess)
            return bail(LVR_EHIP);
    if (hipHostMalloc((void**)&ctx->h_frameSamples,
                      lvr_ctx::kFrameSlots * sizeof(uint64_t), hipHostMallocDefault) !=
        hipSuccess)
        return bail(LVR_ENOMEM);
    if (hipHostMalloc((void**)&ctx->h_counters, 4 * sizeof(unsigned long long),
                      hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&ctx->d_counters, 4 * sizeof(unsigned long long)) !=
            hipSuccess)
        return bail(LVR_ENOMEM);
    *out = ctx;
    return LVR_OK;
}

void lvr_ctx_destroy(lvr_ctx* ctx)
{
    if (!ctx)
        return;
    (void)hipSetDevice(ctx->device);
    for (lvr_batch* b : ctx->readers)
    {
        (void)hipEventSynchronize(lvr::batchDoneEvent(b));
        lvr::batchForgetCtx(b, ctx);
    }
    if (ctx->stream)
        (void)hipStreamSynchronize(ctx->stream);
    if (ctx->uploadStream)
        (void)hipStreamSynchronize(ctx->uploadStream);
    if (ctx->histStream)
        (void)hipStreamSynchronize(ctx->histStream);
    if (ctx->selStream)
        (void)hipStreamSynchronize(ctx->selStream);
    lvr::destroyPyramid(ctx);
    lvr::destroyJpeg(ctx);
    for (float4* l : ctx->d_layers)
        if (l)
            (void)hipFree(l);
    for (hipEvent_t e : ctx->compositeEv)
        if (e)
            (void)hipEventDestroy(e);
    for (void* p : ctx->retired)
        (void)hipFree(p);
    ctx->retired.clear();
    for (void* p : ctx->retiredHost)
        (void)hipHostFree(p);
    ctx->retiredHost.clear();
    (void)hipFree(ctx->d_pool);
    (void)hipFree(ctx->d_gen);
    (void)hipFree(ctx->d_cls);
    (void)hipFree(ctx->pt.d_slots);
    (void)hipFree(ctx->pt.d_update);
    if (ctx->pt.h_update)
        (void)hipHostFree(ctx->pt.h_update);
    if (ctx->pt.updateDone)
        (void)hipEventDestroy(ctx->pt.updateDone);
    if (ctx->tableReady)
        (void)hipEventDestroy(ctx->tableReady);
    (void)hipFree(ctx->d_tf);
    (void)hipFree(ctx->d_frame);
    if (ctx->recStream)
    {
        (void)hipStreamSynchronize(ctx->recStream);
        (void)hipStreamDestroy(ctx->recStream);
    }
    for (auto& rs : ctx->recSets)
    {
        void* bufs[] = {rs.req,        rs.bricks,     rs.keys,      rs.ranges,    rs.tileCount,
                        rs.tileOffset, rs.tileList,   rs.superCount, rs.superOffset,
                        rs.superList,  rs.counters,   rs.superOrder};
        for (void* p : bufs)
            (void)hipFree(p);
        if (rs.ready)
            (void)hipEventDestroy(rs.ready);
        if (rs.free)
            (void)hipEventDestroy(rs.free);
    }
    (void)hipFree(ctx->d_counters);
    (void)hipFree(ctx->d_hist);
    (void)hipFree(ctx->d_histCache);
    for (int r = 0; r < 2; ++r)
    {
        if (ctx->h_histRing[r])
            (void)hipHostFree(ctx->h_histRing[r]);
        if (ctx->histRingDone[r])
            (void)hipEventDestroy(ctx->histRingDone[r]);
    }
    (void)hipFree(ctx->d_selA);
    (void)hipFree(ctx->d_selB);
    (void)hipFree(ctx->d_selOut);
    (void)hipFree(ctx->d_selCount);
    (void)hipFree(ctx->d_staging);
    (void)hipFree(ctx->d_stagingSlots);
    if (ctx->h_staging)
        (void)hipHostFree(ctx->h_staging);
    if (ctx->h_counters)
        (void)hipHostFree(ctx->h_counters);
    if (ctx->noallocArmed)
        (void)hipEventSynchronize(ctx->noallocCopied);
    if (ctx->h_noalloc)
        (void)hipHostFree(ctx->h_noalloc);
    if (ctx->noallocCopied)
        (void)hipEventDestroy(ctx->noallocCopied);
    ctx->frameStaging.destroy();
    if (ctx->h_sel)
        (void)hipHostFree(ctx->h_sel);
    if (ctx->selDone)
        (void)hipEventDestroy(ctx->selDone);
    if (ctx->renderDone)
        (void)hipEventDestroy(ctx->renderDone);
    for (auto& f : ctx->frames)
    {
        for (auto& e : f.ev)
            if (e)
                (void)hipEventDestroy(e);
        if (f.done)
            (void)hipEventDestroy(f.done);
    }
    if (ctx->h_frameSamples)
        (void)hipHostFree(ctx->h_frameSamples);
    if (ctx->stream)
        (void)hipStreamDestroy(ctx->stream);
    if (ctx->uploadStream)
        (void)hipStreamDestroy(ctx->uploadStream);
    if (ctx->histStream)
        (void)hipStreamDestroy(ctx->histStream);
    if (ctx->selStream)
        (void)hipStreamDestroy(ctx->selStream);
    if (ctx->histDone)
        (void)hipEventDestroy(ctx->histDone);
    delete ctx;
}

int lvr_ctx_query(lvr_ctx* ctx)
{
    if (!ctx)
        return LVR_EINVAL;
    if (!ctx->stream)
        return 1;
    const hipError_t e = hipStreamQuery(ctx->stream);
    if (e == hipSuccess)
        return 1;
    if (e == hipErrorNotReady)
        return 0;
    return lvr::hipFail(ctx, e, "hipStreamQuery");
}

const char* lvr_last_error(const lvr_ctx* ctx)
{
    return ctx ? ctx->error.c_str() : "null context";
}

int lvr_ctx_sync(lvr_ctx* ctx)
{
    if (!ctx)
        return LVR_EINVAL;
    LVR_HIP(ctx, hipSetDevice(ctx->device));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->uploadStream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->histStream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->selStream));
    if (ctx->recStream)
        LVR_HIP(ctx, hipStreamSynchronize(ctx->recStream));
    for (void* p : ctx->retired)
        (void)hipFree(p);
    ctx->retired.clear();
    for (void* p : ctx->retiredHost)
        (void)hipHostFree(p);
    ctx->retiredHost.clear();
    return LVR_OK;
}

int lvr_slot_bytes(const lvr_volume_info* info, uint64_t* out)
{
    if (!info || !out)
        return LVR_EINVAL;
    lvr_volume_info vi = *info;
    if (lvr_fill_regular_volume_info(&vi) != LVR_OK)
        return LVR_EINVAL;
    const lvr::BrickLayout L = lvr::makeLayout(vi.max_block);
    *out = (uint64_t(lvr::layoutVoxels(L)) * lvr::bytesPerVoxel(vi.data_type) + 255) &
           ~uint64_t(255);
    return LVR_OK;
}

int lvr_pool_set_policy(lvr_ctx* ctx, int policy)
{
    if (!ctx || (policy != 0 && policy != 1))
        return LVR_EINVAL;
    ctx->policy = policy;
    return LVR_OK;
}

int lvr_pool_stats(const lvr_ctx* ctx, uint64_t* slots, uint64_t* used,
                   uint64_t* slot_bytes)
{
    if (!ctx)
        return LVR_EINVAL;
    if (slots)
        *slots = ctx->nSlots;
    if (used)
        *used = ctx->resident.size();
    if (slot_bytes)
        *slot_bytes = ctx->slotStride;
    return LVR_OK;
}

int lvr_brick_upload(lvr_ctx* ctx, uint64_t node_id, const void* voxels,
                     size_t bytes)
{
    return lvr_brick_upload_many(ctx, &node_id, 1, voxels, bytes);
}

int lvr_brick_upload_many(lvr_ctx* ctx, const uint64_t* node_ids, int n,
                          const void* voxels, size_t bytes_per_brick)
{
    if (!ctx || n < 0 || (n > 0 && (!node_ids || !voxels)))
        return LVR_EINVAL;
    if (bytes_per_brick != ctx->brickBytes)
        return lvr::fail(ctx, LVR_EINVAL, "brick size does not match "
                                          "maximumBlockSize^3 * bytesPerVoxel");
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    if (n == 0)
        return LVR_OK;
    // frames queued on the render stream, and batches on their own streams,
    // may still read slots evicted here
    LVR_HIP(ctx, hipEventRecord(ctx->renderDone, ctx->stream));
    LVR_HIP(ctx, hipStreamWaitEvent(ctx->uploadStream, ctx->renderDone, 0));
    rc = lvr::waitPoolReaders(ctx, ctx->uploadStream);
    if (rc)
        return rc;
    // stage in chunks of up to 64 MiB: pinned host -> device linear -> tiled slots
    const size_t chunkBricks = std::max<size_t>(1, (64u << 20) / ctx->brickBytes);
    const size_t chunk = std::min<size_t>(chunkBricks, size_t(n));
    rc = lvr::ensureStaging(ctx, chunk * (ctx->brickBytes + sizeof(int32_t)));
    if (rc)
        return rc;
    rc = lvr::ensureDeviceStaging(ctx, chunk);
    if (rc)
        return rc;
    const uint8_t* src = static_cast<const uint8_t*>(voxels);
    for (int base = 0; base < n; base += int(chunkBricks))
    {
        const int cnt = std::min<int>(int(chunkBricks), n - base);
        LVR_HIP(ctx, hipStreamSynchronize(ctx->uploadStream));
        std::memcpy(ctx->h_staging, src + size_t(base) * ctx->brickBytes,
                    size_t(cnt) * ctx->brickBytes);
        int32_t* slots = reinterpret_cast<int32_t*>(ctx->h_staging + chunk * ctx->brickBytes);
        for (int i = 0; i < cnt; ++i)
        {
            rc = lvr::acquireSlot(ctx, node_ids[base + i], &slots[i]);
            if (rc)
                return rc;
        }
        LVR_HIP(ctx, hipMemcpyAsync(ctx->d_staging, ctx->h_staging,
                                    size_t(cnt) * ctx->brickBytes, hipMemcpyHostToDevice,
                                    ctx->uploadStream));
        LVR_HIP(ctx, hipMemcpyAsync(ctx->d_stagingSlots, slots, sizeof(int32_t) * cnt,
                                    hipMemcpyHostToDevice, ctx->uploadStream));
        rc = lvr::convertLayout(ctx, cnt, true, ctx->uploadStream);
        if (rc)
            return rc;
        // each brick's histogram into the cache as it lands in its slot
        rc = lvr::cacheBrickHistograms(ctx, node_ids + base, slots, cnt, ctx->uploadStream);
        if (rc)
            return rc;
    }
    LVR_HIP(ctx, hipStreamSynchronize(ctx->uploadStream));
    return LVR_OK;
}

int lvr_brick_evict(lvr_ctx* ctx, uint64_t node_id)
{
    if (!ctx)
        return LVR_EINVAL;
    auto it = ctx->resident.find(node_id);
    if (it == ctx->resident.end())
        return lvr::fail(ctx, LVR_ENOENT, "brick not resident");
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    // kernels already queued may still read the slot
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    rc = lvr::waitPoolReaders(ctx, nullptr);
    if (rc)
        return rc;
    if (it->second.pinned)
        return lvr::fail(ctx, LVR_EINVAL, "brick is pinned");
    ++ctx->poolGeneration;
    ctx->freeSlots.push_back(it->second.slot);
    uint64_t idx;
    if (lvr::pageIndex(ctx, node_id, &idx))
        ctx->lru.erase(int32_t(idx));
    ctx->resident.erase(it);
    rc = lvr::setPage(ctx, node_id, -1);
    if (rc)
        return rc;
    LVR_HIP(ctx, hipStreamSynchronize(ctx->uploadStream));
    return LVR_OK;
}

int lvr_brick_resident(const lvr_ctx* ctx, uint64_t node_id)
{
    if (!ctx)
        return LVR_EINVAL;
    return ctx->resident.count(node_id) ? 1 : 0;
}

int lvr_brick_download(lvr_ctx* ctx, uint64_t node_id, void* out, size_t bytes)
{
    if (!ctx || !out || bytes != ctx->brickBytes)
        return LVR_EINVAL;
    const int32_t slot = lvr::slotOf(ctx, node_id);
    if (slot < 0)
        return lvr::fail(ctx, LVR_ENOENT, "brick not resident");
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    LVR_HIP(ctx, hipStreamSynchronize(ctx->uploadStream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    rc = lvr::ensureDeviceStaging(ctx, 1);
    if (rc)
        return rc;
    LVR_HIP(ctx, hipMemcpyAsync(ctx->d_stagingSlots, &slot, sizeof(int32_t),
                                hipMemcpyHostToDevice, ctx->stream));
    rc = lvr::convertLayout(ctx, 1, false, ctx->stream);
    if (rc)
        return rc;
    LVR_HIP(ctx, hipMemcpyAsync(out, ctx->d_staging, bytes, hipMemcpyDeviceToHost,
                                ctx->stream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LVR_OK;
}

static int tfUpload(lvr_ctx* ctx, const std::vector<float>& rgba, int n,
                    float range0, float range1)
{
    // the render kernel stages the whole table in LDS as float4 (128 KiB max)
    if (n > 8192)
        return lvr::fail(ctx, LVR_EINVAL, "transfer function above 8192 entries");
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    // on the device the table carries a copy of its first entry before it
    // and of its last after it (the layout the classification reads)
    std::vector<float> padded(size_t(n + 2) * 4);
    std::copy(rgba.begin(), rgba.begin() + 4, padded.begin());
    std::copy(rgba.begin(), rgba.begin() + size_t(n) * 4, padded.begin() + 4);
    std::copy(rgba.begin() + size_t(n - 1) * 4, rgba.begin() + size_t(n) * 4,
              padded.begin() + size_t(n + 1) * 4);
    if (n + 2 > ctx->tfCapacity)
    {
        LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        (void)hipFree(ctx->d_tf);
        ctx->d_tf = nullptr;
        LVR_HIP(ctx, hipMalloc((void**)&ctx->d_tf, padded.size() * sizeof(float)));
        ctx->tfCapacity = n + 2;
    }
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LVR_HIP(ctx, hipMemcpyAsync(ctx->d_tf, padded.data(), padded.size() * sizeof(float),
                                hipMemcpyHostToDevice, ctx->stream));
    LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->nTf = n;
    ctx->range0 = range0;
    ctx->range1 = range1;
    ctx->tfSet = true;
    ++ctx->tfGeneration;
    return LVR_OK;
}

int lvr_tf_set(lvr_ctx* ctx, const float* rgba, int n, float range0,
               float range1)
{
    if (!ctx || !rgba || n < 1)
        return LVR_EINVAL;
    if (!std::isfinite(range0) || !std::isfinite(range1))
        return lvr::fail(ctx, LVR_EINVAL, "data range not finite");
    std::vector<float> v(rgba, rgba + size_t(n) * 4);
    return tfUpload(ctx, v, n, range0, range1);
}

int lvr_tf_set_rgba8(lvr_ctx* ctx, const uint8_t* lut, int n, float range0,
                     float range1)
{
    if (!ctx || !lut || n < 1)
        return LVR_EINVAL;
    if (!std::isfinite(range0) || !std::isfinite(range1))
        return lvr::fail(ctx, LVR_EINVAL, "data range not finite");
    std::vector<float> v(size_t(n) * 4);
    for (size_t i = 0; i < v.size(); ++i)
        v[i] = float(lut[i]) / 255.0f;
    return tfUpload(ctx, v, n, range0, range1);
}
}
