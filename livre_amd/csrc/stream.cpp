// Out-of-core brick streaming and the coarser-LOD rendering set.
//
// The reference's asynchronous pipeline (RenderPipeline.cpp:160-225) loads
// missing bricks on an upload thread (DataUploadFilter.cpp:104-120) and
// renders, meanwhile, the nearest loaded ancestor of every visible brick
// that is not there yet (RenderingSetGeneratorFilter.cpp:39-108). Here:
//  - host worker threads call the DataSource (getData-style callback) into
//    pinned staging buffers, most wanted bricks first;
//  - the caller's thread moves finished host bricks to the device: it takes
//    pool slots (LRU eviction, never a pinned brick), issues
//    hipMemcpyAsync from the pinned buffers and the layout conversion on the
//    context's upload stream, and marks the bricks resident once that
//    stream has passed them (event), so transfers overlap the ray casting on
//    the render stream;
//  - the rendering set of a frame replaces every missing visible brick by its
//    nearest resident ancestor, as the reference does.
// A brick being loaded occupies a pinned slot but counts as missing.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "lvr_internal.h"

namespace
{
struct Upload
{
    hipEvent_t done = nullptr;
    std::vector<uint64_t> ids;
    std::vector<int> buffers; // pinned staging buffers the copies read
};

// Parent of a node (NodeId::getParent, NodeId.cpp:77-88).
uint64_t parentOf(uint64_t id)
{
    const lvr::NodeKey k = lvr::decodeNode(id);
    return lvr::encodeNode(k.level - 1, k.x >> 1, k.y >> 1, k.z >> 1, k.t);
}

// The rendering set rule with a residency predicate.
template <typename Resident>
int renderingSet(const uint64_t* visible, int n, Resident resident, uint64_t* out, int cap,
                 int* n_out, int* n_available)
{
    std::unordered_set<uint64_t> set;
    set.reserve(size_t(n) * 2);
    int available = 0;
    for (int i = 0; i < n; ++i)
    {
        // collectLoadedTextures: the node itself or its nearest loaded
        // ancestor (RenderingSetGeneratorFilter.cpp:51-69)
        uint64_t cur = visible[i];
        for (;;)
        {
            if (resident(cur))
            {
                set.insert(cur);
                break;
            }
            if (lvr::decodeNode(cur).level == 0)
                break;
            cur = parentOf(cur);
        }
        available += set.count(visible[i]) ? 1 : 0;
    }
    // entries with an ancestor in the set are dropped, only when the set's
    // size differs from the visible count (RenderingSetGeneratorFilter.cpp:
    // 81-96 skips the pass otherwise)
    std::vector<uint64_t> keep;
    keep.reserve(set.size());
    const bool prune = set.size() != size_t(n);
    for (uint64_t id : set)
    {
        bool covered = false;
        if (prune)
            for (uint64_t a = id; lvr::decodeNode(a).level > 0 && !covered;)
            {
                a = parentOf(a);
                covered = set.count(a) > 0;
            }
        if (!covered)
            keep.push_back(id);
    }
    // deterministic output order (the reference's is hash-map order; the
    // caller orders front to back)
    std::sort(keep.begin(), keep.end());
    if (n_out)
        *n_out = int(keep.size());
    if (n_available)
        *n_available = available;
    if (int(keep.size()) > cap)
        return LVR_ERANGE;
    std::copy(keep.begin(), keep.end(), out);
    return LVR_OK;
}
} // namespace

struct lvr_stream
{
    lvr_ctx* ctx = nullptr;
    lvr_brick_source_fn source = nullptr;
    void* user = nullptr;
    std::string error;

    // pinned staging buffers, one brick each
    std::vector<uint8_t*> buffers;
    std::vector<int> freeBuffers;

    // work shared with the source threads
    std::mutex m;
    std::condition_variable cv;
    std::deque<uint64_t> queue;                   // wanted, most wanted first
    std::unordered_set<uint64_t> generating;      // taken by a thread
    std::vector<std::pair<uint64_t, int>> ready;  // generated: (id, buffer)
    std::unordered_set<uint64_t> readyIds;        // ids in `ready`
    std::vector<std::pair<uint64_t, int>> failed; // source errors: (id, code)
    bool stop = false;
    std::vector<std::thread> threads;

    // device side (caller's thread only)
    std::unordered_set<uint64_t> loading; // slots taken, copies in flight
    // per octree node (page-table index): stamp of the last request that
    // saw it, and 1 while its copy is in flight
    std::vector<uint32_t> seen;
    std::vector<uint8_t> loadingFlag;
    uint32_t stamp = 0;
    std::deque<Upload> uploads;
    std::vector<hipEvent_t> spareEvents;
    int32_t* h_slots = nullptr; // pinned slot lists of the conversions
    size_t slotCap = 0;
    size_t slotHead = 0;
    lvr_stream_stats stats{};
};

namespace
{
int stfail(lvr_stream* s, int code, const std::string& msg)
{
    s->error = msg;
    return code;
}

void worker(lvr_stream* s)
{
    const size_t bytes = s->ctx->brickBytes;
    for (;;)
    {
        uint64_t id = 0;
        int buf = -1;
        {
            std::unique_lock<std::mutex> lk(s->m);
            s->cv.wait(lk, [&] {
                return s->stop || (!s->queue.empty() && !s->freeBuffers.empty());
            });
            if (s->stop)
                return;
            id = s->queue.front();
            s->queue.pop_front();
            buf = s->freeBuffers.back();
            s->freeBuffers.pop_back();
            s->generating.insert(id);
        }
        const int rc = s->source(s->user, id, s->buffers[size_t(buf)], bytes);
        {
            std::lock_guard<std::mutex> lk(s->m);
            s->generating.erase(id);
            if (rc == 0)
            {
                s->ready.emplace_back(id, buf);
                s->readyIds.insert(id);
            }
            else
            {
                s->failed.emplace_back(id, rc);
                s->freeBuffers.push_back(buf);
            }
        }
        s->cv.notify_all();
    }
}

// Bricks whose copies have completed become resident (unpinned).
int retireUploads(lvr_stream* s, bool block, int* completed)
{
    while (!s->uploads.empty())
    {
        Upload& u = s->uploads.front();
        const hipError_t q = block ? hipEventSynchronize(u.done) : hipEventQuery(u.done);
        if (q == hipErrorNotReady)
            break;
        if (q != hipSuccess)
            return stfail(s, LVR_EHIP, "upload event");
        for (uint64_t id : u.ids)
        {
            (void)lvr::pinBrick(s->ctx, id, -1);
            s->loading.erase(id);
            uint64_t idx;
            if (lvr::pageIndex(s->ctx, id, &idx))
                s->loadingFlag[idx] = 0;
        }
        s->stats.bricks_loaded += u.ids.size();
        s->stats.bytes_loaded += u.ids.size() * s->ctx->brickBytes;
        if (completed)
            *completed += int(u.ids.size());
        {
            std::lock_guard<std::mutex> lk(s->m);
            for (int b : u.buffers)
                s->freeBuffers.push_back(b);
        }
        s->cv.notify_all();
        s->spareEvents.push_back(u.done);
        s->uploads.pop_front();
    }
    return LVR_OK;
}

// Move generated bricks to the device: slots, async copies, conversion.
int issueUploads(lvr_stream* s)
{
    std::vector<std::pair<uint64_t, int>> ready;
    {
        std::lock_guard<std::mutex> lk(s->m);
        ready.swap(s->ready);
        s->readyIds.clear();
        s->stats.source_errors += s->failed.size();
        s->failed.clear();
    }
    if (ready.empty())
        return LVR_OK;
    lvr_ctx* ctx = s->ctx;
    const size_t chunk = ctx->stagingBricks;
    size_t pos = 0;
    std::vector<std::pair<uint64_t, int>> back; // no slot free: retry later
    while (pos < ready.size())
    {
        const size_t cnt = std::min(chunk, ready.size() - pos);
        Upload u;
        // the device staging area is reused in stream order; the slot list
        // lives in a pinned ring until its copy has run
        if (s->slotHead + cnt > s->slotCap)
        {
            // every earlier conversion must have read its slot list
            int rc = retireUploads(s, true, nullptr);
            if (rc)
                return rc;
            s->slotHead = 0;
        }
        int32_t* slots = s->h_slots + s->slotHead;
        size_t k = 0;
        for (size_t i = 0; i < cnt; ++i)
        {
            const uint64_t id = ready[pos + i].first;
            const int buf = ready[pos + i].second;
            if (ctx->resident.count(id) && !s->loading.count(id))
            {
                // already there (loaded meanwhile): drop the copy
                std::lock_guard<std::mutex> lk(s->m);
                s->freeBuffers.push_back(buf);
                continue;
            }
            if (lvr::evictableSlots(ctx) == 0)
            {
                back.push_back(ready[pos + i]);
                continue;
            }
            int32_t slot = -1;
            int rc = lvr::acquireSlotFor(ctx, id, &slot);
            if (rc)
                return stfail(s, rc, ctx->error);
            (void)lvr::pinBrick(ctx, id, 1);
            s->loading.insert(id);
            {
                uint64_t idx;
                if (lvr::pageIndex(ctx, id, &idx))
                    s->loadingFlag[idx] = 1;
            }
            slots[k] = slot;
            LVR_HIP(ctx, hipMemcpyAsync(ctx->d_staging + k * ctx->brickBytes,
                                        s->buffers[size_t(buf)], ctx->brickBytes,
                                        hipMemcpyHostToDevice, ctx->uploadStream));
            u.ids.push_back(id);
            u.buffers.push_back(buf);
            ++k;
        }
        pos += cnt;
        if (k == 0)
            continue;
        LVR_HIP(ctx, hipMemcpyAsync(ctx->d_stagingSlots, slots, sizeof(int32_t) * k,
                                    hipMemcpyHostToDevice, ctx->uploadStream));
        s->slotHead += k;
        int rc = lvr::convertLayout(ctx, int(k), true, ctx->uploadStream);
        if (rc)
            return stfail(s, rc, ctx->error);
        if (s->spareEvents.empty())
        {
            hipEvent_t e;
            LVR_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
            s->spareEvents.push_back(e);
        }
        u.done = s->spareEvents.back();
        s->spareEvents.pop_back();
        LVR_HIP(ctx, hipEventRecord(u.done, ctx->uploadStream));
        s->uploads.push_back(std::move(u));
        ++s->stats.transfers;
    }
    if (!back.empty())
    {
        std::lock_guard<std::mutex> lk(s->m);
        for (const auto& r : back)
        {
            s->ready.push_back(r);
            s->readyIds.insert(r.first);
        }
        s->stats.pool_full += back.size();
    }
    return LVR_OK;
}
} // namespace

extern "C" {

int lvr_noise_source(void* user, uint64_t node_id, void* out, size_t bytes)
{
    const lvr_noise_source_data* d = static_cast<const lvr_noise_source_data*>(user);
    return d ? lvr_noise_brick_host(&d->info, &d->params, node_id, out, bytes) : LVR_EINVAL;
}

int lvr_stream_create(lvr_ctx* ctx, lvr_brick_source_fn source, void* user, int threads,
                      int staging_bricks, lvr_stream** out)
{
    if (!ctx || !source || !out || threads < 1 || threads > 256 || staging_bricks < 1)
        return LVR_EINVAL;
    *out = nullptr;
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    lvr_stream* s = new lvr_stream();
    s->ctx = ctx;
    s->source = source;
    s->user = user;
    // one conversion moves up to 64 bricks (device staging of the context)
    rc = lvr::ensureDeviceStaging(ctx, 64);
    if (rc)
    {
        delete s;
        return rc;
    }
    s->buffers.resize(size_t(staging_bricks), nullptr);
    for (int i = 0; i < staging_bricks; ++i)
    {
        if (hipHostMalloc((void**)&s->buffers[size_t(i)], ctx->brickBytes,
                          hipHostMallocDefault) != hipSuccess)
        {
            lvr_stream_destroy(s);
            return LVR_ENOMEM;
        }
        s->freeBuffers.push_back(i);
    }
    s->seen.assign(size_t(ctx->pt.levelOffset.back()), 0u);
    s->loadingFlag.assign(size_t(ctx->pt.levelOffset.back()), uint8_t(0));
    s->slotCap = size_t(staging_bricks) + 64;
    if (hipHostMalloc((void**)&s->h_slots, sizeof(int32_t) * s->slotCap,
                      hipHostMallocDefault) != hipSuccess)
    {
        lvr_stream_destroy(s);
        return LVR_ENOMEM;
    }
    // the stream's own pins keep loading bricks in place; frames pin theirs
    ctx->pinsGuardReaders = true;
    for (int i = 0; i < threads; ++i)
        s->threads.emplace_back(worker, s);
    *out = s;
    return LVR_OK;
}

void lvr_stream_destroy(lvr_stream* s)
{
    if (!s)
        return;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cv.notify_all();
    for (std::thread& t : s->threads)
        t.join();
    (void)retireUploads(s, true, nullptr);
    for (uint64_t id : s->loading)
        (void)lvr::pinBrick(s->ctx, id, -1);
    for (hipEvent_t e : s->spareEvents)
        (void)hipEventDestroy(e);
    for (uint8_t* b : s->buffers)
        if (b)
            (void)hipHostFree(b);
    if (s->h_slots)
        (void)hipHostFree(s->h_slots);
    s->ctx->pinsGuardReaders = false;
    delete s;
}

const char* lvr_stream_last_error(const lvr_stream* s)
{
    return s ? s->error.c_str() : "null stream";
}

int lvr_stream_request_ex(lvr_stream* s, const uint64_t* ids, int n, int ancestors)
{
    if (!s || n < 0 || (n > 0 && !ids))
        return LVR_EINVAL;
    lvr_ctx* ctx = s->ctx;
    if (++s->stamp == 0) // wrapped: forget every old stamp
    {
        std::fill(s->seen.begin(), s->seen.end(), 0u);
        s->stamp = 1;
    }
    const uint32_t stamp = s->stamp;
    // wanted bricks bucketed by level (coarse first), each once
    std::vector<std::vector<uint64_t>> byLevel(ctx->vi.depth);
    for (int i = 0; i < n; ++i)
    {
        uint64_t a = ids[i];
        for (;;)
        {
            uint64_t idx;
            if (!lvr::pageIndex(ctx, a, &idx))
                return stfail(s, LVR_EINVAL, "node id outside the octree");
            if (s->seen[idx] == stamp)
                break; // this node and its ancestors are already handled
            s->seen[idx] = stamp;
            if (ctx->pt.h_slots[idx] >= 0)
            {
                if (!s->loadingFlag[idx])
                    lvr::touch(ctx, a); // wanted: last in line for eviction
            }
            else
                byLevel[lvr::decodeNode(a).level].push_back(a);
            if (!ancestors || lvr::decodeNode(a).level == 0)
                break;
            a = parentOf(a);
        }
    }
    {
        std::lock_guard<std::mutex> lk(s->m);
        // the newest wishes replace the queue; bricks being generated or
        // waiting for a slot finish
        s->queue.clear();
        for (const auto& level : byLevel)
            for (uint64_t id : level)
                if (!s->generating.count(id) && !s->readyIds.count(id))
                    s->queue.push_back(id);
        s->stats.requested += s->queue.size();
    }
    s->cv.notify_all();
    return LVR_OK;
}

int lvr_stream_request(lvr_stream* s, const uint64_t* ids, int n)
{
    return lvr_stream_request_ex(s, ids, n, 0);
}

int lvr_stream_pump(lvr_stream* s, int* completed)
{
    if (!s)
        return LVR_EINVAL;
    if (completed)
        *completed = 0;
    if (hipSetDevice(s->ctx->device) != hipSuccess)
        return stfail(s, LVR_EHIP, "hipSetDevice");
    int rc = retireUploads(s, false, completed);
    if (rc)
        return rc;
    return issueUploads(s);
}

int lvr_stream_wait(lvr_stream* s)
{
    if (!s)
        return LVR_EINVAL;
    for (;;)
    {
        int rc = lvr_stream_pump(s, nullptr);
        if (rc)
            return rc;
        bool idle, stuck;
        {
            std::unique_lock<std::mutex> lk(s->m);
            idle = s->queue.empty() && s->generating.empty() && s->ready.empty();
            // generated bricks that no slot can take while nothing is in
            // flight: the pool is full of pinned bricks
            stuck = !s->ready.empty() && s->uploads.empty();
            if (!idle && !stuck)
                s->cv.wait_for(lk, std::chrono::milliseconds(1));
        }
        if (idle)
            return retireUploads(s, true, nullptr);
        if (stuck)
            return stfail(s, LVR_ENOMEM, "brick pool full of pinned bricks");
        if (!s->uploads.empty())
        {
            rc = retireUploads(s, true, nullptr);
            if (rc)
                return rc;
        }
    }
}

int lvr_stream_rendering_set(lvr_stream* s, const uint64_t* visible, int n, uint64_t* out,
                             int cap, int* n_out, int* n_available)
{
    if (!s || n < 0 || (n > 0 && !visible) || cap < 0 || (cap > 0 && !out))
        return LVR_EINVAL;
    const lvr_ctx* ctx = s->ctx;
    const int rc = renderingSet(
        visible, n,
        [&](uint64_t id) {
            uint64_t idx;
            return lvr::pageIndex(ctx, id, &idx) && ctx->pt.h_slots[idx] >= 0 &&
                   !s->loadingFlag[idx];
        },
        out, cap, n_out, n_available);
    if (rc)
        return stfail(s, rc, "rendering set larger than the output buffer");
    return LVR_OK;
}

int lvr_rendering_set(const uint64_t* visible, int n, const uint64_t* resident, int n_resident,
                      uint64_t* out, int cap, int* n_out, int* n_available)
{
    if (n < 0 || (n > 0 && !visible) || n_resident < 0 || (n_resident > 0 && !resident) ||
        cap < 0 || (cap > 0 && !out))
        return LVR_EINVAL;
    std::unordered_set<uint64_t> res(resident, resident + n_resident);
    return renderingSet(
        visible, n, [&](uint64_t id) { return res.count(id) > 0; }, out, cap, n_out,
        n_available);
}

int lvr_stream_get_stats(lvr_stream* s, lvr_stream_stats* out)
{
    if (!s || !out)
        return LVR_EINVAL;
    *out = s->stats;
    std::lock_guard<std::mutex> lk(s->m);
    out->queued = s->queue.size();
    out->generating = s->generating.size();
    out->ready = s->ready.size();
    out->in_flight = s->loading.size();
    return LVR_OK;
}

int lvr_brick_pin(lvr_ctx* ctx, const uint64_t* ids, int n, int delta)
{
    if (!ctx || n < 0 || (n > 0 && !ids) || (delta != 1 && delta != -1))
        return LVR_EINVAL;
    for (int i = 0; i < n; ++i)
    {
        const int rc = lvr::pinBrick(ctx, ids[i], delta);
        if (rc)
        {
            for (int j = 0; j < i; ++j) // undo
                (void)lvr::pinBrick(ctx, ids[j], -delta);
            return rc;
        }
    }
    return LVR_OK;
}
}
