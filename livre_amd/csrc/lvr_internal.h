// Internal declarations shared by the translation units of liblivre_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <list>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/livre_amd.h"
#include "lvr_layout.h"

namespace lvr
{
// ------------------------------------------------------------------ NodeId
// 64-bit brick key with GCC's bitfield layout from the LSB: level 4 bits,
// x, y, z 14 bits each, time step 18 bits (livre/data/NodeId.h:34-48,
// livre/data/types.h:80-93).
struct NodeKey
{
    uint32_t level, x, y, z, t;
};

__host__ __device__ inline NodeKey decodeNode(uint64_t id)
{
    NodeKey k;
    k.level = uint32_t(id & 0xFu);
    k.x = uint32_t((id >> 4) & 0x3FFFu);
    k.y = uint32_t((id >> 18) & 0x3FFFu);
    k.z = uint32_t((id >> 32) & 0x3FFFu);
    k.t = uint32_t(id >> 46);
    return k;
}

__host__ __device__ inline uint64_t encodeNode(uint32_t level, uint32_t x,
                                               uint32_t y, uint32_t z,
                                               uint32_t t)
{
    return uint64_t(level & 0xFu) | (uint64_t(x & 0x3FFFu) << 4) |
           (uint64_t(y & 0x3FFFu) << 18) | (uint64_t(z & 0x3FFFu) << 32) |
           (uint64_t(t & 0x3FFFFu) << 46);
}

size_t bytesPerVoxel(int dataType);

// World box of a node, as internalNodeToLODNode computes it
// (livre/data/DataSourcePlugin.cpp:40-71), in FP32 with the same order of
// operations. Usable on host and device.
__host__ __device__ inline void nodeWorldBox(const float* worldSize, const uint32_t* rootBlocks,
                                             const NodeKey& k, float* mn, float* mx)
{
    const uint32_t scale = 1u << k.level;
    const uint32_t bx = rootBlocks[0] * scale;
    const uint32_t by = rootBlocks[1] * scale;
    const uint32_t bz = rootBlocks[2] * scale;
    uint32_t bmax = bx;
    if (by > bmax)
        bmax = by;
    if (bz > bmax)
        bmax = bz;
    const uint32_t pos[3] = {k.x, k.y, k.z};
    const float denom = float(bmax);
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const float cmin = float(pos[i]) / denom;
        const float cmax = float(pos[i] + 1u) / denom;
        const float half = worldSize[i] * 0.5f;
        mn[i] = cmin * worldSize[i] - half;
        mx[i] = cmax * worldSize[i] - half;
    }
}

__host__ __device__ inline void nodeWorldBox(const lvr_volume_info& vi, const NodeKey& k,
                                             float* mn, float* mx)
{
    nodeWorldBox(vi.world_size, vi.root_blocks, k, mn, mx);
}

// What the host sends per brick of a frame: the NodeId and its page-table
// index (-1 for an id outside the octree). 16 bytes.
struct BrickKey
{
    uint64_t id;
    int32_t page;
    uint32_t pad;
};

// Per-brick record of one frame, in the order of the ordered list, built on
// the device from the BrickKey (brickRecordsKernel) when the frame's GPU work
// starts; the slot comes from the device page table at that point, so a
// brick evicted after the host prepared the frame (or a batch replayed after
// an eviction) reads -1 and is skipped. 64 bytes.
struct BrickEntry
{
    float bmin[3];
    float bmax[3];
    float ext[3];   // bmax - bmin, rounded as the shader's (aabbMax - aabbMin)
    // 1 / ext, correctly rounded; with kBrickPow2 set, (1 / ext) * texExt,
    // exact because ext is a power of two
    float rcp[3];
    int32_t slot;   // pool slot from the device page table, -1 if not resident
    uint32_t level;
    uint32_t flags;
    uint32_t pad;
};

// BrickEntry::flags: every extent is a power of two, so the texture position
// (pos - aabbMin) / ext * texExt is one multiply by rcp, rounded once, exactly
// as the division and the multiply are.
constexpr uint32_t kBrickPow2 = 1u;

// Dense page table over the octree: per level,
// levelOffset[L] + x + nx * (y + ny * z) holds a slot index or -1. The host
// copy is authoritative; entries changed since the last frame are listed in
// `dirty` and written into the device copy by a scatter kernel on the render
// stream before the next launch that reads it (applyPageTable).
struct PageTable
{
    std::vector<uint64_t> levelOffset; // depth + 1 entries
    std::vector<uint32_t> levelDim;    // 3 per level
    int32_t* d_slots = nullptr;
    std::vector<int32_t> h_slots;
    uint64_t size = 0;
    std::vector<uint32_t> dirty;   // page indices changed on the host
    std::vector<uint8_t> isDirty;  // per page: listed in `dirty`
    // fixed staging of (page, slot) pairs, allocated with the context
    static constexpr int kUpdateChunk = 32768;
    int2* h_update = nullptr;  // pinned
    int2* d_update = nullptr;
    hipEvent_t updateDone = nullptr; // last scatter that read h_update
    bool updateArmed = false;
};

struct PoolEntry
{
    int32_t slot;
    int pinned;
};

// Eviction order of the resident bricks (front = next to evict) as a doubly
// linked list over page indices, in arrays sized by the octree's node count:
// moving a brick to the back rewrites three array entries, where a heap list
// would chase nodes scattered over memory for every rendered brick.
struct DenseLru
{
    std::vector<int32_t> prev, next; // -1 = none
    std::vector<uint64_t> ids;       // node id of each listed page index
    int32_t head = -1, tail = -1;

    void init(size_t pages)
    {
        prev.assign(pages, -1);
        next.assign(pages, -1);
        ids.assign(pages, 0);
        head = tail = -1;
    }
    void pushBack(int32_t i, uint64_t id)
    {
        ids[size_t(i)] = id;
        prev[size_t(i)] = tail;
        next[size_t(i)] = -1;
        if (tail >= 0)
            next[size_t(tail)] = i;
        else
            head = i;
        tail = i;
    }
    void erase(int32_t i)
    {
        const int32_t p = prev[size_t(i)], n = next[size_t(i)];
        if (p >= 0)
            next[size_t(p)] = n;
        else
            head = n;
        if (n >= 0)
            prev[size_t(n)] = p;
        else
            tail = p;
        prev[size_t(i)] = next[size_t(i)] = -1;
    }
    void moveToBack(int32_t i)
    {
        if (i == tail)
            return;
        const uint64_t id = ids[size_t(i)];
        erase(i);
        pushBack(i, id);
    }
};

// Everything one frame (one request of a batch) needs on the device.
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
    // host-computed per-frame constants (scalar operands in the kernels):
    // texMax - texMin, float(texDim), float(texDim - 1), 1 / nSamples,
    // float(nTf), maxSamples / nSamples
    float texExt[3], texDimF[3], texLastF[3];
    float stepSize, nTfF, alphaCorrection;
    int32_t nSamples, maxSamples;
    float exitA;
    float mult, add;
    int32_t nTf;
    int32_t tileW, tileH, tilesX, tilesY;
    int32_t nBricks;
    int32_t linear;
    int32_t depth; // voxel fetches in flight per ray (NEAREST march)
    int32_t classify; // 1 tabulated, 2 per sample (8/16-bit NEAREST)
    int32_t wavesPerSimd; // resident wavefronts per SIMD (single-frame launch)
    int32_t nClip;
    float clip[24];
    uint64_t slotStride;
    BrickLayout layout; // tiled voxel order inside a pool slot
    // byte offset of voxel (x, y, z) in a slot, as
    //   bpv * (x + 4 y + 16 z) + tileK[0] * (x >> 2) + tileK[1] * (y >> 2)
    //   + tileK[2] * (z >> 2)
    // with tileK = bpv * {60, 64 T0 - 16, 64 T0 T1 - 64} (T = tiles per axis):
    // the 4x4x4 micro-tile index times 64 plus the voxel inside the tile
    uint32_t tileK[3];
    // position of this frame in a batch
    int32_t brickBase; // first brick record / tile range
    int32_t tileBase;  // first tile of the per-tile arrays
    int32_t superBase; // first super-tile (binning) of the per-super-tile arrays
    int32_t superX, superY; // super-tiles of kSuperTile x kSuperTile tiles
    int32_t itemBase;  // first work item
    int32_t nItems;    // work items (16 per 4x4 super-tile)
    int32_t dataType;
    int32_t request;   // index in the batch
    // volume shape for the device-built brick records
    float worldSize[3];
    uint32_t rootBlocks[3];
    const uint8_t* pool;
    const int32_t* pageTable; // device page table of the volume: slot or -1
    const float4* tf;  // nTf + 2 RGBA entries: tf[0], the nTf entries, tf[nTf - 1]
    const float4* cls; // classified table per raw value (integer NEAREST), or null
    uint32_t* frame;   // RGBA8 output, row 0 = bottom
    float4* layer;     // or, when set, the unconverted RGBA32F accumulation
                       // (a sort-last partial), row 0 = bottom
    // the order in which the work queue hands out the kSuper x kSuper
    // super-tiles of work items, or null for screen order
    const int32_t* superOrder;
};

// Pinned staging of the per-frame host records (brick entries, screen
// ranges, request tables): hipMemcpyAsync from pinned memory is a plain DMA
// that returns at once, where a pageable source is first staged by the host
// in chunks. Two buffers are used in turn, and one is rewritten only after
// the copies that read it have completed.
struct RawPyramid;

struct PinnedRing
{
    uint8_t* buf[2] = {nullptr, nullptr};
    size_t cap[2] = {0, 0};
    hipEvent_t done[2] = {nullptr, nullptr};
    bool armed[2] = {false, false};
    int next = 0;
    // outgrown buffers: hipHostFree waits for the whole device, so they are
    // kept until destroy()
    std::vector<uint8_t*> retired;

    // a buffer of at least `bytes` whose previous copies have completed;
    // *slot names it for release()
    hipError_t acquire(size_t bytes, uint8_t** out, int* slot);
    // the copies reading buffer `slot` are queued on st
    hipError_t release(int slot, hipStream_t st);
    void destroy();
};
} // namespace lvr

struct lvr_ctx
{
    int device = 0;
    int numCUs = 256;
    hipStream_t stream = nullptr;
    hipStream_t uploadStream = nullptr;
    // histogram kernels read the pool beside the ray casting of the frame
    hipStream_t histStream = nullptr;
    // selection reads only its own buffers, so it runs on a stream of its own
    // and does not wait for frames still rendering on `stream`
    hipStream_t selStream = nullptr;
    hipEvent_t histDone = nullptr; // end of the last histogram on histStream
    hipEvent_t renderDone = nullptr; // recorded on `stream` before slot reuse
    lvr_volume_info vi{};
    size_t voxelBytes = 0; // bytes per voxel
    size_t brickBytes = 0; // maximumBlockSize^3 * bytesPerVoxel
    size_t slotStride = 0; // tiled brick bytes rounded to 256
    lvr::BrickLayout layout{};
    uint64_t nSlots = 0;
    uint8_t* d_pool = nullptr;
    lvr::PageTable pt;
    // recorded after the last page-table scatter, on the stream it ran on:
    // record passes on other streams wait for it before they read the table
    hipEvent_t tableReady = nullptr;
    bool tableRecorded = false;
    std::vector<int32_t> freeSlots;
    std::unordered_map<uint64_t, lvr::PoolEntry> resident;
    lvr::DenseLru lru; // over page indices
    int policy = 0;

    // pinned staging for uploads, device staging for layout conversion
    uint8_t* h_staging = nullptr;
    size_t stagingBytes = 0;
    uint8_t* d_staging = nullptr;
    int32_t* d_stagingSlots = nullptr;
    size_t stagingBricks = 0;

    // transfer function
    float* d_tf = nullptr; // float4 entries
    int nTf = 0;
    int tfCapacity = 0; // entries of d_tf (the table and its two edge copies)
    uint64_t tfGeneration = 0; // bumped by every transfer-function upload
    // classified table (r*a, g*a, b*a, a) per raw value of 8/16-bit volumes
    // for the current TF, data range and sampling rate
    float4* d_cls = nullptr;
    uint64_t clsKey[4] = {~0ull, 0, 0, 0};
    float range0 = 0.f, range1 = 0.f;
    bool tfSet = false;

    // per-frame buffers
    uint32_t* d_frame = nullptr;
    int fbW = 0, fbH = 0;
    size_t fbCapacity = 0;
    // the current frame's records (aliases into recSets)
    lvr::BrickEntry* d_bricks = nullptr;
    int4* d_ranges = nullptr; // conservative tile range per brick
    // the current frame's tile binning (aliases into recSets): ordered brick
    // lists per tile and, the first level, per super-tile
    int32_t* d_tileCount = nullptr;
    int32_t* d_tileOffset = nullptr;
    int32_t* d_tileList = nullptr;
    int32_t* d_superCount = nullptr;
    int32_t* d_superOffset = nullptr;
    int32_t* d_superList = nullptr;
    unsigned long long* curCounters = nullptr; // the current frame's counters
    lvr::FrameUniforms* d_req = nullptr; // the frame as a one-request table
    // Two sets of frame records and binning in turn: frame t+1's records
    // are copied and binned on recStream while frame t ray-casts (the
    // binning kernels fit beside the persistent ray cast on each CU); a
    // set's copies wait until the frame that last used it is finished, and
    // the frame's ray cast waits for its binning.
    struct RecordSet
    {
        lvr::FrameUniforms* req = nullptr;
        size_t reqCap = 0;
        lvr::BrickKey* keys = nullptr;
        size_t keyCap = 0;
        lvr::BrickEntry* bricks = nullptr;
        size_t brickCap = 0;
        int4* ranges = nullptr;
        size_t rangeCap = 0;
        int32_t *tileCount = nullptr, *tileOffset = nullptr, *tileList = nullptr;
        size_t tileCountCap = 0, tileOffsetCap = 0, tileListCap = 0;
        int32_t *superCount = nullptr, *superOffset = nullptr, *superList = nullptr;
        size_t superCountCap = 0, superOffsetCap = 0, superListCap = 0;
        int32_t* superOrder = nullptr; // work super-tiles: the heaviest first
        size_t superOrderCap = 0;
        unsigned long long* counters = nullptr; // [0] samples, [1] work queue
        size_t countersCap = 0;
        hipEvent_t ready = nullptr; // binned (recStream)
        hipEvent_t free = nullptr;  // its frame finished, counters read (stream)
        bool used = false;
    };
    RecordSet recSets[2];
    int recNext = 0;
    hipStream_t recStream = nullptr;
    std::vector<lvr::BrickKey> hostBricks;
    std::vector<int4> hostRanges;
    lvr::PinnedRing frameStaging; // the frame's records on their way to d_*
    // lvr_render_noalloc: device address of the library's static scratch
    // array on this context's device (null if it could not be taken), the
    // fixed pinned staging of one frame's uniforms, keys and ranges, and the
    // event of the last copy that read that staging
    uint8_t* d_scratch = nullptr;
    uint8_t* h_noalloc = nullptr;
    hipEvent_t noallocCopied = nullptr;
    bool noallocArmed = false;
    lvr::RawPyramid* pyramid = nullptr; // levels of the last raw source loaded
    lvr_jpeg_encoder* jpeg = nullptr;   // frame grab encoder on `stream`
    // sort-last partial images (RGBA32F), grow-only
    static constexpr int kMaxLayers = 16;
    float4* d_layers[kMaxLayers] = {};
    size_t layerCap[kMaxLayers] = {};
    // viewport of each layer's last partial render (0 x 0: never rendered)
    int layerW[kMaxLayers] = {}, layerH[kMaxLayers] = {};
    hipEvent_t compositeEv[2] = {nullptr, nullptr};
    // device buffers replaced while queued frames may still read them, and
    // outgrown pinned buffers (hipHostFree would wait for the whole device);
    // freed where every stream has drained
    std::vector<void*> retired;
    std::vector<void*> retiredHost;
    // selection result (counters, then ids), written by the kernel into
    // pinned host memory; one selection in flight
    uint8_t* h_sel = nullptr;
    hipEvent_t selDone = nullptr;
    bool selPending = false;
    float selRange[2] = {0.f, 1.f}; // sort-last share of the pending selection
    unsigned long long* d_counters = nullptr; // [0] samples, [1] work queue
    unsigned long long* h_counters = nullptr; // pinned

    // histogram
    unsigned long long* d_hist = nullptr;
    int histBins = 0; // allocated bins
    int histLast = 0; // bins of the last histogram
    // job lists of the histogram kernels: two pinned, device-mapped buffers
    // used in turn, each with the event of the last kernel that read it
    uint8_t* h_histRing[2] = {nullptr, nullptr};
    size_t histRingBytes[2] = {0, 0};
    hipEvent_t histRingDone[2] = {nullptr, nullptr};
    int histRingNext = 0;
    // running histogram of the last set (d_hist holds its sum)
    std::vector<uint64_t> histSet;   // sorted resident ids
    bool histValid = false;
    uint64_t histGeneration = 0;
    float histRange0 = 0.f, histRange1 = 0.f;
    uint64_t histDeltaBricks = 0;
    // per-slot histogram cache (HistogramObject per brick): u32 counts of
    // the interior voxels of the brick in the slot, unscaled; histCachedId[s]
    // is the brick whose counts slot s holds (~0: none); histSetSlots gives
    // the slot each brick of histSet was counted from
    uint32_t* d_histCache = nullptr;
    int histCacheBins = 0;
    bool histCacheTried = false;
    std::vector<uint64_t> histCachedId;
    std::vector<uint32_t> histRowVersion; // bumped whenever a row is (re)binned
    std::vector<int32_t> histSetSlots;
    std::vector<uint32_t> histSetVersions; // row version each was counted with
    float histCacheR0 = 0.f, histCacheR1 = 0.f;
    uint64_t histBinned = 0; // bricks binned into the cache by the last call
    // bumped whenever brick contents change (upload, generation, eviction)
    uint64_t poolGeneration = 0;

    // selection
    uint64_t* d_selA = nullptr;
    uint64_t* d_selB = nullptr;
    uint64_t* d_selOut = nullptr;
    int32_t* d_selCount = nullptr; // [0] next frontier, [1] visible
    size_t selCapacity = 0;

    // statistics of the last kFrameSlots frames, so a caller that keeps
    // frames in flight reads each one's after the next has been submitted
    static constexpr int kFrameSlots = 4;
    struct FrameRecord
    {
        uint64_t bricks = 0, entries = 0;
        float hostMs = 0.f; // host time of the render call
        bool timed = false, histTimed = false, armed = false;
        // binning start (recStream), ray-cast start, ray-cast end (on
        // `stream`), histogram start and end (on `histStream`), binning end
        // (recStream)
        hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        hipEvent_t done = nullptr; // the frame's composite steps are in h_frameSamples
    };
    FrameRecord frames[kFrameSlots];
    uint64_t* h_frameSamples = nullptr; // pinned, one per slot
    uint64_t frameCount = 0;            // frames rendered
    bool timing = false;

    // batches launched on their own streams that read this pool; slot reuse
    // waits for them unless the owner pins every brick a launch reads
    std::vector<lvr_batch*> readers;
    bool pinsGuardReaders = false;
    uint64_t pinnedBricks = 0; // resident bricks with a pin count above 0

    // device scratch of the procedural sources (grow-only)
    void* d_gen = nullptr;
    size_t genCapacity = 0;

    std::string error;
};

namespace lvr
{
int fail(lvr_ctx* ctx, int code, const std::string& msg);
int hipFail(lvr_ctx* ctx, hipError_t e, const char* what);
int ensureDevice(lvr_ctx* ctx);
int32_t slotOf(const lvr_ctx* ctx, uint64_t id);
// Dense index of a node in the page table (false outside the octree).
bool pageIndex(const lvr_ctx* ctx, uint64_t id, uint64_t* idx);
void touch(lvr_ctx* ctx, uint64_t id);
// slotOf + touch with one lookup
int32_t slotAndTouch(lvr_ctx* ctx, uint64_t id);
// Device levels of a raw source's LOD pyramid (raw.hip), freed with the
// context or by lvr_pyramid_release.
void destroyPyramid(lvr_ctx* ctx);
// The frame grab's JPEG encoder (jpeg.hip), freed with the context.
void destroyJpeg(lvr_ctx* ctx);
// Write the page-table entries changed on the host into the device table on
// stream `st`, and make `st` wait for the last such update wherever it ran:
// kernels queued on `st` afterwards read the current table. Only the brick
// record pass reads the table, so updates never race a ray cast.
int applyPageTable(lvr_ctx* ctx, hipStream_t st);
int acquireSlotFor(lvr_ctx* ctx, uint64_t id, int32_t* slot);
// Undo acquireSlotFor for a brick whose contents were never written.
int releaseSlot(lvr_ctx* ctx, uint64_t id);
// Change a resident brick's pin count (pinned bricks are never evicted).
int pinBrick(lvr_ctx* ctx, uint64_t id, int delta);
// Slots that can take a new brick without touching pinned ones.
uint64_t evictableSlots(const lvr_ctx* ctx);
// Order slot reuse after every batch still reading the pool: stream `st`
// waits for them, or the host does when st is null.
int waitPoolReaders(lvr_ctx* ctx, hipStream_t st);
// HistogramObject per brick, computed as the bricks land in their slots:
// bin the bricks ids[0..n) just written into slots[0..n) (a host array) into
// the per-slot histogram cache, on stream st after the writes. A no-op where
// the cache is not used (float data, rows above half the pool's bytes).
int cacheBrickHistograms(lvr_ctx* ctx, const uint64_t* ids, const int32_t* slots, int n,
                         hipStream_t st);
// Device scratch of at least `bytes` owned by the context (nullptr + error).
void* genScratch(lvr_ctx* ctx, size_t bytes);
// lvr_render_noalloc set-up, called by lvr_ctx_create with the context's
// device current: the device address of the static scratch array (taken
// once per device) and the context's fixed pinned staging.
int noallocSetup(lvr_ctx* ctx);
// Batch internals used by the pool and the server.
hipEvent_t batchDoneEvent(lvr_batch* b);
void batchForgetCtx(lvr_batch* b, lvr_ctx* ctx);
hipStream_t batchStream(lvr_batch* b);
uint32_t* batchFrame(lvr_batch* b, int request);
int ensureDeviceStaging(lvr_ctx* ctx, size_t bricks);
// d_staging (linear bricks) <-> pool slots d_stagingSlots, on stream st
int convertLayout(lvr_ctx* ctx, int nBricks, bool toTiled, hipStream_t st);
} // namespace lvr

#define LVR_HIP(ctx, call)                                                     \
    do                                                                         \
    {                                                                          \
        hipError_t _e = (call);                                                \
        if (_e != hipSuccess)                                                  \
            return lvr::hipFail((ctx), _e, #call);                             \
    } while (0)
