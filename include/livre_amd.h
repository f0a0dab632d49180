/*
 * livre_amd.h — C ABI of the MI355X (gfx950) render hot path for Livre.
 *
 * The library replaces the GL ray caster behind livre::Renderer
 * (livre/core/render/Renderer.h:38-109, livre/eq/render/RayCastRenderer.h:31-74)
 * and the GPU brick cache behind TexturePool/TextureObject
 * (livre/core/render/TexturePool.h:35-80, livre/lib/cache/TextureObject.cpp:76-120).
 * Bricks are addressed by Livre's 64-bit NodeId (livre/data/NodeId.h:34-48).
 *
 * Conventions
 *  - Every entry point returns 0 on success and a negative LVR_E* code on
 *    failure; lvr_last_error() gives the message. No C++ exception crosses
 *    the ABI (the reference throws std::runtime_error and logs per-brick
 *    errors, RayCastRenderer.cpp:117-138, 517-522).
 *  - Matrices are column-major float[16] (OpenGL / vmmlib layout).
 *  - Bricks hold (block + 2*overlap)^3 voxels, x fastest, exactly what
 *    DataSourcePlugin::getData returns (DataSourcePlugin.h:81) and what
 *    glTexSubImage3D consumes (TextureObject.cpp:107-109).
 *  - Images are RGBA8, 4 bytes per pixel, row 0 = bottom row (GL window
 *    coordinates, what Equalizer reads back).
 *  - A context belongs to one device and is not thread-safe: the host
 *    serialises calls on it (the reference renders on one pipe thread).
 */
#ifndef LIVRE_AMD_H
#define LIVRE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define LVR_OK 0
#define LVR_EINVAL -1  /* invalid argument */
#define LVR_EHIP -2    /* HIP runtime failure */
#define LVR_ENOMEM -3  /* brick pool or device memory exhausted */
#define LVR_ENOENT -4  /* brick not resident */
#define LVR_ERANGE -5  /* output buffer too small */

/* livre::DataType (livre/data/VolumeInformation.h:29-39) */
#define LVR_DT_FLOAT 0
#define LVR_DT_UINT8 1
#define LVR_DT_UINT16 2
#define LVR_DT_UINT32 3
#define LVR_DT_INT8 4
#define LVR_DT_INT16 5
#define LVR_DT_INT32 6

typedef struct lvr_ctx lvr_ctx;

/* livre::VolumeInformation (livre/data/VolumeInformation.h:42-111), the
 * fields the render path reads. The caller fills voxels, max_block, overlap
 * and data_type; lvr_fill_regular_volume_info computes the rest. */
typedef struct
{
    uint32_t voxels[3];
    uint32_t max_block[3]; /* maximumBlockSize = block + 2 * overlap */
    uint32_t overlap[3];
    int32_t data_type;
    float world_size[3];
    float world_space_per_voxel;
    uint32_t depth; /* RootNode::getDepth(): number of octree levels */
    uint32_t root_blocks[3];
} lvr_volume_info;

/* Replaces fillRegularVolumeInfo (livre/data/DataSourcePlugin.cpp:73-95). */
int lvr_fill_regular_volume_info(lvr_volume_info* info);

/* Replaces DataSourcePlugin::internalNodeToLODNode's world box
 * (livre/data/DataSourcePlugin.cpp:40-71): box = {min xyz, max xyz}. */
int lvr_node_world_box(const lvr_volume_info* info, uint64_t node_id,
                       float box[6]);

/* livre::Frustum (livre/data/Frustum.h:37-111): the quantities the render
 * path uses, computed from modelview and projection. */
typedef struct
{
    float mv[16], proj[16];
    float inv_mv[16], inv_proj[16];
    float mvp[16];
    float eye[3];       /* Frustum::getEyePos */
    float view_dir[3];  /* Frustum::getViewDir */
    float near_plane, far_plane, left, right, bottom, top;
    float planes[6][4]; /* normalized culling planes: l, r, b, t, near, far */
} lvr_frustum;

/* Replaces Frustum::Frustum(mv, proj) (livre/data/Frustum.cpp:24-42). */
int lvr_frustum_init(lvr_frustum* f, const float mv[16], const float proj[16]);

/* ---------------------------------------------------------------- context */
/* One context per volume and device: owns the HBM brick pool of
 * pool_bytes, its device page table and the render scratch buffers.
 * Replaces TexturePool(dataSource) + the texture Cache of gpu-cache-mem
 * bytes (TexturePool.cpp:33-84, Window.cpp:54-72). */
int lvr_ctx_create(const lvr_volume_info* info, size_t pool_bytes, int device,
                   lvr_ctx** out);
void lvr_ctx_destroy(lvr_ctx* ctx);
const char* lvr_last_error(const lvr_ctx* ctx);
/* Synchronise the context's stream. */
int lvr_ctx_sync(lvr_ctx* ctx);
/* 1 when the work queued on the context's render stream has finished, 0
 * while it runs (never blocks). */
int lvr_ctx_query(lvr_ctx* ctx);

/* HBM bytes one brick occupies in the pool (the pool stores bricks in its own
 * tiled layout, padded, 256-byte aligned): pool_bytes = n * this holds n
 * bricks. Replaces the texture size accounting of TextureObject.cpp:36-41. */
int lvr_slot_bytes(const lvr_volume_info* info, uint64_t* out);

/* Cache policy of the pool: 0 = LRU (touched on render, the default),
 * 1 = load order (the reference LRUCachePolicy, which never refreshes on
 * get, Cache.cpp:41-96, 158-170). Pinned bricks are never evicted. */
int lvr_pool_set_policy(lvr_ctx* ctx, int policy);
/* slots: capacity in bricks; used: resident bricks. */
int lvr_pool_stats(const lvr_ctx* ctx, uint64_t* slots, uint64_t* used,
                   uint64_t* slot_bytes);

/* --------------------------------------------------------------- bricks */
/* Upload one brick of getData() bytes into a pool slot, evicting by the
 * policy when full. Replaces TextureObject::Impl::loadTextureToGPU
 * (TextureObject.cpp:94-120) and TexturePool::generate (TexturePool.cpp:90-119).
 * bytes must equal maximumBlockSize^3 * bytesPerVoxel. Synchronous with
 * respect to the host buffer. */
int lvr_brick_upload(lvr_ctx* ctx, uint64_t node_id, const void* voxels,
                     size_t bytes);
/* Same, for n bricks stored back to back in one host buffer; the copy is
 * staged through pinned memory on the upload stream. */
int lvr_brick_upload_many(lvr_ctx* ctx, const uint64_t* node_ids, int n,
                          const void* voxels, size_t bytes_per_brick);
/* Release a slot (TexturePool::release, TexturePool.cpp:121-126). */
int lvr_brick_evict(lvr_ctx* ctx, uint64_t node_id);
/* 1 if resident, 0 if not, negative on error. */
int lvr_brick_resident(const lvr_ctx* ctx, uint64_t node_id);
/* Copy a resident brick back to the host (for checks). */
int lvr_brick_download(lvr_ctx* ctx, uint64_t node_id, void* out,
                       size_t bytes);

/* ------------------------------------------------------- procedural data */
/* Procedural DataSource plugins for synthetic volumes: integer value noise
 * (5 octaves by default) remapped through a 4096-entry uint16 table. Every
 * voxel is a pure function of (seed, level, voxel coordinates), so a brick
 * computed on the host equals the one computed on the device bit for bit.
 * There is no reference counterpart: the reference's MemoryDataSource fills
 * bricks with one constant (MemoryDataSource.cpp:38-61). */
typedef struct
{
    uint64_t seed;
    int32_t octaves;    /* number of octaves, 1..8 */
    int32_t base_cells; /* lattice cells across the largest axis, octave 0 */
    uint16_t remap[4097]; /* monotone table over the octave sum, 4096 segments */
} lvr_noise_params;

/* Fill the remap table: log-normal y = exp(gain * z) of the standardised
 * octave sum, mapped linearly to the full range of the volume's type. */
int lvr_noise_params_init(lvr_noise_params* p, uint64_t seed, int octaves,
                          int base_cells, float gain);
/* Host generation of one brick (DataSourcePlugin::getData). */
int lvr_noise_brick_host(const lvr_volume_info* info,
                         const lvr_noise_params* p, uint64_t node_id,
                         void* out, size_t bytes);
/* Device generation straight into pool slots (no PCIe traffic). */
int lvr_noise_bricks_device(lvr_ctx* ctx, const lvr_noise_params* p,
                            const uint64_t* node_ids, int n);

/* Two more procedural sources, host/device bit-identical like the noise:
 * LVR_FIELD_BLOBS sums `count` anisotropic rotated bumps plus uniform noise
 * of amplitude `noise` (scaled to the type's range; floats in [0, 1]);
 * LVR_FIELD_SINES is the mean of `count` low-frequency plane waves, in
 * [-1, 1]. Supported types: u8, u16, float. No reference counterpart (same
 * reason as the noise source). */
#define LVR_FIELD_BLOBS 0
#define LVR_FIELD_SINES 1
typedef struct
{
    int32_t kind;  /* LVR_FIELD_BLOBS or LVR_FIELD_SINES */
    int32_t count; /* blobs or waves, 1..4096 */
    uint64_t seed;
    float noise;   /* blobs: uniform noise amplitude, relative to one bump */
    int32_t pad;
} lvr_field_params;

int lvr_field_brick_host(const lvr_volume_info* info,
                         const lvr_field_params* p, uint64_t node_id,
                         void* out, size_t bytes);
int lvr_field_bricks_device(lvr_ctx* ctx, const lvr_field_params* p,
                            const uint64_t* node_ids, int n);

/* Blobs plus value noise, the out-of-core configuration's source ("generated
 * per brick from a seed, blobs plus value noise"). A voxel is the integer
 * blend (blob * w + noise * (256 - w) + 128) >> 8 of two 16-bit values at
 * its centre, then converted to the volume's type as the noise is:
 *   - noise: the value-noise source above with `noise`;
 *   - blob: a smooth field of `blob_count` Gaussian blobs (LVR_FIELD_BLOBS
 *     with seed `blob_seed` and no uniform noise, scaled to 0..65535),
 *     tabulated at create time on a lattice of blob_cells + 1 nodes per axis
 *     (node i at i / blob_cells of the largest axis) and interpolated
 *     trilinearly in 16.16 fixed point, like the noise octaves.
 * Host and device bricks are bit-identical. blob_count 1..4096, blob_cells
 * 1..1024, blob_weight 0..256 (0: the noise source exactly). The handle is
 * read-only after create and may be shared by source threads. No reference
 * counterpart (the reference's MemoryDataSource fills bricks with one
 * constant, MemoryDataSource.cpp:38-61). */
typedef struct lvr_blobnoise lvr_blobnoise;
int lvr_blobnoise_create(const lvr_volume_info* info,
                         const lvr_noise_params* noise, uint64_t blob_seed,
                         int blob_count, int blob_cells, int blob_weight,
                         lvr_blobnoise** out);
void lvr_blobnoise_destroy(lvr_blobnoise* bn);
int lvr_blobnoise_brick_host(const lvr_blobnoise* bn, uint64_t node_id,
                             void* out, size_t bytes);
/* Device generation into the context's pool; the context's volume must be
 * the one the handle was made for. */
int lvr_blobnoise_bricks_device(lvr_ctx* ctx, const lvr_blobnoise* bn,
                                const uint64_t* node_ids, int n);

/* ------------------------------------------------------- raw / NRRD data */
/* RawDataSource (livre/data/RawDataSource.cpp:62-219): a memory-mapped raw
 * (x fastest) or NRRD volume. uri:
 *   [raw://]/path.raw|.img[?output=type]#nx,ny,nz[,type]   (type default uint8)
 *   [raw://]/path.nrrd[?output=type]     (type, dimension 3, sizes, endian,
 *                                         encoding raw, optional datafile)
 * with the reference's type names (char, int8, unsigned char, uint8, short,
 * int16, unsigned short, uint16, int, int32, unsigned int, uint32, float, and
 * NRRD's aliases such as uchar). An output type narrower than the input keeps
 * the high bits (uint16 -> uint8, uint32 -> uint8 or uint16, as _scale does).
 * The volume information is the reference's: the whole volume is one brick
 * (depth 1, no overlap). On failure *out is null and error receives the
 * reason. */
typedef struct lvr_raw_source lvr_raw_source;
typedef struct
{
    lvr_volume_info info;  /* one brick, output type */
    int32_t input_type;    /* LVR_DT_* of the file */
    int32_t big_endian;    /* NRRD "endian: big"; bytes are passed on as stored */
    uint64_t data_bytes;   /* voxels * input bytes per voxel */
} lvr_raw_info;
int lvr_raw_open(const char* uri, lvr_raw_source** out, char* error, size_t error_len);
void lvr_raw_close(lvr_raw_source* src);
int lvr_raw_get_info(const lvr_raw_source* src, lvr_raw_info* info);
/* DataSourcePlugin::getData of the one brick (node id of level 0 at the
 * origin), an lvr_brick_source_fn with user = the source. */
int lvr_raw_brick(void* user, uint64_t node_id, void* out, size_t bytes);

/* LOD pyramid of a raw source (no reference counterpart: the reference
 * serves a raw volume as one brick). The file becomes the finest level of the
 * regular octree lvr_pyramid_volume_info describes (block and overlap chosen
 * by the caller; zero padding up to the octree's extent); each coarser level
 * is the 2x2x2 mean of the next (integers rounded half up, floats
 * ((a+b)+(c+d))+((e+f)+(g+h)) times 1/8); a brick holds its level's voxels
 * [b*block - overlap, (b+1)*block + overlap), clamped to the level. The
 * levels are built in HBM on the first load (about 8/7 of the padded finest
 * level) and kept until lvr_pyramid_release or another source is loaded;
 * bricks are cut straight into pool slots. The context must have been
 * created with the pyramid's volume information; cubic brick grids only. */
int lvr_pyramid_volume_info(const lvr_raw_source* src, uint32_t block, uint32_t overlap,
                            lvr_volume_info* info);
int lvr_pyramid_load(lvr_ctx* ctx, const lvr_raw_source* src, const uint64_t* node_ids,
                     int n);
int lvr_pyramid_release(lvr_ctx* ctx);

/* ----------------------------------------------------- transfer function */
/* FP32 RGBA transfer function of n entries over the data range
 * [range0, range1] (RayCastRenderer.cpp:161-193, 280-285). Sampled with
 * linear filtering and clamp-to-edge, as the GL 1D texture is. As in the
 * reference, the range is used only when range1 > 0 and range1 > range0;
 * otherwise frames map the volume type's full range (numeric_limits min and
 * max, RayCastRenderer.cpp:236-285; for float volumes FLT_MIN and FLT_MAX).
 * Float histograms need a valid range. */
int lvr_tf_set(lvr_ctx* ctx, const float* rgba, int n, float range0,
               float range1);
/* RGBA8 LUT as TransferFunction1D::getLUT produces (TransferFunction1D.cpp:
 * 92-107); entries are read as unorm8, the GL_RGBA texture format. */
int lvr_tf_set_rgba8(lvr_ctx* ctx, const uint8_t* lut, int n, float range0,
                     float range1);

/* ------------------------------------------------------------ selection */
/* SelectVisibles + DFSTraversal (livre/data/SelectVisibles.cpp:45-131,
 * livre/data/DFSTraversal.cpp:29-76) on the device, level by level, with
 * the reference's FP32 expressions. Output in the reference's DFS order.
 * n_out receives the count; LVR_ERANGE if cap is too small. Runs on a
 * stream of its own: it does not wait for frames still rendering. */
int lvr_select_visibles(lvr_ctx* ctx, const lvr_frustum* f,
                        uint32_t window_height, float sse, uint32_t min_lod,
                        uint32_t max_lod, uint64_t* out, int cap, int* n_out);
/* lvr_select_visibles in two steps, for pipelined frames: _async queues the
 * selection (the kernel writes its result into pinned host memory) and
 * returns at once; _wait returns that result, as lvr_select_visibles does.
 * One selection may be pending per context (LVR_EINVAL otherwise). A
 * selection queued one frame ahead runs whenever the GPU has room, typically
 * between two ray casts. No reference counterpart (the reference selects on
 * the render thread). */
int lvr_select_visibles_async(lvr_ctx* ctx, const lvr_frustum* f,
                              uint32_t window_height, float sse, uint32_t min_lod,
                              uint32_t max_lod);
int lvr_select_visibles_wait(lvr_ctx* ctx, uint64_t* out, int cap, int* n_out);

/* The selection with everything SelectVisibles takes (SelectVisibles.cpp:
 * 30-41): clip planes, and the sort-last range [range0, range1) of the
 * DFS-ordered visibles that visitPost keeps (lines 109-131: indices
 * [range0 * N, range1 * N), the products in float, truncated). A node wholly
 * in the half-space a clip plane removes (n . x + d < 0 at the box corner
 * furthest along n, the half-space fragRayCast.glsl:152-167 clips away) is
 * culled with its subtree, as _clipPlanes.isOutside(worldBox) culls it
 * (line 69). n_out receives the count of the range. */
typedef struct
{
    uint32_t window_height;
    float sse;
    uint32_t min_lod, max_lod;
    int32_t n_clip_planes;  /* 0..6 */
    float clip_planes[24];  /* normal xyz, d */
    float range0, range1;   /* 0 <= range0 <= range1 <= 1; {0, 1} keeps all */
} lvr_select_params;
int lvr_select_visibles_ex(lvr_ctx* ctx, const lvr_frustum* f, const lvr_select_params* params,
                           uint64_t* out, int cap, int* n_out);
/* The same in two steps (lvr_select_visibles_wait returns the range). */
int lvr_select_visibles_ex_async(lvr_ctx* ctx, const lvr_frustum* f,
                                 const lvr_select_params* params);

/* RayCastRenderer::order (RayCastRenderer.cpp:57-81, 153-159): front to
 * back by |MV * brick centre|, with libstdc++ std::sort on the given order. */
int lvr_order(const lvr_volume_info* info, const lvr_frustum* f,
              const uint64_t* ids, int n, uint64_t* out);

/* ---------------------------------------------------------------- render */
/* nSamplesPerRay of a frame (RayCastRenderer::Impl::onFrameStart,
 * RayCastRenderer.cpp:216-234): `requested` when above 0, otherwise
 * max(maxVoxelDim / 2^(depth - maxLOD - 1), 512) truncated, maxLOD being the
 * finest level among the n rendered bricks. Host only. */
int lvr_samples_per_ray(const lvr_volume_info* info, const uint64_t* ids, int n,
                        int32_t requested, int32_t* out);

typedef struct
{
    int32_t n_samples_per_ray;   /* nSamplesPerRay, at most 1048576; 0 =
                                    lvr_samples_per_ray's rule over the frame's
                                    bricks (RayCastRenderer.cpp:216-234) */
    int32_t max_samples_per_ray; /* opacity reference, 32 in the reference */
    float early_exit;            /* EARLY_EXIT, 0.999 (fragRayCast.glsl:12) */
    int32_t linear_filter;       /* 0: GL_NEAREST (reference default), 1: trilinear */
    int32_t n_clip_planes;       /* 0..6 */
    float clip_planes[24];       /* normal xyz, d */
    int32_t tile_w, tile_h;      /* wavefront ray tile (w * h = 64); 0, 0 = 8 x 8 */
    int32_t samples_per_thread;  /* voxel fetches in flight per ray, 1..3 (NEAREST);
                                    0 = 2. With 2 and per-sample classification a
                                    ray may run one brick ahead of the others of
                                    its wavefront tile. GL_LINEAR: 1 brick by brick,
                                    2 (or 3) with the one-brick-ahead window, the
                                    next sample's 8 texels in flight either way;
                                    identical images always */
    int32_t classify;            /* 8/16-bit NEAREST: 1 = classification tabulated
                                    per raw value (one table read per sample),
                                    2 = computed per sample from the TF,
                                    0 = automatic (table for 8-bit data);
                                    identical images either way */
    int32_t waves_per_simd;      /* wavefronts per SIMD the ray cast keeps
                                    resident: 2 to 8; 0 = 5 for NEAREST
                                    with 2 samples per thread and per-sample
                                    classification, 4 otherwise. More hide
                                    more latency, fewer thrash the L1 less */
} lvr_render_params;

/* Render one frame of the ordered bricks (fragRayCast.glsl:113-207 per brick,
 * RayCastRenderer.cpp:467-571) into the context framebuffer, followed by
 * the RGBA32F -> RGBA8 copy (fragTexCopy.glsl:11-15). viewport = {x, y, w, h};
 * only w and h are used (the frame covers the whole viewport). The host sends
 * each brick's NodeId and page-table index; the device builds the brick
 * records and reads each brick's slot from the device page table when the
 * frame's GPU work starts (the reference looks each texture up by NodeId when
 * it draws the brick, RayCastRenderer.cpp:510-522). Bricks that are not
 * resident then are skipped, as the reference skips a missing texture
 * (RayCastRenderer.cpp:517-522). Asynchronous on the context stream. */
int lvr_render(lvr_ctx* ctx, const lvr_frustum* f, const int viewport[4],
               const uint64_t* ordered_ids, int n,
               const lvr_render_params* params);
/* lvr_render without allocating: no reference counterpart. For callers that
 * cannot allocate device or pinned memory during a frame (code under HIP
 * graph capture, tight serving loops, harnesses that forbid it). The same
 * frame as lvr_render, written to out_rgba8 (caller device memory of
 * w * h * 4 bytes, 4-byte aligned, row 0 = bottom) instead of the context's
 * framebuffer; out_samples (caller device memory, 8-byte aligned, or NULL)
 * receives the composite-step count. The call allocates nothing: the kernels'
 * scratch is one static device array of the library (about 26 MB per device,
 * held until the process ends) and the host staging is fixed at
 * lvr_ctx_create. Capacity per call: 32768 bricks, 262144 ray tiles (4096 x
 * 4096 pixels at 8 x 8 rays), 4M (tile, brick) pairs; beyond it LVR_ERANGE.
 * Every call on one device shares that array, which keeps the last call's
 * contents: calls that overlap on one device (contexts on different streams)
 * race, so serialise them. Asynchronous on the context's render stream; no
 * frame statistics are recorded. */
int lvr_render_noalloc(lvr_ctx* ctx, const lvr_frustum* f, const int viewport[4],
                       const uint64_t* ordered_ids, int n, const lvr_render_params* params,
                       void* out_rgba8, uint64_t* out_samples);

/* ------------------------------------------------ sort-last compositing */
/* Sort-last decomposition (SelectVisibles::Impl::visitPost keeps the range
 * [r0 * N, r1 * N) of the visibles, SelectVisibles.cpp:112-131; Equalizer's
 * compositor blends the partial frames in depth order, Channel.cpp:464-603).
 * lvr_render_partial ray-casts ordered_ids, a part of the frame's global
 * front-to-back order kept in that order (a contiguous range of it, or the
 * bricks of one part of lvr_sortlast_parts), like lvr_render, into the
 * context's layer `layer` (0..15): w * h premultiplied RGBA32F accumulations,
 * row 0 = bottom, not converted to RGBA8. Sample positions are those of the
 * whole frame (global-grid snapping), so the layers of consecutive ranges,
 * composited front first, or the layers of the parts, composited in
 * lvr_sortlast_order, give the whole frame up to FP rounding and the early
 * exit each layer takes on its own. The RGBA8 framebuffer is left as it was;
 * the frame statistics (lvr_frame_stats_get, the samples of lvr_frame_read)
 * are the partial's. */
int lvr_render_partial(lvr_ctx* ctx, const lvr_frustum* f, const int viewport[4],
                       const uint64_t* ordered_ids, int n, const lvr_render_params* params,
                       int layer);
/* Static sort-last decomposition (SelectVisibles.cpp:112-131 built with
 * LIVRE_STATIC_DECOMPOSITION: a node belongs to one channel whatever the
 * camera, by its own position, NodeId.cpp:129-138). Here the parts are the
 * halves of the octree's block grid split at its middle along x, then y,
 * then z: n_parts = 1, 2, 4 or 8 (8: the level-1 octants). part_of[i] is
 * the part of ids[i] (bit 2 x, bit 1 y, bit 0 z of the halves used, 1 = the
 * high half), so a brick keeps its part, and a GPU its bricks, as the
 * camera moves. A node that straddles a split (the root, or a level-0 node
 * of a root grid with an odd block count along a split axis) gets -1: it is
 * a frame's only node when it is visible with a power-of-two root grid.
 * LVR_EINVAL for another n_parts or an id outside the octree. */
int lvr_sortlast_parts(const lvr_volume_info* info, const uint64_t* ids, int n, int n_parts,
                       int32_t* part_of);
/* The front-to-back order of the parts for an eye position (world space):
 * at each split the half holding the eye first, x before y before z (the
 * traversal of the splits as a BSP tree), so every ray meets the layers in
 * this order. order receives n_parts part indices. */
int lvr_sortlast_order(const lvr_volume_info* info, const float eye[3], int n_parts,
                       int32_t* order);
/* Device pointer of a layer (w * h float4), e.g. for an exchange between
 * GPUs; null until the layer is first rendered. The pointer stays valid until
 * a partial render into the same layer at a larger viewport replaces the
 * buffer (the old one is freed at the next lvr_ctx_sync). */
int lvr_partial_device_ptr(lvr_ctx* ctx, int layer, void** ptr);
/* Copy a layer to host memory: width * height premultiplied RGBA32F pixels,
 * row 0 = bottom (bytes >= width * height * 16, LVR_ERANGE otherwise;
 * out_rgba32f may be NULL to learn the size). Synchronous. No reference
 * counterpart: Equalizer reads partial frames back inside its compositor. */
int lvr_partial_read(lvr_ctx* ctx, int layer, float* out_rgba32f, size_t bytes, int* width,
                     int* height);
/* Front-to-back "over" of n (1..16) RGBA32F device images of width x height
 * (front first) into the context's RGBA8 framebuffer with fragTexCopy's
 * rounding; read it with lvr_frame_read. A layer of this context must have
 * been rendered at exactly width x height, and a layer buffer that has since
 * been replaced is rejected (LVR_EINVAL); other device pointers must hold
 * width * height float4. The kernel runs reps (>= 1) times,
 * each time with the same result; kernel_ms, when not NULL, receives the
 * mean time per launch over the reps back-to-back launches (the call then
 * waits for them). */
int lvr_composite_layers(lvr_ctx* ctx, const void* const* layers, int n, int width,
                         int height, int reps, float* kernel_ms);

/* Copy the last frame to the host (w * h * 4 bytes) and return the number of
 * executed composite steps (fragRayCast.glsl:184-204 iterations). */
int lvr_frame_read(lvr_ctx* ctx, uint8_t* out_rgba8, size_t bytes,
                   uint64_t* samples);
/* Device pointer of the framebuffer (uint32 RGBA8 per pixel). */
int lvr_frame_device_ptr(lvr_ctx* ctx, void** ptr, int* width, int* height);

/* Statistics of the last frame. Kernel times come from HIP events recorded
 * on the context's stream around each stage while timing is enabled (the
 * reference has no timer on this path; it only counts cache statistics,
 * CacheStatistics.cpp:57-73). lvr_frame_stats_get synchronises. */
typedef struct
{
    uint64_t samples;      /* executed composite steps */
    uint64_t bricks;       /* resident bricks that reached the screen */
    uint64_t tile_entries; /* (tile, brick) pairs binned */
    uint64_t brick_bytes;  /* maximumBlockSize^3 * bytesPerVoxel per brick */
    float bin_ms;          /* binning kernels */
    float raycast_ms;      /* ray-cast kernel */
    float histogram_ms;    /* histogram kernels of the last lvr_histogram */
    float host_ms;         /* host time of the lvr_render call: preparation of the
                              brick keys and screen ranges, staging, launches */
} lvr_frame_stats;
int lvr_set_timing(lvr_ctx* ctx, int enable);
/* Statistics of the last rendered frame (waits for it). */
int lvr_frame_stats_get(lvr_ctx* ctx, lvr_frame_stats* out);
/* Statistics of the frame rendered `back` frames before the last one
 * (0 = the last, up to 3), for callers that keep frames in flight
 * (RenderPipeline::renderAsync). Waits for that frame only; LVR_ERANGE once
 * it is no longer kept. No reference counterpart (the GL path reads
 * nothing back per frame). */
int lvr_frame_stats_back(lvr_ctx* ctx, int back, lvr_frame_stats* out);

/* Autotuner of the launch shape (no reference counterpart: the GL path has
 * no tunable launch). Renders the given frame with every ray tile in
 * {8x8, 16x4, 4x16, 32x2, 2x32}, every waves_per_simd in {4, 5, 6, 8}, every
 * samples_per_thread in {1, 2, 3} (NEAREST) or {1, 2} (GL_LINEAR), and for
 * 8/16-bit NEAREST both classify modes; times the ray-cast kernel with HIP events (median of reps)
 * and writes the fastest tile_w, tile_h, samples_per_thread, classify and
 * waves_per_simd into *params.
 * All candidates produce the same image. The timings of the candidates are
 * copied to results (up to cap entries); *n_results = number of candidates.
 * Leaves the last candidate's frame in the framebuffer. */
typedef struct
{
    int32_t tile_w, tile_h, samples_per_thread, classify, waves_per_simd;
    float raycast_ms;
} lvr_tune_entry;
int lvr_autotune(lvr_ctx* ctx, const lvr_frustum* f, const int viewport[4],
                 const uint64_t* ordered_ids, int n, int reps,
                 lvr_render_params* params, lvr_tune_entry* results, int cap,
                 int* n_results);

/* --------------------------------------------------------------- batches */
/* Many frames (requests) of different volumes, viewports, cameras and
 * transfer functions rendered by one persistent launch: binning of every
 * request in one pass, then one work queue over all requests' ray tiles.
 * The launch sequence is captured as a HIP graph and replayed while the
 * batch keeps its shape. No reference counterpart: Livre renders one frame
 * per channel per call (Channel.cpp:283-336). Requests use nearest
 * filtering; u8, u16 and f32 volumes; transfer functions up to 256 entries
 * are staged in LDS, larger ones (up to 8192) are read from memory. */
typedef struct lvr_batch lvr_batch;
typedef struct
{
    uint64_t requests;
    uint64_t samples;      /* composite steps of all requests */
    uint64_t bricks;       /* rendered bricks of all requests */
    uint64_t tile_entries; /* (tile, brick) pairs binned */
    float gpu_ms;          /* binning + ray casting on the device */
    uint64_t graph_captures, graph_launches;
    uint64_t solo_requests; /* requests launched on their own (set_solo_tiles) */
} lvr_batch_stats;

int lvr_batch_create(int device, lvr_batch** out);
void lvr_batch_destroy(lvr_batch* batch);
const char* lvr_batch_last_error(const lvr_batch* batch);
/* 1 (default): capture and replay the launch sequence as a HIP graph. */
int lvr_batch_set_graph(lvr_batch* batch, int enable);
/* Requests that the single-frame kernel's two-brick window covers (NEAREST,
 * 2 samples per thread, classification per sample) and that hold at least
 * min_tiles ray tiles are launched on their own, one single-frame launch
 * each, instead of in the batch kernel (which marches brick by brick); a
 * lone such request always is. Without reserved CUs these launches are
 * dealt, largest first, to the batch's stream and up to three more that
 * start after the binning, so that one launch's tail overlaps the next and
 * the shared launch. 0 (default) = four tiles per resident wavefront of a
 * launch (2 workgroups of 2 x waves_per_simd wavefronts per CU), negative =
 * only a lone request. A batch with such requests is not captured as a
 * graph. Frames and sample counts do not depend on the split. No reference
 * counterpart. */
int lvr_batch_set_solo_tiles(lvr_batch* batch, int min_tiles);
/* Queue one frame of ctx's volume; request receives its index. The TF is
 * n_tf FP32 RGBA entries over [range0, range1] (as lvr_tf_set). Adding after
 * a run starts a new batch. One launch renders all requests, so their
 * waves_per_simd must agree, and so must samples_per_thread among requests
 * of the same voxel type (8-bit, 16-bit, float; LVR_EINVAL otherwise). The
 * batch pipelines 2 or 3 fetches per ray (1 runs as 2, with the same
 * samples). */
int lvr_batch_add(lvr_batch* batch, lvr_ctx* ctx, const lvr_frustum* f,
                  const int viewport[4], const uint64_t* ordered_ids, int n,
                  const lvr_render_params* params, const float* tf_rgba,
                  int n_tf, float range0, float range1, int* request);
int lvr_batch_clear(lvr_batch* batch);
/* Launch every queued request (asynchronous). */
int lvr_batch_run(lvr_batch* batch);
/* Wait for the launch; fills stats (may be NULL). */
int lvr_batch_wait(lvr_batch* batch, lvr_batch_stats* stats);
/* RGBA8 frame (w * h * 4 bytes, row 0 = bottom) and composite steps of one
 * request, after lvr_batch_wait. */
int lvr_batch_read(lvr_batch* batch, int request, uint8_t* out_rgba8,
                   size_t bytes, uint64_t* samples);
/* 1 when the last launch has finished, 0 while it runs (never blocks). */
int lvr_batch_query(lvr_batch* batch);
/* Render request `request` (added since the last run) into caller device
 * memory of width * height * 4 bytes instead of the batch's own frames. */
int lvr_batch_set_output(lvr_batch* batch, int request, void* device_rgba8);
/* Keep n CUs free of this batch's launches (a CU mask on its stream), for
 * work other streams run while a launch is in flight. 0 = use every CU. */
int lvr_batch_reserve_cus(lvr_batch* batch, int n);

/* --------------------------------------------------------------- serving */
/* Request loop for concurrent camera sessions on one device, by continuous
 * batching (the reference serves one session per process through
 * Client::run / Config::frame, Client.cpp:89-122, Config.cpp:298-342, with
 * no batching). Sessions are admitted against a byte budget of device memory
 * (the volumes' brick pools + 3 frames of w*h*4 per session); later ones wait
 * until a session closes. Submitted frame requests queue; lvr_server_step
 * retires finished launches and puts every queued request of an admitted
 * session (up to max_batch) into the next persistent batched launch while at
 * most three other launches are on the GPU. A newer request of a session
 * supersedes its older one still queued (status 1). Missing bricks are
 * produced on the device by the volume's procedural source, evicting
 * least-recently-used bricks; bricks read by queued or running launches are
 * pinned. Frames render straight into per-session device framebuffers
 * (five per session, so the finished one is never one being written);
 * launches leave a few CUs to the selection and source kernels. Not
 * thread-safe: one host thread drives a server. While a context belongs to a
 * server, other batches must not read its pool. */
typedef struct lvr_server lvr_server;
typedef struct
{
    int64_t request;
    int32_t session;
    int32_t status;      /* 0 rendered, 1 superseded, < 0 error code */
    uint64_t samples;    /* executed composite steps */
    double arrive_ms, launch_ms, done_ms; /* lvr_server_now_ms clock */
} lvr_server_frame;
typedef struct
{
    uint64_t admitted, frames, batches, samples, superseded, failed;
    uint64_t bricks_generated;
    uint64_t budget_bytes, used_bytes, pending, waiting_sessions, in_flight;
    /* host time spent in the scheduler, ms: visible-set selection, ordering,
     * brick residency (misses + pins), batch assembly, launch, retirement */
    double select_ms, order_ms, residency_ms, add_ms, run_ms, retire_ms;
    /* GPU time of the launches (binning + ray casting), summed, ms */
    double gpu_ms;
} lvr_server_stats;
double lvr_server_now_ms(void);
int lvr_server_create(int device, size_t budget_bytes, int max_batch,
                      lvr_server** out);
void lvr_server_destroy(lvr_server* server);
const char* lvr_server_last_error(const lvr_server* server);
int lvr_server_set_coalesce(lvr_server* server, int enable);
/* Serve a volume from ctx's pool. noise or field (at most one) names the
 * procedural source of missing bricks; with neither, the caller keeps the
 * needed bricks resident. The pool's bytes count against the budget. The
 * server queues the selection of a pending request on ctx ahead of its
 * launch (lvr_select_visibles_async), so between a submit and the return of
 * lvr_server_drain the caller starts no selection of its own on ctx. */
int lvr_server_add_volume(lvr_server* server, lvr_ctx* ctx,
                          const lvr_noise_params* noise,
                          const lvr_field_params* field, int* volume);
/* Open a session of width x height frames with its own transfer function and
 * render parameters (NEAREST). *admitted = 1 if it fits the budget now. */
int lvr_server_open(lvr_server* server, int volume, int width, int height,
                    const float* tf_rgba, int n_tf, float range0, float range1,
                    const lvr_render_params* params, int* session,
                    int* admitted);
int lvr_server_close(lvr_server* server, int session);
/* Queue a frame request of a session; arrive_ms stamps it (server clock). */
int lvr_server_submit(lvr_server* server, int session, const lvr_frustum* f,
                      double arrive_ms, int64_t* request);
/* Non-blocking scheduler step: retire finished launches, launch the queue. */
int lvr_server_step(lvr_server* server, int* launched, int* retired);
/* Step until nothing admitted is queued or running (blocks). */
int lvr_server_drain(lvr_server* server);
/* Take up to cap finished (or superseded / failed) requests. */
int lvr_server_poll(lvr_server* server, lvr_server_frame* out, int cap, int* n);
/* The session's last finished frame (RGBA8, row 0 = bottom). */
int lvr_server_read(lvr_server* server, int session, uint8_t* out_rgba8,
                    size_t bytes, uint64_t* samples);
int lvr_server_get_stats(lvr_server* server, lvr_server_stats* out);


/* --------------------------------------------------------- out-of-core */
/* DataSourcePlugin::getData as a callback (DataSourcePlugin.h:81): write the
 * brick of node_id, bytes = maximumBlockSize^3 * bytesPerVoxel, x fastest;
 * return 0 or an error code. Called from the stream's worker threads. */
typedef int (*lvr_brick_source_fn)(void* user, uint64_t node_id, void* out,
                                   size_t bytes);
/* The value-noise source as such a callback; user = lvr_noise_source_data. */
typedef struct
{
    lvr_volume_info info;
    lvr_noise_params params;
} lvr_noise_source_data;
int lvr_noise_source(void* user, uint64_t node_id, void* out, size_t bytes);
/* The blobs-plus-noise source as such a callback; user = lvr_blobnoise*. */
int lvr_blobnoise_source(void* user, uint64_t node_id, void* out, size_t bytes);

/* Asynchronous brick streaming into a context's pool (the reference's async
 * RenderPipeline + DataUploadFilter, RenderPipeline.cpp:160-225,
 * DataUploadFilter.cpp:104-120): `threads` host threads call the source into
 * `staging_bricks` pinned buffers; lvr_stream_pump (caller's thread) takes
 * pool slots (LRU, never a pinned brick) and issues hipMemcpyAsync + the
 * layout conversion on the context's upload stream; bricks count as
 * resident once that stream has passed them. While a stream exists, a frame
 * must pin the bricks it renders (lvr_brick_pin) until it has finished. */
typedef struct lvr_stream lvr_stream;
typedef struct
{
    uint64_t requested, bricks_loaded, bytes_loaded, transfers;
    uint64_t source_errors, pool_full;
    uint64_t queued, generating, ready, in_flight;
} lvr_stream_stats;
int lvr_stream_create(lvr_ctx* ctx, lvr_brick_source_fn source, void* user,
                      int threads, int staging_bricks, lvr_stream** out);
void lvr_stream_destroy(lvr_stream* stream);
const char* lvr_stream_last_error(const lvr_stream* stream);
/* Replace the wish list: the given bricks that are neither resident nor on
 * their way, coarse levels first, then in the given order. Resident bricks
 * among the given ones become the most recently used (evicted last). */
int lvr_stream_request(lvr_stream* stream, const uint64_t* ids, int n);
/* Same; with ancestors != 0 the wish list also holds every ancestor of the
 * given bricks (coarse first), so fallbacks refine level by level while the
 * full-resolution bricks stream in. */
int lvr_stream_request_ex(lvr_stream* stream, const uint64_t* ids, int n,
                          int ancestors);
/* Non-blocking: retire finished transfers, start the generated ones.
 * *completed = bricks that became resident. */
int lvr_stream_pump(lvr_stream* stream, int* completed);
/* Pump until every wished brick is resident (blocks). */
int lvr_stream_wait(lvr_stream* stream);
/* RenderingSetGenerator::generateRenderingSet (RenderingSetGeneratorFilter.cpp:
 * 39-108): every visible brick, or its nearest resident ancestor; then,
 * when that set's size differs from n, the bricks with an ancestor in the
 * set are dropped. Sorted by id (the caller orders front to back).
 * *n_available = visible bricks resident themselves. */
int lvr_stream_rendering_set(lvr_stream* stream, const uint64_t* visible,
                             int n, uint64_t* out, int cap, int* n_out,
                             int* n_available);
/* The same rule against an explicit resident set (no device). */
int lvr_rendering_set(const uint64_t* visible, int n, const uint64_t* resident,
                      int n_resident, uint64_t* out, int cap, int* n_out,
                      int* n_available);
int lvr_stream_get_stats(lvr_stream* stream, lvr_stream_stats* out);
/* Pin (delta = +1) or unpin (-1) resident bricks: pinned bricks are never
 * evicted (a texture held by a renderer, Cache.cpp:148-149). */
int lvr_brick_pin(lvr_ctx* ctx, const uint64_t* ids, int n, int delta);

/* ------------------------------------------------------------- histogram */
/* Per-frame histogram of the given bricks (HistogramObject.cpp:140-224 +
 * HistogramFilter.cpp:78-149): counts of interior voxels scaled by
 * 8^(depth - level - 1). 8-bit data: 256 bins, bin = value; 16-bit data:
 * 65536 bins, bin = value; float data: 256 bins over the TF range. The
 * kernels run on the context's histogram stream, beside the ray casting of
 * the frame; the result stays on the device until lvr_histogram_read. */
int lvr_histogram(lvr_ctx* ctx, const uint64_t* ids, int n, int bins);
int lvr_histogram_read(lvr_ctx* ctx, uint64_t* out, int bins);

/* ------------------------------------------------- frame grab as JPEG */
/* FrameGrabber::_encodeJpeg (livre/eq/FrameGrabber.cpp:49-84) compresses the
 * read-back frame with libjpeg-turbo's tjCompress2: TJPF_BGRA, TJSAMP_444,
 * quality 100, flags TJXOP_ROT180 (= 6, read as TJFLAG_BOTTOMUP). These entry
 * points write the same JFIF file on the device: baseline JPEG, 4:4:4, the
 * accurate integer DCT (tjCompress2's choice at quality >= 96), standard
 * tables scaled to `quality`. notifyNewImage (FrameGrabber.cpp:86-109) sends
 * the bytes on; here they land in a host buffer. */
#define LVR_PIXEL_RGBA 0
#define LVR_PIXEL_BGRA 1
typedef struct lvr_jpeg_encoder lvr_jpeg_encoder;
typedef struct
{
    float device_ms;  /* kernels of the last encode (HIP events), the sum of: */
    float dct_ms;     /* colour conversion, DCT, quantisation, code lengths */
    float offsets_ms; /* DC code lengths, bit offsets of the blocks */
    float emit_ms;    /* Huffman codes into the bit stream */
    float stuff_ms;   /* 0x00 after each 0xFF, EOI */
    int32_t width, height;
    uint64_t bytes; /* file length of the last encode */
} lvr_jpeg_stats;
/* Encoder with its own stream and grow-only device scratch. */
int lvr_jpeg_encoder_create(int device, lvr_jpeg_encoder** out);
void lvr_jpeg_encoder_destroy(lvr_jpeg_encoder* enc);
/* Encode device pixels (4 bytes each, 4-byte aligned, pitch a multiple of 4,
 * 1 <= width, height <= 16384). bottom_up != 0 takes row height-1 as the
 * top row (TJFLAG_BOTTOMUP). *size receives the file length; out may be NULL
 * to learn it (the encode still runs); LVR_ERANGE if cap is too small.
 * Synchronous: returns once the bytes are in `out`. */
int lvr_jpeg_encode(lvr_jpeg_encoder* enc, const void* device_pixels, int width, int height,
                    int pitch_bytes, int pixel_format, int bottom_up, int quality,
                    uint8_t* out, size_t cap, size_t* size);
/* The same from host pixels, as FrameGrabber::_encodeJpeg receives them
 * (FrameGrabber.cpp:49-53: the read-back eq::Image); they are copied to the
 * encoder's device buffer first. */
int lvr_jpeg_encode_host(lvr_jpeg_encoder* enc, const void* host_pixels, int width,
                         int height, int pitch_bytes, int pixel_format, int bottom_up,
                         int quality, uint8_t* out, size_t cap, size_t* size);
/* Upper bound of the file length for a width x height image. */
size_t lvr_jpeg_max_size(int width, int height);
/* Last error of enc; with enc NULL, why the last create on this thread failed. */
const char* lvr_jpeg_error(const lvr_jpeg_encoder* enc);
int lvr_jpeg_get_stats(const lvr_jpeg_encoder* enc, lvr_jpeg_stats* out);
/* Grab the context's last frame (RGBA8, row 0 = bottom, as lvr_frame_read
 * returns it) as the reference's frame grabber encodes it; quality 100 in
 * the reference. Runs on the context's stream after the frame. */
int lvr_frame_grab_jpeg(lvr_ctx* ctx, int quality, uint8_t* out, size_t cap, size_t* size);
int lvr_frame_grab_stats(lvr_ctx* ctx, lvr_jpeg_stats* out);

/* ---------------------------------------------------- session front end */
/* The messages of a session as Livre's ZeroEQ communicator handles them
 * (livre/eq/zeroeq/communicator.cpp); the transport (ZeroEQ publish /
 * subscribe, the HTTP server) is not part of this library. */
/* LookOut → modelview (onCamera, communicator.cpp:106-133): the client's
 * view matrix (meters, the data's frame) to the renderer's (Livre space):
 * the camera position -R^-1 t scaled by meter_to_data_unit, taken through
 * data_to_livre (VolumeInformation::dataToLivreTransform; NULL = identity),
 * then -R (that) as the new translation. Matrices are 16 floats,
 * column-major. LVR_EINVAL for a singular rotation or a zero ratio. */
int lvr_lookout_to_modelview(const float lookout[16], const float data_to_livre[16],
                             float meter_to_data_unit, float out_modelview[16]);
/* Modelview → LookOut (_createLookOut, communicator.cpp:257-283), the
 * inverse conversion, for publishing the camera. */
int lvr_modelview_to_lookout(const float modelview[16], const float data_to_livre[16],
                             float meter_to_data_unit, float out_lookout[16]);
/* TransferFunction1D message (the communicator subscribes to the render
 * settings' transfer function, communicator.cpp:236): the session's TF and
 * data range for the requests staged from now on. */
int lvr_server_set_tf(lvr_server* server, int session, const float* tf_rgba, int n_tf,
                      float range0, float range1);
/* VolumeRendererParameters and ClipPlanes messages (communicator.cpp:
 * 235-237): the session's render parameters (NEAREST, up to 6 clip planes)
 * for the requests staged from now on. */
int lvr_server_set_params(lvr_server* server, int session, const lvr_render_params* params);
/* ImageJPEG request (communicator.cpp:197-200 → Config::renderJPEG,
 * Config.cpp:217-240, and FrameGrabber): the session's last finished frame as
 * a JPEG file, encoded on the device from the session's framebuffer (rows
 * bottom-up, as lvr_frame_grab_jpeg; *size receives the length; LVR_ERANGE
 * if cap is too small, out may be NULL to learn it). */
int lvr_server_grab_jpeg(lvr_server* server, int session, lvr_jpeg_encoder* enc, int quality,
                         uint8_t* out, size_t cap, size_t* size);

#ifdef __cplusplus
}
#endif
#endif
