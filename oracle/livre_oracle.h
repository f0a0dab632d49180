/*
 * livre_oracle.h — CPU restatement of the Livre render hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests/, the smoke
 * check and the CPU-baseline leg of bench.py may load it; the product
 * (livre_amd/) never links, loads or calls it.
 *
 * Every function restates a reference function; the citation is given at the
 * definition in livre_oracle.cpp. Matrices are column-major float[16] (OpenGL
 * and vmmlib layout). Bricks are (block + 2*overlap)^3 voxels, x fastest.
 */
#ifndef LIVRE_ORACLE_H
#define LIVRE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors livre::DataType (livre/data/VolumeInformation.h:29-39). */
enum
{
    ORC_DT_FLOAT = 0,
    ORC_DT_UINT8 = 1,
    ORC_DT_UINT16 = 2,
    ORC_DT_UINT32 = 3,
    ORC_DT_INT8 = 4,
    ORC_DT_INT16 = 5,
    ORC_DT_INT32 = 6
};

typedef struct
{
    /* inputs */
    uint32_t voxels[3];
    uint32_t max_block[3]; /* maximumBlockSize = block + 2*overlap */
    uint32_t overlap[3];
    int32_t data_type;
    /* outputs of orc_fill_volume_info */
    float world_size[3];
    float world_space_per_voxel;
    uint32_t depth; /* RootNode::getDepth(), i.e. number of levels */
    uint32_t root_blocks[3];
} orc_volume_info;

/* Shader uniforms of one frame (fragRayCast.glsl:29-51). */
typedef struct
{
    float inv_proj[16];
    float inv_mv[16];
    float eye[3];
    float near_plane;
    float far_plane;
    int32_t width;
    int32_t height;
} orc_camera;

typedef struct
{
    int32_t n_samples_per_ray;   /* nSamplesPerRay */
    int32_t max_samples_per_ray; /* the opacity reference, 32 in Livre */
    float early_exit;            /* EARLY_EXIT, 0.999 in Livre */
    float range0, range1;        /* dataSourceRange */
    int32_t linear_filter;       /* 0: GL_NEAREST (reference default), 1: trilinear */
    int32_t n_clip_planes;
    float clip_planes[24];
    int32_t window[4]; /* x0, y0, x1, y1 (exclusive); all 0 = whole frame */
} orc_render_params;

int orc_fill_volume_info(orc_volume_info* vi);
int orc_node_world_box(const orc_volume_info* vi, uint64_t node_id, float box[6]);

/* Frustum quantities as vmmlib/livre compute them, for cross-checks. */
typedef struct
{
    float near_plane, far_plane, left, right, bottom, top;
    float near_plane_eq[4]; /* normalized near culling plane */
    float eye[3];
    float view_dir[3];
} orc_frustum_info;
int orc_frustum_info_get(const float mv[16], const float proj[16],
                         orc_frustum_info* out);

int orc_select_visibles(const orc_volume_info* vi, const float mv[16],
                        const float proj[16], uint32_t window_height,
                        float sse, uint32_t min_lod, uint32_t max_lod,
                        float range0, float range1, int n_clip,
                        const float* clip_planes, uint64_t* out, int cap,
                        int* n_out);

int orc_order(const orc_volume_info* vi, const float mv[16],
              const uint64_t* ids, int n, uint64_t* out);

/* brick_ptrs[i] points to the voxels of ordered[i]; a NULL pointer means the
 * brick is missing and is skipped (RayCastRenderer.cpp:517-522). out_rgba is
 * W*H*4 floats, row 0 = bottom row (GL window coordinates). */
int orc_render_f32(const orc_volume_info* vi, const orc_camera* cam,
                   const uint64_t* ordered, int n,
                   const void* const* brick_ptrs, const float* tf_rgba,
                   int n_tf, const orc_render_params* p, float* out_rgba,
                   uint64_t* samples, int n_threads);
int orc_render_f64(const orc_volume_info* vi, const orc_camera* cam,
                   const uint64_t* ordered, int n,
                   const void* const* brick_ptrs, const float* tf_rgba,
                   int n_tf, const orc_render_params* p, float* out_rgba,
                   uint64_t* samples, int n_threads);

/* Framebuffer write of fragTexCopy.glsl: unorm8 = round(clamp(c,0,1)*255). */
void orc_to_rgba8(const float* rgba, int n_pixels, uint8_t* out);

/* Per-frame histogram of the rendered set (HistogramObject + HistogramFilter).
 * bins: 256 for 8-bit and float data, 65536 for 16-bit data (bin = value). */
int orc_histogram(const orc_volume_info* vi, const uint64_t* ids, int n,
                  const void* const* brick_ptrs, int bins, float range0,
                  float range1, uint64_t* out);

/* TransferFunction1D default LUT and getLUT() truncation. */
void orc_default_tf(float* rgba256);
void orc_tf_lut(const float* rgba, int n, uint8_t* lut);

/* Frame grab as JPEG (FrameGrabber.cpp:49-84 through libjpeg-turbo): baseline
 * 4:4:4, accurate integer DCT, standard tables scaled to `quality`. Pixels are
 * 4 bytes (RGBA, or BGRA when bgra != 0); bottom_up takes row height-1 first.
 * *size receives the file length; out may be NULL to query it. Returns 0,
 * -1 on bad arguments, -2 when cap is too small. */
int orc_jpeg_encode(const uint8_t* pixels, int width, int height, int pitch,
                    int bgra, int bottom_up, int quality, uint8_t* out,
                    size_t cap, size_t* size);

#ifdef __cplusplus
}
#endif
#endif
