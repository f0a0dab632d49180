// Procedural DataSource plugins for the blob and smooth-field volumes (the
// fields themselves are in field_core.h).
//
// Like the value-noise source, every voxel of every brick (any level, halo
// included) is a pure function of (parameters, level, voxel coordinates) and
// is evaluated at the voxel's centre, so bricks are independent and can be
// produced on the device straight into the pool, equal to the bricks the
// host computes.
#include <vector>

#include "field_core.h"
#include "lvr_internal.h"

namespace lvr
{
namespace
{
using namespace field;

struct Geo
{
    int64_t origin[3];
    uint32_t level;
};

Geo brickOrigin(const lvr_volume_info& vi, uint64_t id)
{
    const NodeKey k = decodeNode(id);
    const uint32_t pos[3] = {k.x, k.y, k.z};
    Geo g;
    for (int a = 0; a < 3; ++a)
        g.origin[a] = int64_t(pos[a]) * (vi.max_block[a] - 2 * vi.overlap[a]) -
                      int64_t(vi.overlap[a]);
    g.level = k.level;
    return g;
}

template <typename T>
__global__ void fieldBricksKernel(FieldShape s, const Blob* __restrict__ blobs,
                                  const Wave* __restrict__ waves, const Geo* __restrict__ geos,
                                  const int32_t* __restrict__ slots, uint8_t* __restrict__ pool,
                                  size_t slotStride, BrickLayout L)
{
    const Geo g = geos[blockIdx.y];
    T* out = reinterpret_cast<T*>(pool + size_t(slots[blockIdx.y]) * slotStride);
    const uint32_t n = L.dim[0] * L.dim[1] * L.dim[2];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const uint32_t x = i % L.dim[0];
        const uint32_t y = (i / L.dim[0]) % L.dim[1];
        const uint32_t z = i / (L.dim[0] * L.dim[1]);
        const int64_t c[3] = {g.origin[0] + x, g.origin[1] + y, g.origin[2] + z};
        out[tiledIndex(L, x, y, z)] = fieldVoxel<T>(s, blobs, waves, g.level, c);
    }
}
} // namespace
} // namespace lvr

extern "C" {

int lvr_field_brick_host(const lvr_volume_info* info, const lvr_field_params* p,
                         uint64_t node_id, void* out, size_t bytes)
{
    if (!info || !p || !out || p->count < 1 || p->count > 4096 ||
        (p->kind != LVR_FIELD_BLOBS && p->kind != LVR_FIELD_SINES))
        return LVR_EINVAL;
    lvr_volume_info vi = *info;
    if (lvr_fill_regular_volume_info(&vi) != LVR_OK)
        return LVR_EINVAL;
    const int dt = vi.data_type;
    if (dt != LVR_DT_UINT8 && dt != LVR_DT_UINT16 && dt != LVR_DT_FLOAT)
        return LVR_EINVAL;
    const size_t n = size_t(vi.max_block[0]) * vi.max_block[1] * vi.max_block[2];
    if (bytes != n * lvr::bytesPerVoxel(dt) || lvr::decodeNode(node_id).level >= vi.depth)
        return LVR_EINVAL;
    const lvr::FieldSetup f = lvr::makeField(vi, *p);
    const lvr::Geo g = lvr::brickOrigin(vi, node_id);
    size_t i = 0;
    for (uint32_t z = 0; z < vi.max_block[2]; ++z)
        for (uint32_t y = 0; y < vi.max_block[1]; ++y)
            for (uint32_t x = 0; x < vi.max_block[0]; ++x, ++i)
            {
                const int64_t c[3] = {g.origin[0] + x, g.origin[1] + y, g.origin[2] + z};
                if (dt == LVR_DT_UINT8)
                    static_cast<uint8_t*>(out)[i] = lvr::fieldVoxel<uint8_t>(
                        f.shape, f.blobs.data(), f.waves.data(), g.level, c);
                else if (dt == LVR_DT_UINT16)
                    static_cast<uint16_t*>(out)[i] = lvr::fieldVoxel<uint16_t>(
                        f.shape, f.blobs.data(), f.waves.data(), g.level, c);
                else
                    static_cast<float*>(out)[i] = lvr::fieldVoxel<float>(
                        f.shape, f.blobs.data(), f.waves.data(), g.level, c);
            }
    return LVR_OK;
}

int lvr_field_bricks_device(lvr_ctx* ctx, const lvr_field_params* p, const uint64_t* node_ids,
                            int n)
{
    if (!ctx || !p || n < 0 || (n > 0 && !node_ids) || p->count < 1 || p->count > 4096 ||
        (p->kind != LVR_FIELD_BLOBS && p->kind != LVR_FIELD_SINES))
        return LVR_EINVAL;
    const int dt = ctx->vi.data_type;
    if (dt != LVR_DT_UINT8 && dt != LVR_DT_UINT16 && dt != LVR_DT_FLOAT)
        return lvr::fail(ctx, LVR_EINVAL, "field sources support u8, u16, f32");
    int rc = lvr::ensureDevice(ctx);
    if (rc)
        return rc;
    const lvr::FieldSetup f = lvr::makeField(ctx->vi, *p);
    const int chunk = 4096;
    auto up256 = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t blobBytes = up256(sizeof(lvr::Blob) * (f.blobs.size() + 1));
    const size_t waveBytes = up256(sizeof(lvr::Wave) * (f.waves.size() + 1));
    const size_t geoBytes = up256(sizeof(lvr::Geo) * chunk);
    uint8_t* scratch = static_cast<uint8_t*>(
        lvr::genScratch(ctx, blobBytes + waveBytes + geoBytes + sizeof(int32_t) * chunk));
    if (!scratch)
        return LVR_ENOMEM;
    lvr::Blob* d_blobs = reinterpret_cast<lvr::Blob*>(scratch);
    lvr::Wave* d_waves = reinterpret_cast<lvr::Wave*>(scratch + blobBytes);
    lvr::Geo* d_geo = reinterpret_cast<lvr::Geo*>(scratch + blobBytes + waveBytes);
    int32_t* d_slots = reinterpret_cast<int32_t*>(scratch + blobBytes + waveBytes + geoBytes);
    // slots taken here may still be read by batches on other streams
    rc = lvr::waitPoolReaders(ctx, ctx->stream);
    if (rc)
        return rc;
    if ((!f.blobs.empty() &&
         hipMemcpyAsync(d_blobs, f.blobs.data(), sizeof(lvr::Blob) * f.blobs.size(),
                        hipMemcpyHostToDevice, ctx->stream) != hipSuccess) ||
        (!f.waves.empty() &&
         hipMemcpyAsync(d_waves, f.waves.data(), sizeof(lvr::Wave) * f.waves.size(),
                        hipMemcpyHostToDevice, ctx->stream) != hipSuccess))
        return lvr::fail(ctx, LVR_EHIP, "field source upload");
    std::vector<lvr::Geo> geo;
    std::vector<int32_t> slots;
    for (int base = 0; base < n && rc == LVR_OK; base += chunk)
    {
        const int cnt = std::min(chunk, n - base);
        geo.resize(size_t(cnt));
        slots.resize(size_t(cnt));
        int taken = 0;
        for (int i = 0; i < cnt && rc == LVR_OK; ++i)
        {
            if (lvr::decodeNode(node_ids[base + i]).level >= ctx->vi.depth)
                rc = lvr::fail(ctx, LVR_EINVAL, "node id outside the octree");
            else
            {
                geo[size_t(i)] = lvr::brickOrigin(ctx->vi, node_ids[base + i]);
                rc = lvr::acquireSlotFor(ctx, node_ids[base + i], &slots[size_t(i)]);
                taken += rc == LVR_OK;
            }
        }
        if (rc == LVR_OK &&
            (hipMemcpyAsync(d_geo, geo.data(), sizeof(lvr::Geo) * cnt, hipMemcpyHostToDevice,
                            ctx->stream) != hipSuccess ||
             hipMemcpyAsync(d_slots, slots.data(), sizeof(int32_t) * cnt, hipMemcpyHostToDevice,
                            ctx->stream) != hipSuccess))
            rc = lvr::fail(ctx, LVR_EHIP, "field source upload");
        if (rc)
        {
            // bricks of this chunk never received their voxels
            const std::string msg = ctx->error;
            for (int i = 0; i < taken; ++i)
                (void)lvr::releaseSlot(ctx, node_ids[base + i]);
            ctx->error = msg;
            break;
        }
        const dim3 grid(64, unsigned(cnt));
        if (dt == LVR_DT_UINT8)
            hipLaunchKernelGGL(lvr::fieldBricksKernel<uint8_t>, grid, dim3(256), 0, ctx->stream,
                               f.shape, d_blobs, d_waves, d_geo, d_slots, ctx->d_pool,
                               ctx->slotStride, ctx->layout);
        else if (dt == LVR_DT_UINT16)
            hipLaunchKernelGGL(lvr::fieldBricksKernel<uint16_t>, grid, dim3(256), 0, ctx->stream,
                               f.shape, d_blobs, d_waves, d_geo, d_slots, ctx->d_pool,
                               ctx->slotStride, ctx->layout);
        else
            hipLaunchKernelGGL(lvr::fieldBricksKernel<float>, grid, dim3(256), 0, ctx->stream,
                               f.shape, d_blobs, d_waves, d_geo, d_slots, ctx->d_pool,
                               ctx->slotStride, ctx->layout);
        if (hipGetLastError() != hipSuccess)
            rc = lvr::fail(ctx, LVR_EHIP, "field source kernel");
        // each brick's histogram into the cache as it lands in its slot
        if (rc == LVR_OK)
            rc = lvr::cacheBrickHistograms(ctx, node_ids + base, slots.data(), cnt, ctx->stream);
        // the host vectors and the scratch are reused by the next chunk
        if (rc == LVR_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = lvr::fail(ctx, LVR_EHIP, "field source kernel");
    }
    return rc;
}
}
