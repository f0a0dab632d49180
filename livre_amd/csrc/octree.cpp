// Octree shape, node boxes, frustum and brick ordering: the host part of the
// render path that the reference computes on the CPU per frame.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "lvr_internal.h"
#include "radix.h"

namespace lvr
{
size_t bytesPerVoxel(int dataType)
{
    switch (dataType)
    {
    case LVR_DT_UINT8:
    case LVR_DT_INT8:
        return 1;
    case LVR_DT_UINT16:
    case LVR_DT_INT16:
        return 2;
    case LVR_DT_UINT32:
    case LVR_DT_INT32:
    case LVR_DT_FLOAT:
        return 4;
    default:
        return 0;
    }
}

namespace
{
inline float at(const float* m, int row, int col)
{
    return m[col * 4 + row];
}

// General 4x4 inverse by cofactors in double precision, rounded to float.
bool inverse4(const float* src, float* dst)
{
    double m[16], inv[16];
    for (int i = 0; i < 16; ++i)
        m[i] = src[i];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] +
             m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] +
             m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] +
             m[13] * m[3] * m[10];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] +
             m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] -
             m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] +
             m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] +
             m[12] * m[7] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] +
             m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] -
             m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] +
             m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] +
             m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] -
             m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] +
              m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] -
              m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] +
              m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] +
              m[12] * m[6] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] +
              m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] -
              m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] +
              m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] +
                       m[3] * inv[12];
    if (det == 0.0 || !std::isfinite(det))
        return false;
    for (int i = 0; i < 16; ++i)
        dst[i] = float(inv[i] / det);
    return true;
}
} // namespace
} // namespace lvr

extern "C" {

int lvr_fill_regular_volume_info(lvr_volume_info* info)
{
    if (!info || lvr::bytesPerVoxel(info->data_type) == 0)
        return LVR_EINVAL;
    uint32_t block[3];
    for (int i = 0; i < 3; ++i)
    {
        if (info->max_block[i] <= 2u * info->overlap[i] || info->voxels[i] == 0)
            return LVR_EINVAL;
        block[i] = info->max_block[i] - 2u * info->overlap[i];
    }
    // worldSpacePerVoxel = 1 / max(voxels); worldSize = voxels * that
    const uint32_t vmax =
        std::max(info->voxels[0], std::max(info->voxels[1], info->voxels[2]));
    info->world_space_per_voxel = 1.0f / float(vmax);
    uint32_t levels = ~0u;
    for (int i = 0; i < 3; ++i)
    {
        info->world_size[i] = float(info->voxels[i]) * info->world_space_per_voxel;
        const uint32_t nBlocks =
            uint32_t(std::ceil(float(info->voxels[i]) / float(block[i])));
        const uint32_t l = uint32_t(std::ceil(std::log2(double(nBlocks))));
        levels = std::min(levels, l);
    }
    for (int i = 0; i < 3; ++i)
        info->root_blocks[i] = uint32_t(
            std::ceil(float(info->voxels[i] >> levels) / float(block[i])));
    info->depth = levels + 1;
    if (info->depth > 15)
        return LVR_EINVAL; // NodeId holds 4 level bits, 15 = invalid
    return LVR_OK;
}

int lvr_node_world_box(const lvr_volume_info* info, uint64_t node_id,
                       float box[6])
{
    if (!info || !box)
        return LVR_EINVAL;
    const lvr::NodeKey k = lvr::decodeNode(node_id);
    if (k.level >= info->depth)
        return LVR_EINVAL;
    lvr::nodeWorldBox(*info, k, box, box + 3);
    return LVR_OK;
}

// Frustum::Impl (livre/data/Frustum.cpp:24-42) over vmmlib's Frustumf(P)
// and FrustumCullerf(P * MV); vmmlib f4b3706 is not vendored, its published
// formulas are used: near = P23 / (P22 - 1), far = P23 / (P22 + 1),
// top/bottom = near * (P12 +/- 1) / P11, left/right = near * (P02 -/+ 1) /
// P00, and normalized Gribb-Hartmann planes from the rows of P * MV.
int lvr_frustum_init(lvr_frustum* f, const float mv[16], const float proj[16])
{
    using lvr::at;
    if (!f || !mv || !proj)
        return LVR_EINVAL;
    std::memcpy(f->mv, mv, sizeof f->mv);
    std::memcpy(f->proj, proj, sizeof f->proj);
    if (!lvr::inverse4(mv, f->inv_mv) || !lvr::inverse4(proj, f->inv_proj))
        return LVR_EINVAL;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
        {
            float s = 0.f;
            for (int k = 0; k < 4; ++k)
                s += at(proj, i, k) * at(mv, k, j);
            f->mvp[j * 4 + i] = s;
        }
    f->near_plane = at(proj, 2, 3) / (at(proj, 2, 2) - 1.0f);
    f->far_plane = at(proj, 2, 3) / (at(proj, 2, 2) + 1.0f);
    f->bottom = f->near_plane * (at(proj, 1, 2) - 1.0f) / at(proj, 1, 1);
    f->top = f->near_plane * (at(proj, 1, 2) + 1.0f) / at(proj, 1, 1);
    f->left = f->near_plane * (at(proj, 0, 2) - 1.0f) / at(proj, 0, 0);
    f->right = f->near_plane * (at(proj, 0, 2) + 1.0f) / at(proj, 0, 0);
    const float* m = f->mvp;
    const int rowOf[6] = {0, 0, 1, 1, 2, 2};
    const float sign[6] = {1.f, -1.f, 1.f, -1.f, 1.f, -1.f};
    for (int p = 0; p < 6; ++p)
    {
        float* pl = f->planes[p];
        for (int j = 0; j < 4; ++j)
        {
            const float r3 = at(m, 3, j), ri = at(m, rowOf[p], j);
            pl[j] = sign[p] > 0.f ? r3 + ri : r3 - ri;
        }
        const float len = std::sqrt(pl[0] * pl[0] + pl[1] * pl[1] + pl[2] * pl[2]);
        for (int j = 0; j < 4; ++j)
            pl[j] /= len;
    }
    for (int i = 0; i < 3; ++i)
    {
        f->eye[i] = f->inv_mv[12 + i];
        f->view_dir[i] = f->inv_mv[8 + i];
    }
    return LVR_OK;
}

// RayCastRenderer::order with DistanceOperator (RayCastRenderer.cpp:57-81,
// 153-159): std::sort of the given order by |MV * centre|. The distances are
// computed once; the comparisons, and so libstdc++'s tie handling, are the
// same as recomputing them per comparison.
int lvr_order(const lvr_volume_info* info, const lvr_frustum* f,
              const uint64_t* ids, int n, uint64_t* out)
{
    if (!info || !f || (n > 0 && (!ids || !out)) || n < 0)
        return LVR_EINVAL;
    std::vector<std::pair<float, uint64_t>> v(static_cast<size_t>(n));
    std::vector<std::pair<uint32_t, uint64_t>> bits(static_cast<size_t>(n)), tmp;
    bool plain = true; // every distance a non-negative number (not NaN)
    for (int i = 0; i < n; ++i)
    {
        float mn[3], mx[3];
        lvr::nodeWorldBox(*info, lvr::decodeNode(ids[i]), mn, mx);
        float c[3];
        for (int a = 0; a < 3; ++a)
            c[a] = (mn[a] + mx[a]) * 0.5f;
        float r[4];
        for (int a = 0; a < 4; ++a)
            r[a] = lvr::at(f->mv, a, 0) * c[0] + lvr::at(f->mv, a, 1) * c[1] +
                   lvr::at(f->mv, a, 2) * c[2] + lvr::at(f->mv, a, 3);
        const float x = r[0] / r[3], y = r[1] / r[3], z = r[2] / r[3];
        const float d = std::sqrt(x * x + y * y + z * z);
        v[size_t(i)] = {d, ids[i]};
        uint32_t u;
        std::memcpy(&u, &d, sizeof u);
        plain = plain && !(d < 0.0f) && u <= 0x7f800000u;
        bits[size_t(i)] = {u, ids[i]};
    }
    // Non-negative floats order as their bit patterns. With all distances
    // distinct the sorted order is unique, so the radix sort gives exactly
    // what std::sort gives; equal distances keep std::sort's tie order.
    if (plain)
    {
        lvr::radixSortByKey(bits, tmp);
        if (!lvr::hasEqualKeys(bits))
        {
            for (int i = 0; i < n; ++i)
                out[i] = bits[size_t(i)].second;
            return LVR_OK;
        }
    }
    std::sort(v.begin(), v.end(),
              [](const std::pair<float, uint64_t>& a,
                 const std::pair<float, uint64_t>& b) { return a.first < b.first; });
    for (int i = 0; i < n; ++i)
        out[i] = v[size_t(i)].second;
    return LVR_OK;
}

static int splitAxes(int nParts)
{
    switch (nParts)
    {
    case 1: return 0;
    case 2: return 1;
    case 4: return 2;
    case 8: return 3;
    default: return -1;
    }
}

int lvr_sortlast_parts(const lvr_volume_info* info, const uint64_t* ids, int n, int n_parts,
                       int32_t* part_of)
{
    const int axes = splitAxes(n_parts);
    if (!info || n < 0 || (n > 0 && (!ids || !part_of)) || axes < 0 || info->depth < 1)
        return LVR_EINVAL;
    for (int i = 0; i < n; ++i)
    {
        const lvr::NodeKey k = lvr::decodeNode(ids[i]);
        if (k.level >= info->depth)
            return LVR_EINVAL;
        // blocks per axis at the node's level; the split is at half of them:
        // low half if 2 (pos + 1) <= blocks, high half if 2 pos >= blocks
        const uint32_t pos[3] = {k.x, k.y, k.z};
        int32_t part = 0;
        for (int a = 0; a < axes; ++a)
        {
            const uint64_t blocks = uint64_t(info->root_blocks[a]) << k.level;
            const uint64_t p = pos[a];
            int bit;
            if (2 * (p + 1) <= blocks)
                bit = 0;
            else if (2 * p >= blocks)
                bit = 1;
            else
            {
                part = -1;
                break;
            }
            part = (part << 1) | bit;
        }
        part_of[i] = part;
    }
    return LVR_OK;
}

int lvr_sortlast_order(const lvr_volume_info* info, const float eye[3], int n_parts,
                       int32_t* order)
{
    const int axes = splitAxes(n_parts);
    if (!info || !eye || !order || axes < 0 || info->depth < 1)
        return LVR_EINVAL;
    // the split planes in world space: the minimum corner of the level-1
    // node at the middle block position (internalNodeToLODNode's box)
    lvr::NodeKey mid;
    mid.level = 1;
    mid.x = info->root_blocks[0];
    mid.y = info->root_blocks[1];
    mid.z = info->root_blocks[2];
    mid.t = 0;
    float mn[3], mx[3];
    lvr::nodeWorldBox(*info, mid, mn, mx);
    int nearBit[3];
    for (int a = 0; a < 3; ++a)
        nearBit[a] = eye[a] < mn[a] ? 0 : 1;
    // BSP traversal: near half before far half at every split, x outermost
    for (int r = 0; r < n_parts; ++r)
    {
        int32_t part = 0;
        for (int a = 0; a < axes; ++a)
        {
            const int far = (r >> (axes - 1 - a)) & 1;
            part = (part << 1) | (far ? 1 - nearBit[a] : nearBit[a]);
        }
        order[r] = part;
    }
    return LVR_OK;
}

int lvr_samples_per_ray(const lvr_volume_info* info, const uint64_t* ids, int n,
                        int32_t requested, int32_t* out)
{
    if (!info || !out || n < 0 || (n > 0 && !ids) || requested < 0 || info->depth < 1)
        return LVR_EINVAL;
    if (requested > 0)
    {
        *out = requested;
        return LVR_OK;
    }
    // RayCastRenderer::Impl::onFrameStart (RayCastRenderer.cpp:216-234): the
    // finest level among the rendered bricks, its voxel count along the
    // longest axis, at least minSamplesPerRay = 512; the float is stored to
    // an unsigned integer (truncated)
    uint32_t maxLod = 0;
    for (int i = 0; i < n; ++i)
    {
        const uint32_t level = lvr::decodeNode(ids[i]).level;
        if (level >= info->depth)
            return LVR_EINVAL;
        maxLod = std::max(maxLod, level);
    }
    const float maxVoxelDim =
        float(std::max(info->voxels[0], std::max(info->voxels[1], info->voxels[2])));
    const float maxVoxelsAtLod = maxVoxelDim / float(1u << (info->depth - maxLod - 1));
    *out = int32_t(uint32_t(std::max(maxVoxelsAtLod, 512.0f)));
    return LVR_OK;
}
}
