/*
 * livre_oracle.cpp — CPU restatement of the Livre render hot path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker that the HIP product is
 * compared against, and the CPU baseline that bench.py times. Nothing in
 * livre_amd/ links or loads it.
 *
 * Built with -ffp-contract=off so that every FP32 expression rounds exactly as
 * written, in the order the reference writes it (no fused multiply-add).
 *
 * Parity pinning (what checks each function):
 *   orc_fill_volume_info   tests/data/volumeInformation.cpp:57-106,
 *                          tests/data/dataSource.cpp:46 (golden fixtures)
 *   orc_select_visibles    tests/data/lodSelection.cpp:70-207 (seven golden
 *                          NodeId sets, including FP32 knife-edge nodes);
 *                          clip planes by brute force over the box corners
 *                          and the march's own clipping (Lexis' isOutside
 *                          itself is not vendored: parity unpinned there)
 *   orc_order              brute-force property (non-decreasing eye distance)
 *   orc_render_*           closed forms on tiny scenes (empty TF, constant
 *                          volume series, brick-split invariance, exit
 *                          threshold bound) and the FP64 instance
 *   orc_histogram          tests/lib/cache.cpp:106-134 (max index 17,
 *                          count 2^24) and brute-force counting
 *   orc_default_tf/lut     tests/core/transferFunction.cpp:31-72 + tf_f.1dt
 */
#include "livre_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace
{
/* ---------------------------------------------------------------- NodeId */
/* 64-bit key, GCC bitfield order from the LSB (livre/data/NodeId.h:34-48,
 * livre/data/types.h:80-93): level 4 bits, x/y/z 14 bits each, time 18. */
struct Node
{
    uint32_t level, x, y, z, t;
};

inline Node decode(uint64_t id)
{
    Node n;
    n.level = uint32_t(id & 0xF);
    n.x = uint32_t((id >> 4) & 0x3FFF);
    n.y = uint32_t((id >> 18) & 0x3FFF);
    n.z = uint32_t((id >> 32) & 0x3FFF);
    n.t = uint32_t(id >> 46);
    return n;
}

inline uint64_t encode(uint32_t level, uint32_t x, uint32_t y, uint32_t z,
                       uint32_t t)
{
    return uint64_t(level & 0xF) | (uint64_t(x & 0x3FFF) << 4) |
           (uint64_t(y & 0x3FFF) << 18) | (uint64_t(z & 0x3FFF) << 32) |
           (uint64_t(t & 0x3FFFF) << 46);
}

/* ------------------------------------------------------------- LODNode */
struct Box
{
    float mn[3], mx[3];
};

/* DataSourcePlugin::internalNodeToLODNode (DataSourcePlugin.cpp:40-71) */
Box world_box(const orc_volume_info& vi, const Node& n)
{
    uint32_t bricks[3];
    for (int i = 0; i < 3; ++i)
        bricks[i] = vi.root_blocks[i] * (1u << n.level);
    /* vmmlib find_max_index: first index holding the maximum */
    int index = 0;
    for (int i = 1; i < 3; ++i)
        if (bricks[i] > bricks[index])
            index = i;
    const uint32_t pos[3] = {n.x, n.y, n.z};
    Box b;
    for (int i = 0; i < 3; ++i)
    {
        const float cmin = float(pos[i]) / float(bricks[index]);
        const float cmax = float(pos[i] + 1u) / float(bricks[index]);
        const float half = vi.world_size[i] * 0.5f;
        b.mn[i] = cmin * vi.world_size[i] - half;
        b.mx[i] = cmax * vi.world_size[i] - half;
    }
    return b;
}

/* -------------------------------------------------------------- matrices */
inline float m_at(const float* m, int row, int col) { return m[col * 4 + row]; }

/* vmmlib Matrix * Matrix: r(i,j) = sum_k a(i,k) b(k,j), k ascending. */
void mat_mul(const float* a, const float* b, float* r)
{
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
        {
            float s = 0.f;
            for (int k = 0; k < 4; ++k)
                s += m_at(a, i, k) * m_at(b, k, j);
            r[j * 4 + i] = s;
        }
}

/* ---------------------------------------------------------------- Frustum */
/* livre::Frustum (Frustum.cpp:24-76) over vmmlib Frustumf(projection) and
 * FrustumCullerf(P * MV). vmmlib is not vendored (pinned f4b3706,
 * .gitsubprojects:2); its published algorithms are restated:
 *  - Frustum(projection): near = P23/(P22-1), far = P23/(P22+1),
 *    bottom/top = near*(P12-/+1)/P11, left/right = near*(P02-/+1)/P00
 *  - FrustumCuller: Gribb-Hartmann planes from the rows of P*MV,
 *    left = r3+r0, right = r3-r0, bottom = r3+r1, top = r3-r1,
 *    near = r3+r2, far = r3-r2, each divided by the length of its normal.
 * The golden LOD sets (tests/data/lodSelection.cpp) pin this formulation:
 * several nodes sit exactly at the SSE threshold in real arithmetic. */
struct Frustum
{
    float nearp, farp, left, right, bottom, top;
    float planes[6][4]; /* left, right, bottom, top, near, far */
    float eye[3];
    float dir[3];
};

void normalize_plane(float* p)
{
    const float len = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    p[0] /= len;
    p[1] /= len;
    p[2] /= len;
    p[3] /= len;
}

/* 4x4 inverse in double, rounded to float (only eye position and view
 * direction of the selection use it, see Frustum.cpp:36-41). */
void invert(const float* m, double* inv)
{
    double a[16];
    for (int i = 0; i < 16; ++i)
        a[i] = m[i];
    double r[16];
    r[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] +
           a[9] * a[7] * a[14] + a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
    r[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] -
           a[8] * a[7] * a[14] - a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
    r[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] +
           a[8] * a[7] * a[13] + a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
    r[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] -
            a[8] * a[6] * a[13] - a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
    r[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] -
           a[9] * a[3] * a[14] - a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
    r[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] +
           a[8] * a[3] * a[14] + a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
    r[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] -
           a[8] * a[3] * a[13] - a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
    r[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] +
            a[8] * a[2] * a[13] + a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
    r[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] +
           a[5] * a[3] * a[14] + a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
    r[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] -
           a[4] * a[3] * a[14] - a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
    r[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] +
            a[4] * a[3] * a[13] + a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
    r[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] -
            a[4] * a[2] * a[13] - a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
    r[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] -
           a[5] * a[3] * a[10] - a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
    r[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] +
           a[4] * a[3] * a[10] + a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
    r[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] -
            a[4] * a[3] * a[9] - a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
    r[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] +
            a[4] * a[2] * a[9] + a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
    const double det = a[0] * r[0] + a[1] * r[4] + a[2] * r[8] + a[3] * r[12];
    for (int i = 0; i < 16; ++i)
        inv[i] = r[i] / det;
}

Frustum make_frustum(const float* mv, const float* proj)
{
    Frustum f;
    const float p22 = m_at(proj, 2, 2), p23 = m_at(proj, 2, 3);
    f.nearp = p23 / (p22 - 1.0f);
    f.farp = p23 / (p22 + 1.0f);
    f.bottom = f.nearp * (m_at(proj, 1, 2) - 1.0f) / m_at(proj, 1, 1);
    f.top = f.nearp * (m_at(proj, 1, 2) + 1.0f) / m_at(proj, 1, 1);
    f.left = f.nearp * (m_at(proj, 0, 2) - 1.0f) / m_at(proj, 0, 0);
    f.right = f.nearp * (m_at(proj, 0, 2) + 1.0f) / m_at(proj, 0, 0);

    float mvp[16];
    mat_mul(proj, mv, mvp);
    float row[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            row[i][j] = m_at(mvp, i, j);
    for (int j = 0; j < 4; ++j)
    {
        f.planes[0][j] = row[3][j] + row[0][j];
        f.planes[1][j] = row[3][j] - row[0][j];
        f.planes[2][j] = row[3][j] + row[1][j];
        f.planes[3][j] = row[3][j] - row[1][j];
        f.planes[4][j] = row[3][j] + row[2][j];
        f.planes[5][j] = row[3][j] - row[2][j];
    }
    for (int i = 0; i < 6; ++i)
        normalize_plane(f.planes[i]);

    double inv[16];
    invert(mv, inv);
    /* eye = translation of inverse MV; dir = its column 2 */
    for (int i = 0; i < 3; ++i)
    {
        f.eye[i] = float(inv[12 + i]);
        f.dir[i] = float(inv[8 + i]);
    }
    return f;
}

inline float plane_dot(const float* p, const float* v)
{
    return p[0] * v[0] + p[1] * v[1] + p[2] * v[2] + p[3] * 1.0f;
}

/* vmmlib AABB::computeNearFar: per axis the corner nearest to / farthest
 * from the plane along its normal. */
void near_far(const Box& b, const float* plane, float* vmin, float* vmax)
{
    for (int i = 0; i < 3; ++i)
    {
        if (plane[i] >= 0.f)
        {
            vmin[i] = b.mn[i];
            vmax[i] = b.mx[i];
        }
        else
        {
            vmin[i] = b.mx[i];
            vmax[i] = b.mn[i];
        }
    }
}

/* vmmlib FrustumCuller::test(AABB) != VISIBILITY_NONE: a box is outside when
 * its corner farthest along some plane normal is behind that plane. */
bool in_frustum(const Frustum& f, const Box& b)
{
    for (int i = 0; i < 6; ++i)
    {
        float vmin[3], vmax[3];
        near_far(b, f.planes[i], vmin, vmax);
        if (plane_dot(f.planes[i], vmax) < 0.f)
            return false;
    }
    return true;
}

/* ---------------------------------------------------------- SelectVisibles */
struct Selector
{
    const orc_volume_info* vi;
    Frustum f;
    uint32_t height;
    float sse;
    uint32_t minLod, maxLod;
    int nClip = 0;
    const float* clip = nullptr; /* n_clip planes: normal xyz, d */
    std::vector<uint64_t> visibles;

    /* SelectVisibles::Impl::isLODVisible (SelectVisibles.cpp:45-64) */
    bool lod_visible(const float* worldCoord, float worldSpacePerVoxel) const
    {
        const float t = f.top;
        const float b = f.bottom;
        const float worldSpacePerPixel = (t - b) / float(height);
        const float pixelPerVoxel = worldSpacePerVoxel / worldSpacePerPixel;
        const float distance = std::abs(plane_dot(f.planes[4], worldCoord));
        const float n = f.nearp;
        const float pixelPerVoxelInDistance = pixelPerVoxel * n / (n + distance);
        return pixelPerVoxelInDistance <= sse;
    }

    /* SelectVisibles::Impl::visit (SelectVisibles.cpp:66-107) */
    bool visit(uint64_t id)
    {
        const Node n = decode(id);
        const Box box = world_box(*vi, n);
        if (!in_frustum(f, box))
            return false;
        /* _clipPlanes.isOutside(worldBox) (SelectVisibles.cpp:69): Lexis'
         * ClipPlanes is not under the reference tree, so the rule restated
         * here is the half-space fragRayCast.glsl:152-167 clips away: the
         * box is outside a plane when even its corner furthest along the
         * normal has n . x + d < 0. */
        for (int c = 0; c < nClip; ++c)
        {
            const float* cp = clip + 4 * c;
            float vmin[3], vmax[3];
            near_far(box, cp, vmin, vmax);
            if (plane_dot(cp, vmax) < 0.f)
                return false;
        }

        float vmin[3], vmax[3];
        near_far(box, f.planes[4], vmin, vmax);
        if (plane_dot(f.planes[4], vmin) < 0.f ||
            plane_dot(f.planes[4], vmax) < 0.f)
        {
            for (int i = 0; i < 3; ++i)
                vmin[i] = f.eye[i] - f.dir[i] * f.nearp;
        }
        float wspv = 0.f;
        for (int i = 0; i < 3; ++i)
        {
            const uint32_t block = vi->max_block[i] - 2u * vi->overlap[i];
            const float size = box.mx[i] - box.mn[i];
            const float v = size / float(block);
            wspv = (i == 0) ? v : std::min(wspv, v);
        }
        bool visible = lod_visible(vmin, wspv);
        visible = (visible && n.level >= minLod) || (n.level == maxLod) ||
                  (n.level == vi->depth - 1);
        if (visible)
            visibles.push_back(id);
        return !visible;
    }

    /* DFSTraversal::Impl::traverse (DFSTraversal.cpp:29-43); children in
     * NodeId::getChildren order, x outer, z inner (NodeId.cpp:92-113). */
    void traverse(uint64_t id, uint32_t depth)
    {
        if (!visit(id))
            return;
        if (depth == 1)
            return;
        const Node n = decode(id);
        for (uint32_t x = 0; x < 2; ++x)
            for (uint32_t y = 0; y < 2; ++y)
                for (uint32_t z = 0; z < 2; ++z)
                    traverse(encode(n.level + 1, n.x * 2 + x, n.y * 2 + y,
                                    n.z * 2 + z, n.t),
                             depth - 1);
    }
};

/* -------------------------------------------------------------- ray march */
template <typename T>
struct V3
{
    T x, y, z;
};
template <typename T>
inline V3<T> operator-(V3<T> a, V3<T> b)
{
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <typename T>
inline V3<T> operator+(V3<T> a, V3<T> b)
{
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <typename T>
inline V3<T> operator*(V3<T> a, T s)
{
    return {a.x * s, a.y * s, a.z * s};
}
template <typename T>
inline T length3(V3<T> a)
{
    return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}
template <typename T>
inline V3<T> normalize3(V3<T> a)
{
    const T l = length3(a);
    return {a.x / l, a.y / l, a.z / l};
}
/* GLSL mod(x, y) = x - y * floor(x / y) */
template <typename T>
inline T glsl_mod(T x, T y)
{
    return x - y * std::floor(x / y);
}

/* intersectBox (fragRayCast.glsl:79-101) */
template <typename T>
inline void intersect_box(V3<T> o, V3<T> d, V3<T> mn, V3<T> mx, T& t0, T& t1)
{
    const T eps = T(0.0000000001f);
    if (d.x == 0)
        d.x = eps;
    if (d.y == 0)
        d.y = eps;
    if (d.z == 0)
        d.z = eps;
    const V3<T> inv = {T(1) / d.x, T(1) / d.y, T(1) / d.z};
    const V3<T> tb = {inv.x * (mn.x - o.x), inv.y * (mn.y - o.y),
                      inv.z * (mn.z - o.z)};
    const V3<T> tt = {inv.x * (mx.x - o.x), inv.y * (mx.y - o.y),
                      inv.z * (mx.z - o.z)};
    const V3<T> tmin = {std::min(tt.x, tb.x), std::min(tt.y, tb.y),
                        std::min(tt.z, tb.z)};
    const V3<T> tmax = {std::max(tt.x, tb.x), std::max(tt.y, tb.y),
                        std::max(tt.z, tb.z)};
    t0 = std::max(std::max(tmin.x, tmin.y), std::max(tmin.x, tmin.z));
    t1 = std::min(std::min(tmax.x, tmax.y), std::min(tmax.x, tmax.z));
}

template <typename T>
inline T voxel_value(const void* p, int type, size_t idx)
{
    switch (type)
    {
    case ORC_DT_UINT8:
        return T(static_cast<const uint8_t*>(p)[idx]);
    case ORC_DT_UINT16:
        return T(static_cast<const uint16_t*>(p)[idx]);
    case ORC_DT_UINT32:
        return T(float(static_cast<const uint32_t*>(p)[idx]));
    case ORC_DT_INT8:
        return T(static_cast<const int8_t*>(p)[idx]);
    case ORC_DT_INT16:
        return T(static_cast<const int16_t*>(p)[idx]);
    case ORC_DT_INT32:
        return T(float(static_cast<const int32_t*>(p)[idx]));
    default:
        return T(static_cast<const float*>(p)[idx]);
    }
}

/* GL texture lookup of the brick: GL_NEAREST picks floor(u*N) clamped to
 * [0, N-1] (GL 4.2 core §3.9.11); GL_LINEAR uses u*N - 0.5 with
 * clamp-to-edge. */
template <typename T>
inline T sample_volume(const void* vox, int type, const uint32_t* N,
                       V3<T> tp, bool linear)
{
    if (!linear)
    {
        const T u[3] = {tp.x, tp.y, tp.z};
        long idx[3];
        for (int i = 0; i < 3; ++i)
        {
            long k = long(std::floor(u[i] * T(N[i])));
            k = std::max(0l, std::min(long(N[i]) - 1, k));
            idx[i] = k;
        }
        const size_t lin = (size_t(idx[2]) * N[1] + size_t(idx[1])) * N[0] +
                           size_t(idx[0]);
        return voxel_value<T>(vox, type, lin);
    }
    const T u[3] = {tp.x * T(N[0]) - T(0.5), tp.y * T(N[1]) - T(0.5),
                    tp.z * T(N[2]) - T(0.5)};
    long i0[3], i1[3];
    T w[3];
    for (int i = 0; i < 3; ++i)
    {
        const T fl = std::floor(u[i]);
        w[i] = u[i] - fl;
        const long k = long(fl);
        i0[i] = std::max(0l, std::min(long(N[i]) - 1, k));
        i1[i] = std::max(0l, std::min(long(N[i]) - 1, k + 1));
    }
    auto at = [&](long x, long y, long z) {
        return voxel_value<T>(vox, type,
                              (size_t(z) * N[1] + size_t(y)) * N[0] + size_t(x));
    };
    const T c00 = at(i0[0], i0[1], i0[2]) * (T(1) - w[0]) +
                  at(i1[0], i0[1], i0[2]) * w[0];
    const T c10 = at(i0[0], i1[1], i0[2]) * (T(1) - w[0]) +
                  at(i1[0], i1[1], i0[2]) * w[0];
    const T c01 = at(i0[0], i0[1], i1[2]) * (T(1) - w[0]) +
                  at(i1[0], i0[1], i1[2]) * w[0];
    const T c11 = at(i0[0], i1[1], i1[2]) * (T(1) - w[0]) +
                  at(i1[0], i1[1], i1[2]) * w[0];
    const T c0 = c00 * (T(1) - w[1]) + c10 * w[1];
    const T c1 = c01 * (T(1) - w[1]) + c11 * w[1];
    return c0 * (T(1) - w[2]) + c1 * w[2];
}

/* texture(transferFnTex, density): 1D, GL_LINEAR, CLAMP_TO_EDGE
 * (RayCastRenderer.cpp:183-185). */
template <typename T>
inline void sample_tf(const float* tf, int n, T density, T* out)
{
    const T u = density * T(n) - T(0.5);
    const T fl = std::floor(u);
    const T w = u - fl;
    long k = long(std::max(T(-1), std::min(T(n), fl)));
    const long a = std::max(0l, std::min(long(n) - 1, k));
    const long b = std::max(0l, std::min(long(n) - 1, k + 1));
    for (int c = 0; c < 4; ++c)
        out[c] = T(tf[a * 4 + c]) * (T(1) - w) + T(tf[b * 4 + c]) * w;
}

template <typename T>
inline V3<T> mat_vec3(const float* m, T x, T y, T z, T w, T* wout)
{
    V3<T> r;
    r.x = T(m[0]) * x + T(m[4]) * y + T(m[8]) * z + T(m[12]) * w;
    r.y = T(m[1]) * x + T(m[5]) * y + T(m[9]) * z + T(m[13]) * w;
    r.z = T(m[2]) * x + T(m[6]) * y + T(m[10]) * z + T(m[14]) * w;
    if (wout)
        *wout = T(m[3]) * x + T(m[7]) * y + T(m[11]) * z + T(m[15]) * w;
    return r;
}

struct BrickGeom
{
    Box box;
    const void* vox;
    int x0, y0, x1, y1; /* conservative pixel bbox, inclusive */
};

/* One pixel of fragRayCast.glsl main (lines 113-207) applied to every brick
 * of the ordered list whose back face covers the pixel, in list order, as
 * the per-brick draws with a memory barrier after each brick do
 * (RayCastRenderer.cpp:467-501, 566-568). */
template <typename T>
void shade_pixel(const orc_volume_info& vi, const orc_camera& cam,
                 const std::vector<BrickGeom>& bricks,
                 const std::vector<int>& list, const float* tf, int ntf,
                 const orc_render_params& p, int px, int py, T* rgba,
                 uint64_t& samples)
{
    const T ndcx = (T(px) + T(0.5)) / T(cam.width) * T(2) - T(1);
    const T ndcy = (T(py) + T(0.5)) / T(cam.height) * T(2) - T(1);
    T w;
    V3<T> eyePos = mat_vec3<T>(cam.inv_proj, ndcx, ndcy, T(-1), T(1), &w);
    eyePos = {eyePos.x / w, eyePos.y / w, eyePos.z / w};

    const V3<T> eye = {T(cam.eye[0]), T(cam.eye[1]), T(cam.eye[2])};
    const V3<T> pixelWorld =
        mat_vec3<T>(cam.inv_mv, eyePos.x, eyePos.y, eyePos.z, T(1), nullptr);
    const V3<T> dir = normalize3(pixelWorld - eye);
    const V3<T> eyeDirN = normalize3(eyePos);
    /* dot((0,0,1),(0,0,-n)) / dot((0,0,1), normalize(eyePos)) */
    const T tNearPlane = -T(cam.near_plane) / eyeDirN.z;
    const T viewCos = -eyeDirN.z;

    const T half[3] = {T(vi.world_size[0]) / T(2), T(vi.world_size[1]) / T(2),
                       T(vi.world_size[2]) / T(2)};
    const V3<T> gmin = {-half[0], -half[1], -half[2]};
    const V3<T> gmax = {half[0], half[1], half[2]};
    T tnearGlobal, tfarGlobal;
    intersect_box(eye, dir, gmin, gmax, tnearGlobal, tfarGlobal);

    const T stepSize = T(1) / T(p.n_samples_per_ray);
    const T alphaCorrection =
        T(p.max_samples_per_ray) / T(p.n_samples_per_ray);
    const T multiplyer = T(1) / (T(p.range1) - T(p.range0));
    const T addedValue = -T(p.range0) / (T(p.range1) - T(p.range0));
    const T exitA = T(p.early_exit);
    const T maxA = T(1) - T(1) / T(256);

    T texMin[3], texMax[3];
    for (int i = 0; i < 3; ++i)
    {
        /* TextureObject::Impl::initialize (TextureObject.cpp:76-92) */
        const float ov = float(vi.overlap[i]) / float(vi.max_block[i]);
        const float sz =
            float(vi.max_block[i] - 2u * vi.overlap[i]) / float(vi.max_block[i]);
        texMin[i] = T(ov);
        texMax[i] = T(ov + sz);
    }

    T acc[4] = {0, 0, 0, 0};
    for (int bi : list)
    {
        const BrickGeom& g = bricks[bi];
        if (!g.vox)
            continue;
        if (acc[3] > exitA) /* imageLoad + discard */
            break;
        const V3<T> bmin = {T(g.box.mn[0]), T(g.box.mn[1]), T(g.box.mn[2])};
        const V3<T> bmax = {T(g.box.mx[0]), T(g.box.mx[1]), T(g.box.mx[2])};
        T tnear, tfar;
        intersect_box(eye, dir, bmin, bmax, tnear, tfar);
        /* rasterisation of the back faces: a fragment exists where the exit
         * point lies between the near and far clip planes */
        const T exitDepth = tfar * viewCos;
        if (!(exitDepth >= T(cam.near_plane) && exitDepth <= T(cam.far_plane)))
            continue;
        if (tnear < tNearPlane)
            tnear = tNearPlane;
        const T residu = glsl_mod(tnear - tnearGlobal, stepSize);
        if (residu > T(0))
            tnear += stepSize - residu;
        if (tnear > tfar)
            continue;
        for (int c = 0; c < p.n_clip_planes; ++c)
        {
            const float* cp = p.clip_planes + 4 * c;
            const V3<T> pn = {T(cp[0]), T(cp[1]), T(cp[2])};
            T rn = dir.x * pn.x + dir.y * pn.y + dir.z * pn.z;
            if (rn == 0)
                rn = T(0.0000000001f);
            const T d = T(cp[3]);
            const T t = -((pn.x * eye.x + pn.y * eye.y + pn.z * eye.z) + d) / rn;
            if (rn > T(0))
                tnear = std::max(tnear, t);
            else
                tfar = std::min(tfar, t);
        }
        if (tnear > tfar)
            continue;

        const V3<T> rayStart = eye + dir * tnear;
        const V3<T> rayStop = eye + dir * tfar;
        V3<T> pos = rayStart;
        const V3<T> step = normalize3(rayStop - rayStart) * stepSize;
        const V3<T> extent = bmax - bmin;
        const V3<T> texExt = {texMax[0] - texMin[0], texMax[1] - texMin[1],
                              texMax[2] - texMin[2]};
        for (T travel = length3(rayStop - rayStart); travel > T(0);
             pos = pos + step, travel -= stepSize)
        {
            /* calcTexturePositionFromAABBPos (fragRayCast.glsl:72-76) */
            const V3<T> tp = {
                (pos.x - bmin.x) / extent.x * texExt.x + texMin[0],
                (pos.y - bmin.y) / extent.y * texExt.y + texMin[1],
                (pos.z - bmin.z) / extent.z * texExt.z + texMin[2]};
            const T value = sample_volume<T>(g.vox, vi.data_type, vi.max_block,
                                             tp, p.linear_filter != 0);
            const T density = value * multiplyer + addedValue;
            T src[4];
            sample_tf<T>(tf, ntf, density, src);
            /* composite (fragRayCast.glsl:103-111) */
            const T alpha =
                T(1) - std::pow(T(1) - std::min(src[3], maxA), alphaCorrection);
            acc[0] += src[0] * alpha * (T(1) - acc[3]);
            acc[1] += src[1] * alpha * (T(1) - acc[3]);
            acc[2] += src[2] * alpha * (T(1) - acc[3]);
            acc[3] += alpha * (T(1) - acc[3]);
            ++samples;
            if (acc[3] > exitA)
                break;
        }
    }
    for (int c = 0; c < 4; ++c)
        rgba[c] = acc[c];
}

template <typename T>
int render(const orc_volume_info* vi, const orc_camera* cam,
           const uint64_t* ordered, int n, const void* const* brick_ptrs,
           const float* tf, int ntf, const orc_render_params* p,
           float* out_rgba, uint64_t* samples_out, int n_threads)
{
    const int W = cam->width, H = cam->height;
    /* projection of the brick corners for conservative screen bounds:
     * clip = P * MV * x = inverse(invMV * invP) ... computed from the
     * inverses in double to stay conservative */
    double invp[16], invmv[16];
    {
        float tmp[16];
        std::memcpy(tmp, cam->inv_proj, sizeof tmp);
        invert(tmp, invp);
        std::memcpy(tmp, cam->inv_mv, sizeof tmp);
        invert(tmp, invmv);
    }
    auto mul4 = [](const double* m, const double* v, double* r) {
        for (int i = 0; i < 4; ++i)
            r[i] = m[i] * v[0] + m[4 + i] * v[1] + m[8 + i] * v[2] +
                   m[12 + i] * v[3];
    };
    std::vector<BrickGeom> bricks(n);
    for (int i = 0; i < n; ++i)
    {
        BrickGeom& g = bricks[i];
        g.box = world_box(*vi, decode(ordered[i]));
        g.vox = brick_ptrs ? brick_ptrs[i] : nullptr;
        double mnx = 1e30, mny = 1e30, mxx = -1e30, mxy = -1e30;
        bool full = false;
        for (int c = 0; c < 8; ++c)
        {
            const double v[4] = {(c & 1) ? g.box.mx[0] : g.box.mn[0],
                                 (c & 2) ? g.box.mx[1] : g.box.mn[1],
                                 (c & 4) ? g.box.mx[2] : g.box.mn[2], 1.0};
            double e[4], clip[4];
            mul4(invmv, v, e);
            mul4(invp, e, clip);
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
        if (full)
        {
            g.x0 = 0;
            g.y0 = 0;
            g.x1 = W - 1;
            g.y1 = H - 1;
        }
        else
        {
            g.x0 = int(std::max(0.0, std::floor(mnx) - 2));
            g.y0 = int(std::max(0.0, std::floor(mny) - 2));
            g.x1 = int(std::min(double(W - 1), std::ceil(mxx) + 2));
            g.y1 = int(std::min(double(H - 1), std::ceil(mxy) + 2));
        }
    }
    const int TS = 16;
    const int tx = (W + TS - 1) / TS, ty = (H + TS - 1) / TS;
    int wx0 = 0, wy0 = 0, wx1 = W, wy1 = H;
    if (p->window[2] > p->window[0] && p->window[3] > p->window[1])
    {
        wx0 = std::max(0, p->window[0]);
        wy0 = std::max(0, p->window[1]);
        wx1 = std::min(W, p->window[2]);
        wy1 = std::min(H, p->window[3]);
    }
    uint64_t total = 0;
#ifdef _OPENMP
    if (n_threads <= 0)
        n_threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads) reduction(+ : total)
#endif
    for (int t = 0; t < tx * ty; ++t)
    {
        const int x0 = (t % tx) * TS, y0 = (t / tx) * TS;
        const int x1 = std::min(W, x0 + TS) - 1, y1 = std::min(H, y0 + TS) - 1;
        if (x1 < wx0 || x0 >= wx1 || y1 < wy0 || y0 >= wy1)
            continue;
        std::vector<int> list;
        for (int i = 0; i < n; ++i)
        {
            const BrickGeom& g = bricks[i];
            if (g.x1 < x0 || g.x0 > x1 || g.y1 < y0 || g.y0 > y1)
                continue;
            list.push_back(i);
        }
        uint64_t s = 0;
        for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x)
            {
                if (x < wx0 || x >= wx1 || y < wy0 || y >= wy1)
                    continue;
                T rgba[4];
                shade_pixel<T>(*vi, *cam, bricks, list, tf, ntf, *p, x, y, rgba,
                               s);
                float* o = out_rgba + (size_t(y) * W + x) * 4;
                for (int c = 0; c < 4; ++c)
                    o[c] = float(rgba[c]);
            }
        total += s;
    }
    if (samples_out)
        *samples_out = total;
    return 0;
}

size_t bytes_per_voxel(int type)
{
    switch (type)
    {
    case ORC_DT_UINT8:
    case ORC_DT_INT8:
        return 1;
    case ORC_DT_UINT16:
    case ORC_DT_INT16:
        return 2;
    default:
        return 4;
    }
}
} // namespace

extern "C" {

/* fillRegularVolumeInfo (DataSourcePlugin.cpp:73-95) */
int orc_fill_volume_info(orc_volume_info* vi)
{
    const uint32_t vmax = std::max(vi->voxels[0],
                                   std::max(vi->voxels[1], vi->voxels[2]));
    vi->world_space_per_voxel = 1.0f / float(vmax);
    for (int i = 0; i < 3; ++i)
        vi->world_size[i] = float(vi->voxels[i]) * vi->world_space_per_voxel;
    uint32_t block[3], levels[3];
    for (int i = 0; i < 3; ++i)
    {
        block[i] = vi->max_block[i] - 2u * vi->overlap[i];
        if (block[i] == 0)
            return -1;
        const uint32_t nb =
            uint32_t(std::ceil(float(vi->voxels[i]) / float(block[i])));
        levels[i] = uint32_t(std::ceil(std::log2(nb)));
    }
    const uint32_t depth = std::min(levels[0], std::min(levels[1], levels[2]));
    for (int i = 0; i < 3; ++i)
        vi->root_blocks[i] = uint32_t(
            std::ceil(float(vi->voxels[i] >> depth) / float(block[i])));
    vi->depth = depth + 1;
    return 0;
}

int orc_node_world_box(const orc_volume_info* vi, uint64_t id, float box[6])
{
    const Box b = world_box(*vi, decode(id));
    for (int i = 0; i < 3; ++i)
    {
        box[i] = b.mn[i];
        box[3 + i] = b.mx[i];
    }
    return 0;
}

int orc_frustum_info_get(const float mv[16], const float proj[16],
                         orc_frustum_info* out)
{
    const Frustum f = make_frustum(mv, proj);
    out->near_plane = f.nearp;
    out->far_plane = f.farp;
    out->left = f.left;
    out->right = f.right;
    out->bottom = f.bottom;
    out->top = f.top;
    for (int i = 0; i < 4; ++i)
        out->near_plane_eq[i] = f.planes[4][i];
    for (int i = 0; i < 3; ++i)
    {
        out->eye[i] = f.eye[i];
        out->view_dir[i] = f.dir[i];
    }
    return 0;
}

/* DFSTraversal::traverse over every root block (DFSTraversal.cpp:63-76) with
 * SelectVisibles, then visitPost's sort-last range split
 * (SelectVisibles.cpp:109-131). */
int orc_select_visibles(const orc_volume_info* vi, const float mv[16],
                        const float proj[16], uint32_t window_height,
                        float sse, uint32_t min_lod, uint32_t max_lod,
                        float range0, float range1, int n_clip,
                        const float* clip_planes, uint64_t* out, int cap,
                        int* n_out)
{
    Selector s;
    s.nClip = n_clip;
    s.clip = clip_planes;
    s.vi = vi;
    s.f = make_frustum(mv, proj);
    s.height = window_height;
    s.sse = sse;
    s.minLod = min_lod;
    s.maxLod = max_lod;
    for (uint32_t x = 0; x < vi->root_blocks[0]; ++x)
        for (uint32_t y = 0; y < vi->root_blocks[1]; ++y)
            for (uint32_t z = 0; z < vi->root_blocks[2]; ++z)
                s.traverse(encode(0, x, y, z, 0), vi->depth);
    const size_t N = s.visibles.size();
    const size_t start = size_t(range0 * float(N));
    const size_t end = size_t(range1 * float(N));
    int k = 0;
    for (size_t i = 0; i < N; ++i)
    {
        if (i < start || i >= end)
            continue;
        if (k >= cap)
            return -1;
        out[k++] = s.visibles[i];
    }
    *n_out = k;
    return 0;
}

/* RayCastRenderer::Impl::order with DistanceOperator
 * (RayCastRenderer.cpp:57-81, 153-159): libstdc++ std::sort (unstable) of
 * the input order, ascending by |MV * centre|. */
int orc_order(const orc_volume_info* vi, const float mv[16],
              const uint64_t* ids, int n, uint64_t* out)
{
    auto dist = [&](uint64_t id) {
        const Box b = world_box(*vi, decode(id));
        float c[3];
        for (int i = 0; i < 3; ++i)
            c[i] = (b.mn[i] + b.mx[i]) * 0.5f;
        float r[4];
        for (int i = 0; i < 4; ++i)
            r[i] = m_at(mv, i, 0) * c[0] + m_at(mv, i, 1) * c[1] +
                   m_at(mv, i, 2) * c[2] + m_at(mv, i, 3);
        const float x = r[0] / r[3], y = r[1] / r[3], z = r[2] / r[3];
        return std::sqrt(x * x + y * y + z * z);
    };
    std::vector<uint64_t> v(ids, ids + n);
    std::sort(v.begin(), v.end(),
              [&](uint64_t a, uint64_t b) { return dist(a) < dist(b); });
    std::copy(v.begin(), v.end(), out);
    return 0;
}

int orc_render_f32(const orc_volume_info* vi, const orc_camera* cam,
                   const uint64_t* ordered, int n,
                   const void* const* brick_ptrs, const float* tf, int n_tf,
                   const orc_render_params* p, float* out_rgba,
                   uint64_t* samples, int n_threads)
{
    return render<float>(vi, cam, ordered, n, brick_ptrs, tf, n_tf, p,
                         out_rgba, samples, n_threads);
}

int orc_render_f64(const orc_volume_info* vi, const orc_camera* cam,
                   const uint64_t* ordered, int n,
                   const void* const* brick_ptrs, const float* tf, int n_tf,
                   const orc_render_params* p, float* out_rgba,
                   uint64_t* samples, int n_threads)
{
    return render<double>(vi, cam, ordered, n, brick_ptrs, tf, n_tf, p,
                          out_rgba, samples, n_threads);
}

/* fragTexCopy.glsl:11-15 into an RGBA8 framebuffer (GL 4.2 §2.1.6.1
 * unsigned normalized conversion). */
void orc_to_rgba8(const float* rgba, int n_pixels, uint8_t* out)
{
    for (int i = 0; i < n_pixels * 4; ++i)
    {
        const float c = std::min(1.0f, std::max(0.0f, rgba[i]));
        out[i] = uint8_t(std::floor(c * 255.0f + 0.5f));
    }
}

/* HistogramObject::Impl::load + binDataSlow (HistogramObject.cpp:30-73,
 * 140-224) accumulated over the rendered set (HistogramFilter.cpp:88-146,
 * full viewport so the centre test passes for every brick). 8-bit: 256 bins,
 * bin = value. 16-bit: 65536 bins, bin = value (the reference's 1024-bin
 * uint16 path divides by a truncated bin width and is not reproducible).
 * float: 256 bins over [range0, range1] as binDataSlow computes them. */
int orc_histogram(const orc_volume_info* vi, const uint64_t* ids, int n,
                  const void* const* brick_ptrs, int bins, float range0,
                  float range1, uint64_t* out)
{
    std::fill(out, out + bins, uint64_t(0));
    const uint32_t* M = vi->max_block;
    const uint32_t* P = vi->overlap;
    for (int b = 0; b < n; ++b)
    {
        const void* vox = brick_ptrs[b];
        if (!vox)
            continue;
        const Node node = decode(ids[b]);
        const uint64_t s1 = uint64_t(1) << (vi->depth - node.level - 1);
        const uint64_t scale = s1 * s1 * s1;
        std::map<double, uint64_t> values;
        std::map<uint32_t, uint64_t> ivalues;
        for (size_t i = P[0]; i < M[0] - P[0]; ++i)
            for (size_t j = P[1]; j < M[1] - P[1]; ++j)
                for (size_t k = P[2]; k < M[2] - P[2]; ++k)
                {
                    const size_t index = i * M[1] * M[2] + j * M[2] + k;
                    switch (vi->data_type)
                    {
                    case ORC_DT_UINT8:
                        ++ivalues[static_cast<const uint8_t*>(vox)[index]];
                        break;
                    case ORC_DT_UINT16:
                        ++ivalues[static_cast<const uint16_t*>(vox)[index]];
                        break;
                    case ORC_DT_FLOAT:
                        ++values[static_cast<const float*>(vox)[index]];
                        break;
                    default:
                        return -2;
                    }
                }
        if (vi->data_type == ORC_DT_UINT8 || vi->data_type == ORC_DT_UINT16)
        {
            /* min/max forced to the type range, so perBinCount is 1 */
            for (const auto& kv : ivalues)
            {
                if (kv.first >= uint32_t(bins))
                    return -3;
                out[kv.first] += scale * kv.second;
            }
        }
        else
        {
            const float minVal =
                std::min(float(values.begin()->first), range0);
            const float maxVal =
                std::max(float(values.rbegin()->first), range1);
            const float range = maxVal - minVal;
            if (minVal != range0 || maxVal != range1)
                return -4; /* incompatible histogram, purged by the filter */
            const float perBinCount = range / float(bins - 1);
            for (const auto& kv : values)
            {
                const size_t bin =
                    size_t((float(kv.first) - minVal) / perBinCount);
                out[std::min(bin, size_t(bins - 1))] += scale * kv.second;
            }
        }
    }
    (void)bytes_per_voxel;
    return 0;
}

/* TransferFunction1D::TransferFunction1D() (TransferFunction1D.cpp:56-90) */
void orc_default_tf(float* rgba)
{
    const float c1[3] = {0.0f, 1.0f, 1.0f};
    const float c2[3] = {1.0f, 0.0f, 1.0f};
    const float alpha1 = 0.39f, alpha2 = 0.03f;
    const float density1 = 127.f, density2 = 255.f;
    float diff[3];
    for (int c = 0; c < 3; ++c)
        diff[c] = c2[c] - c1[c];
    const float alphaDiff = alpha2 - alpha1;
    for (int i = 0; i < 256; ++i)
    {
        float v[4] = {0, 0, 0, 0};
        if (i > 0 && float(i) <= density1)
        {
            v[0] = c1[0];
            v[1] = c1[1];
            v[2] = c1[2];
            v[3] = alpha1 * (float(i) / density1);
        }
        else if (i > 0)
        {
            const float f =
                (float(i) - density1 + 1.f) / (density2 - density1 + 1.f);
            for (int c = 0; c < 3; ++c)
                v[c] = f * diff[c] + c1[c];
            v[3] = f * alphaDiff + alpha1;
        }
        for (int c = 0; c < 4; ++c)
            rgba[i * 4 + c] = v[c];
    }
}

/* TransferFunction1D::getLUT (TransferFunction1D.cpp:92-107) */
void orc_tf_lut(const float* rgba, int n, uint8_t* lut)
{
    for (int i = 0; i < n * 4; ++i)
        lut[i] = uint8_t(rgba[i] * 255.f);
}
}
