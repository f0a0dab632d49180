// Raw and NRRD volumes (RawDataSource, livre/data/RawDataSource.cpp:62-219)
// and an LOD pyramid built from them on the device.
//
// The source follows the reference plugin: the file is memory-mapped, the
// URI names the file, an optional output type (?output=...) and, for raw
// files, the size and input type (#nx,ny,nz[,type]); an NRRD header gives
// type, dimension (3), sizes, endianness and an optional detached data file.
// The volume information is the reference's: one brick holding the whole
// volume (depth 1, no overlap, maximumBlockSize = voxels), and a narrower
// output type keeps the high bits of each value (RawDataSource.cpp:41-61).
// Bytes are passed on as stored, big-endian files included (the reference
// only records the flag).
//
// The pyramid has no reference counterpart: the reference serves a raw
// volume as that single brick. Here the raw volume becomes the finest level
// of a regular octree (fillRegularVolumeInfo over the file's voxels and a
// chosen block size and overlap), padded with zeros to the octree's full
// extent; each coarser level is the 2x2x2 mean of the next one (integers:
// sum + 4 >> 3, round half up; floats: ((a+b)+(c+d))+((e+f)+(g+h)) times
// 1/8 in FP32); a brick of level L holds the level's voxels
// [b * block - overlap, (b + 1) * block + overlap) per axis, clamped to the
// level's extent. The levels are built in HBM from one upload of the file and
// the bricks are cut straight into pool slots.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "lvr_internal.h"

struct lvr_raw_source
{
    int fd = -1;
    void* map = nullptr;
    size_t mapBytes = 0;
    size_t offset = 0; // first data byte in the mapped file
    uint32_t dims[3] = {0, 0, 0};
    int inType = LVR_DT_UINT8;
    int outType = LVR_DT_UINT8;
    int bigEndian = 0;
    lvr_volume_info info{};

    ~lvr_raw_source()
    {
        if (map)
            munmap(map, mapBytes);
        if (fd >= 0)
            close(fd);
    }
};

namespace lvr
{
namespace
{
// RawDataSource::Impl::getDataType (RawDataSource.cpp:129-146), over the
// names NRRD::parseHeader maps every NRRD type to (nrrd.hxx:203-214) and the
// reference's own spellings
bool typeFromName(const std::string& s, int* dt)
{
    static const std::map<std::string, int> names = {
        {"char", LVR_DT_INT8},           {"int8", LVR_DT_INT8},
        {"int8_t", LVR_DT_INT8},         {"signed char", LVR_DT_INT8},
        {"unsigned char", LVR_DT_UINT8}, {"uint8", LVR_DT_UINT8},
        {"uint8_t", LVR_DT_UINT8},       {"uchar", LVR_DT_UINT8},
        {"short", LVR_DT_INT16},         {"int16", LVR_DT_INT16},
        {"int16_t", LVR_DT_INT16},       {"short int", LVR_DT_INT16},
        {"signed short int", LVR_DT_INT16},
        {"unsigned short", LVR_DT_UINT16}, {"uint16", LVR_DT_UINT16},
        {"uint16_t", LVR_DT_UINT16},     {"ushort", LVR_DT_UINT16},
        {"unsigned short int", LVR_DT_UINT16},
        {"int", LVR_DT_INT32},           {"int32", LVR_DT_INT32},
        {"int32_t", LVR_DT_INT32},       {"signed int", LVR_DT_INT32},
        {"unsigned int", LVR_DT_UINT32}, {"uint32", LVR_DT_UINT32},
        {"uint32_t", LVR_DT_UINT32},     {"uint", LVR_DT_UINT32},
        {"float", LVR_DT_FLOAT}};
    const auto it = names.find(s);
    if (it == names.end())
        return false;
    *dt = it->second;
    return true;
}

bool endsWith(const std::string& s, const char* suffix)
{
    const size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

bool parseU32(const std::string& s, uint32_t* v)
{
    if (s.empty())
        return false;
    char* end = nullptr;
    const unsigned long long x = std::strtoull(s.c_str(), &end, 10);
    if (*end != '\0' || x == 0 || x > 0xFFFFFFFFull)
        return false;
    *v = uint32_t(x);
    return true;
}

std::string trim(const std::string& s)
{
    const size_t a = s.find_first_not_of(" \t\r");
    const size_t b = s.find_last_not_of(" \t\r");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// NRRD::parseHeader (nrrd.hxx:160-222): "key: value" fields until the first
// empty line, "#" comments skipped, "key:=value" pairs ignored. Returns the
// byte offset after the header, 0 on error.
size_t parseNrrdHeader(const std::string& path, std::map<std::string, std::string>& fields,
                       std::string* err)
{
    std::ifstream f(path.c_str(), std::ios::binary);
    if (!f)
    {
        *err = "cannot open " + path;
        return 0;
    }
    std::string line;
    if (!std::getline(f, line) || line.compare(0, 4, "NRRD") != 0)
    {
        *err = "not an NRRD file: " + path;
        return 0;
    }
    size_t offset = size_t(f.tellg());
    while (std::getline(f, line))
    {
        const std::streamoff pos = f.tellg();
        offset = pos < 0 ? offset + line.size() : size_t(pos);
        if (!line.empty() && line.back() == '\r')
            line.pop_back();
        if (line.empty())
            break; // data follows
        if (line[0] == '#')
            continue;
        const size_t colon = line.find(':');
        if (colon == std::string::npos || colon + 1 >= line.size())
        {
            *err = "malformed NRRD header line: " + line;
            return 0;
        }
        if (line[colon + 1] == '=')
            continue; // key/value pair
        if (line[colon + 1] != ' ')
        {
            *err = "malformed NRRD header line: " + line;
            return 0;
        }
        fields[line.substr(0, colon)] = trim(line.substr(colon + 2));
    }
    return offset;
}

int openError(std::string msg, const std::string& why, char* err, size_t errLen, int rc)
{
    msg += why;
    if (err && errLen)
    {
        std::strncpy(err, msg.c_str(), errLen - 1);
        err[errLen - 1] = '\0';
    }
    return rc;
}

// ------------------------------------------------------------ device side
struct PyramidBrick
{
    int32_t origin[3]; // first voxel of the brick in its level (may be negative)
    uint32_t level;
};

// raw slab (input type) -> finest level (output type), keeping the high bits
template <typename I, typename O>
__global__ void placeSlabKernel(const I* __restrict__ in, uint32_t nx, uint32_t ny,
                                uint32_t nz, uint64_t z0, O* __restrict__ grid, uint64_t P0,
                                uint64_t P1, int shift)
{
    const uint64_t n = uint64_t(nx) * ny * nz;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
    {
        const uint64_t x = i % nx, y = (i / nx) % ny, z = i / (uint64_t(nx) * ny);
        grid[((z0 + z) * P1 + y) * P0 + x] = O(in[i] >> shift);
    }
}

template <typename T>
__device__ inline T mean8(const T* v)
{
    int64_t s = 0;
    for (int i = 0; i < 8; ++i)
        s += int64_t(v[i]);
    return T((s + 4) >> 3);
}
template <>
__device__ inline float mean8<float>(const float* v)
{
    return (((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]))) * 0.125f;
}

// level L (c0 x c1 x c2) from level L + 1 (twice each dimension)
template <typename T>
__global__ void downsampleKernel(const T* __restrict__ fine, T* __restrict__ coarse,
                                 uint64_t c0, uint64_t c1, uint64_t c2)
{
    const uint64_t n = c0 * c1 * c2;
    const uint64_t f0 = 2 * c0, f1 = 2 * c1;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
    {
        const uint64_t x = i % c0, y = (i / c0) % c1, z = i / (c0 * c1);
        T v[8];
        for (int c = 0; c < 8; ++c)
            v[c] = fine[((2 * z + (c >> 2)) * f1 + 2 * y + ((c >> 1) & 1)) * f0 + 2 * x +
                        (c & 1)];
        coarse[i] = mean8<T>(v);
    }
}

// bricks cut from their levels into pool slots (tiled layout), clamped to
// each level's extent; one workgroup row per brick
template <typename T>
__global__ void cutBricksKernel(const T* const* __restrict__ levels,
                                const uint64_t* __restrict__ levelDim,
                                const PyramidBrick* __restrict__ bricks,
                                const int32_t* __restrict__ slots, uint8_t* __restrict__ pool,
                                size_t slotStride, BrickLayout L)
{
    const PyramidBrick b = bricks[blockIdx.y];
    const T* lv = levels[b.level];
    const uint64_t P = levelDim[b.level];
    T* dst = reinterpret_cast<T*>(pool + size_t(slots[blockIdx.y]) * slotStride);
    const uint32_t M0 = L.dim[0], M1 = L.dim[1], M2 = L.dim[2];
    const uint32_t n = M0 * M1 * M2;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x)
    {
        const uint32_t x = i % M0, y = (i / M0) % M1, z = i / (M0 * M1);
        int64_t g[3] = {int64_t(b.origin[0]) + x, int64_t(b.origin[1]) + y,
                        int64_t(b.origin[2]) + z};
        for (int a = 0; a < 3; ++a)
            g[a] = g[a] < 0 ? 0 : (g[a] >= int64_t(P) ? int64_t(P) - 1 : g[a]);
        dst[tiledIndex(L, x, y, z)] = lv[(uint64_t(g[2]) * P + uint64_t(g[1])) * P + g[0]];
    }
}

template <typename I, typename O>
void launchPlace(const void* in, uint32_t nx, uint32_t ny, uint32_t nz, uint64_t z0,
                 void* grid, uint64_t P0, uint64_t P1, int shift, hipStream_t st)
{
    hipLaunchKernelGGL((placeSlabKernel<I, O>), dim3(4096), dim3(256), 0, st,
                       static_cast<const I*>(in), nx, ny, nz, z0, static_cast<O*>(grid), P0,
                       P1, shift);
}

// the output conversions RawDataSource::Impl::getData supports
bool conversionShift(int in, int out, int* shift)
{
    if (in == out)
        *shift = 0;
    else if (in == LVR_DT_UINT16 && out == LVR_DT_UINT8)
        *shift = 8;
    else if (in == LVR_DT_UINT32 && out == LVR_DT_UINT8)
        *shift = 24;
    else if (in == LVR_DT_UINT32 && out == LVR_DT_UINT16)
        *shift = 16;
    else
        return false;
    return true;
}

template <typename I, typename O>
void hostScale(const void* in, void* out, size_t n, int shift)
{
    const I* src = static_cast<const I*>(in);
    O* dst = static_cast<O*>(out);
    for (size_t i = 0; i < n; ++i)
        dst[i] = O(src[i] >> shift);
}
} // namespace

// device pyramid of one raw source, owned by the context
struct RawPyramid
{
    const lvr_raw_source* src = nullptr;
    std::vector<void*> levels;      // device grids, level 0 first
    std::vector<uint64_t> dims;     // grid edge per level (cubic)
    void* d_levelPtrs = nullptr;    // device copy of `levels`
    uint64_t* d_levelDims = nullptr;
};

void destroyPyramid(lvr_ctx* ctx)
{
    if (!ctx->pyramid)
        return;
    for (void* p : ctx->pyramid->levels)
        (void)hipFree(p);
    (void)hipFree(ctx->pyramid->d_levelPtrs);
    (void)hipFree(ctx->pyramid->d_levelDims);
    delete ctx->pyramid;
    ctx->pyramid = nullptr;
}

// The levels of a source's pyramid in HBM: the file into the finest level,
// then each coarser level from the next.
int buildPyramid(lvr_ctx* ctx, const lvr_raw_source* src)
{
    const lvr_volume_info& vi = ctx->vi;
    const size_t bpv = bytesPerVoxel(vi.data_type);
    const uint32_t block = vi.max_block[0] - 2 * vi.overlap[0];
    int rc = LVR_OK;
    {
        destroyPyramid(ctx);
        ctx->pyramid = new RawPyramid;
        RawPyramid& py = *ctx->pyramid;
        py.src = src;
        // level L spans root_blocks * 2^L blocks per axis
        for (uint32_t l = 0; l < vi.depth; ++l)
        {
            const uint64_t P = uint64_t(vi.root_blocks[0]) * (uint64_t(1) << l) * block;
            void* p = nullptr;
            LVR_HIP(ctx, hipMalloc(&p, P * P * P * bpv));
            py.levels.push_back(p);
            py.dims.push_back(P);
        }
        const uint64_t P = py.dims.back();
        if (P < src->dims[0] || P < src->dims[1] || P < src->dims[2])
            return fail(ctx, LVR_EINVAL, "pyramid grid smaller than the volume");
        void* finest = py.levels.back();
        LVR_HIP(ctx, hipMemsetAsync(finest, 0, P * P * P * bpv, ctx->stream));
        // the file, z-slab by z-slab through pinned staging, into the finest
        // level (output type, zero padding beyond the volume)
        const size_t inBpv = bytesPerVoxel(src->inType);
        const size_t slice = size_t(src->dims[0]) * src->dims[1] * inBpv;
        const uint32_t slabZ =
            uint32_t(std::max<size_t>(1, std::min<size_t>(src->dims[2], (64u << 20) / slice)));
        uint8_t* h = nullptr;
        void* d = nullptr;
        LVR_HIP(ctx, hipHostMalloc((void**)&h, slice * slabZ, hipHostMallocDefault));
        if (hipMalloc(&d, slice * slabZ) != hipSuccess)
        {
            (void)hipHostFree(h);
            return fail(ctx, LVR_ENOMEM, "pyramid staging");
        }
        int shift = 0;
        (void)conversionShift(src->inType, src->outType, &shift);
        const uint8_t* data = static_cast<const uint8_t*>(src->map) + src->offset;
        for (uint32_t z0 = 0; z0 < src->dims[2] && rc == LVR_OK; z0 += slabZ)
        {
            const uint32_t nz = std::min(slabZ, src->dims[2] - z0);
            std::memcpy(h, data + size_t(z0) * slice, slice * nz);
            if (hipMemcpyAsync(d, h, slice * nz, hipMemcpyHostToDevice, ctx->stream) !=
                hipSuccess)
                rc = fail(ctx, LVR_EHIP, "pyramid upload");
            const uint32_t nx = src->dims[0], ny = src->dims[1];
            const hipStream_t st = ctx->stream;
            switch (src->inType * 16 + src->outType)
            {
            case LVR_DT_UINT8 * 16 + LVR_DT_UINT8:
                launchPlace<uint8_t, uint8_t>(d, nx, ny, nz, z0, finest, P, P, shift, st);
                break;
            case LVR_DT_UINT16 * 16 + LVR_DT_UINT16:
                launchPlace<uint16_t, uint16_t>(d, nx, ny, nz, z0, finest, P, P, shift, st);
                break;
            case LVR_DT_UINT16 * 16 + LVR_DT_UINT8:
                launchPlace<uint16_t, uint8_t>(d, nx, ny, nz, z0, finest, P, P, shift, st);
                break;
            case LVR_DT_UINT32 * 16 + LVR_DT_UINT32:
                launchPlace<uint32_t, uint32_t>(d, nx, ny, nz, z0, finest, P, P, shift, st);
                break;
            case LVR_DT_UINT32 * 16 + LVR_DT_UINT16:
                launchPlace<uint32_t, uint16_t>(d, nx, ny, nz, z0, finest, P, P, shift, st);
                break;
            case LVR_DT_UINT32 * 16 + LVR_DT_UINT8:
                launchPlace<uint32_t, uint8_t>(d, nx, ny, nz, z0, finest, P, P, shift, st);
                break;
            case LVR_DT_INT8 * 16 + LVR_DT_INT8:
                launchPlace<int8_t, int8_t>(d, nx, ny, nz, z0, finest, P, P, shift, st);
                break;
            case LVR_DT_INT16 * 16 + LVR_DT_INT16:
                launchPlace<int16_t, int16_t>(d, nx, ny, nz, z0, finest, P, P, shift, st);
                break;
            case LVR_DT_INT32 * 16 + LVR_DT_INT32:
                launchPlace<int32_t, int32_t>(d, nx, ny, nz, z0, finest, P, P, shift, st);
                break;
            default:
                launchPlace<uint32_t, uint32_t>(d, nx, ny, nz, z0, finest, P, P, shift, st);
                break; // float: the bits as they are
            }
            // the staging buffers are reused by the next slab
            if (rc == LVR_OK && (hipGetLastError() != hipSuccess ||
                                 hipStreamSynchronize(ctx->stream) != hipSuccess))
                rc = fail(ctx, LVR_EHIP, "pyramid placement");
        }
        (void)hipHostFree(h);
        (void)hipFree(d);
        if (rc)
            return rc;
        // coarser levels
        for (int l = int(vi.depth) - 2; l >= 0; --l)
        {
            const uint64_t c = py.dims[size_t(l)];
            const void* fine = py.levels[size_t(l) + 1];
            void* coarse = py.levels[size_t(l)];
            const dim3 grid(4096), blk(256);
            switch (vi.data_type)
            {
            case LVR_DT_UINT8:
                hipLaunchKernelGGL(downsampleKernel<uint8_t>, grid, blk, 0, ctx->stream,
                                   static_cast<const uint8_t*>(fine),
                                   static_cast<uint8_t*>(coarse), c, c, c);
                break;
            case LVR_DT_UINT16:
                hipLaunchKernelGGL(downsampleKernel<uint16_t>, grid, blk, 0, ctx->stream,
                                   static_cast<const uint16_t*>(fine),
                                   static_cast<uint16_t*>(coarse), c, c, c);
                break;
            case LVR_DT_UINT32:
                hipLaunchKernelGGL(downsampleKernel<uint32_t>, grid, blk, 0, ctx->stream,
                                   static_cast<const uint32_t*>(fine),
                                   static_cast<uint32_t*>(coarse), c, c, c);
                break;
            case LVR_DT_INT8:
                hipLaunchKernelGGL(downsampleKernel<int8_t>, grid, blk, 0, ctx->stream,
                                   static_cast<const int8_t*>(fine),
                                   static_cast<int8_t*>(coarse), c, c, c);
                break;
            case LVR_DT_INT16:
                hipLaunchKernelGGL(downsampleKernel<int16_t>, grid, blk, 0, ctx->stream,
                                   static_cast<const int16_t*>(fine),
                                   static_cast<int16_t*>(coarse), c, c, c);
                break;
            case LVR_DT_INT32:
                hipLaunchKernelGGL(downsampleKernel<int32_t>, grid, blk, 0, ctx->stream,
                                   static_cast<const int32_t*>(fine),
                                   static_cast<int32_t*>(coarse), c, c, c);
                break;
            default:
                hipLaunchKernelGGL(downsampleKernel<float>, grid, blk, 0, ctx->stream,
                                   static_cast<const float*>(fine),
                                   static_cast<float*>(coarse), c, c, c);
                break;
            }
        }
        LVR_HIP(ctx, hipGetLastError());
        LVR_HIP(ctx, hipMalloc(&py.d_levelPtrs, sizeof(void*) * vi.depth));
        LVR_HIP(ctx, hipMalloc((void**)&py.d_levelDims, sizeof(uint64_t) * vi.depth));
        LVR_HIP(ctx, hipMemcpyAsync(py.d_levelPtrs, py.levels.data(), sizeof(void*) * vi.depth,
                                    hipMemcpyHostToDevice, ctx->stream));
        LVR_HIP(ctx, hipMemcpyAsync(py.d_levelDims, py.dims.data(),
                                    sizeof(uint64_t) * vi.depth, hipMemcpyHostToDevice,
                                    ctx->stream));
        LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }

    return LVR_OK;
}

} // namespace lvr

extern "C" {

int lvr_raw_open(const char* uri, lvr_raw_source** out, char* error, size_t error_len)
{
    using namespace lvr;
    if (!uri || !out)
        return LVR_EINVAL;
    *out = nullptr;
    std::string u(uri);
    const std::string where = std::string("raw source ") + uri + ": ";
    // RawDataSource::handles (RawDataSource.cpp:198-212): scheme raw:// or none
    const size_t scheme = u.find("://");
    if (scheme != std::string::npos)
    {
        if (u.compare(0, scheme, "raw") != 0)
            return openError(where, "unsupported scheme", error, error_len, LVR_EINVAL);
        u = u.substr(scheme + 3);
    }
    std::string fragment, query;
    const size_t hash = u.find('#');
    if (hash != std::string::npos)
    {
        fragment = u.substr(hash + 1);
        u = u.substr(0, hash);
    }
    const size_t q = u.find('?');
    if (q != std::string::npos)
    {
        query = u.substr(q + 1);
        u = u.substr(0, q);
    }
    const std::string path = u;
    const bool isRaw = endsWith(path, ".raw") || endsWith(path, ".img");
    const bool isNrrd = endsWith(path, ".nrrd");
    if (!isRaw && !isNrrd)
        return openError(where, "volume extension does not include raw or nrrd", error,
                         error_len, LVR_EINVAL);

    std::unique_ptr<lvr_raw_source> s(new lvr_raw_source);
    std::string dataFile = path;
    if (isRaw)
    {
        // "#nx,ny,nz[,type]" (RawDataSource.cpp:148-176)
        std::vector<std::string> params;
        std::stringstream ss(fragment);
        std::string item;
        while (std::getline(ss, item, ','))
            params.push_back(item);
        if (params.size() < 3)
            return openError(where, "not enough parameters for the raw file", error,
                             error_len, LVR_EINVAL);
        for (int a = 0; a < 3; ++a)
            if (!parseU32(params[size_t(a)], &s->dims[a]))
                return openError(where, "bad size " + params[size_t(a)], error, error_len,
                                 LVR_EINVAL);
        s->inType = LVR_DT_UINT8;
        if (params.size() > 3 && !typeFromName(params[3], &s->inType))
            return openError(where, "unsupported data format " + params[3], error,
                             error_len, LVR_EINVAL);
    }
    else
    {
        // NRRD (RawDataSource.cpp:178-219)
        std::map<std::string, std::string> fields;
        std::string err;
        const size_t header = parseNrrdHeader(path, fields, &err);
        if (!header)
            return openError(where, err, error, error_len, LVR_EINVAL);
        std::string type = fields["type"];
        if (!typeFromName(type, &s->inType))
            return openError(where, "unsupported data format " + type, error, error_len,
                             LVR_EINVAL);
        if (fields["dimension"] != "3")
            return openError(where, "NRRD is not 3D data", error, error_len, LVR_EINVAL);
        std::stringstream sizes(fields["sizes"]);
        for (int a = 0; a < 3; ++a)
        {
            std::string v;
            if (!(sizes >> v) || !parseU32(v, &s->dims[a]))
                return openError(where, "bad sizes", error, error_len, LVR_EINVAL);
        }
        const auto enc = fields.find("encoding");
        if (enc != fields.end() && enc->second != "raw")
            return openError(where, "only raw encoding is supported", error, error_len,
                             LVR_EINVAL);
        s->bigEndian = fields["endian"] == "big" ? 1 : 0;
        const auto df = fields.find("datafile");
        if (df != fields.end())
        {
            // detached data, relative to the header's directory; it holds
            // only voxels (the reference also skips the header's length
            // there, past the end of a file of exactly the volume's size)
            const size_t slash = path.find_last_of('/');
            dataFile = (slash == std::string::npos ? std::string() : path.substr(0, slash + 1)) +
                       df->second;
            s->offset = 0;
        }
        else
            s->offset = header;
    }
    // ?output=type (RawDataSource.cpp:94-103)
    s->outType = s->inType;
    if (!query.empty())
    {
        const size_t eq = query.find('=');
        if (eq == std::string::npos || query.substr(0, eq) != "output" ||
            !typeFromName(query.substr(eq + 1), &s->outType))
            return openError(where, "bad query " + query, error, error_len, LVR_EINVAL);
        int shift = 0;
        if (!conversionShift(s->inType, s->outType, &shift))
            return openError(where, "unsupported data conversion", error, error_len,
                             LVR_EINVAL);
    }

    s->fd = ::open(dataFile.c_str(), O_RDONLY);
    if (s->fd < 0)
        return openError(where, "cannot open " + dataFile, error, error_len, LVR_ENOENT);
    struct stat st;
    if (fstat(s->fd, &st) != 0)
        return openError(where, "cannot stat " + dataFile, error, error_len, LVR_ENOENT);
    const size_t need = size_t(s->dims[0]) * s->dims[1] * s->dims[2] *
                        bytesPerVoxel(s->inType);
    if (size_t(st.st_size) < s->offset + need)
        return openError(where, "file shorter than the volume", error, error_len,
                         LVR_EINVAL);
    s->mapBytes = size_t(st.st_size);
    s->map = mmap(nullptr, s->mapBytes, PROT_READ, MAP_PRIVATE, s->fd, 0);
    if (s->map == MAP_FAILED)
    {
        s->map = nullptr;
        return openError(where, "cannot mmap " + dataFile, error, error_len, LVR_EINVAL);
    }

    // the reference's volume information: one brick, depth 1, no overlap
    lvr_volume_info& vi = s->info;
    std::memset(&vi, 0, sizeof vi);
    for (int a = 0; a < 3; ++a)
    {
        vi.voxels[a] = s->dims[a];
        vi.max_block[a] = s->dims[a];
        vi.overlap[a] = 0;
    }
    vi.data_type = s->outType;
    if (lvr_fill_regular_volume_info(&vi) != LVR_OK || vi.depth != 1)
        return openError(where, "invalid volume information", error, error_len, LVR_EINVAL);
    *out = s.release();
    return LVR_OK;
}

void lvr_raw_close(lvr_raw_source* src)
{
    delete src;
}

int lvr_raw_get_info(const lvr_raw_source* src, lvr_raw_info* info)
{
    if (!src || !info)
        return LVR_EINVAL;
    std::memset(info, 0, sizeof *info);
    info->info = src->info;
    info->input_type = src->inType;
    info->big_endian = src->bigEndian;
    info->data_bytes = uint64_t(src->dims[0]) * src->dims[1] * src->dims[2] *
                       lvr::bytesPerVoxel(src->inType);
    return LVR_OK;
}

int lvr_raw_brick(void* user, uint64_t node_id, void* out, size_t bytes)
{
    using namespace lvr;
    const lvr_raw_source* s = static_cast<const lvr_raw_source*>(user);
    if (!s || !out)
        return LVR_EINVAL;
    const NodeKey k = decodeNode(node_id);
    if (k.level != 0 || k.x || k.y || k.z || k.t)
        return LVR_ENOENT; // the volume is one brick
    const size_t n = size_t(s->dims[0]) * s->dims[1] * s->dims[2];
    if (bytes != n * bytesPerVoxel(s->outType))
        return LVR_EINVAL;
    const uint8_t* data = static_cast<const uint8_t*>(s->map) + s->offset;
    int shift = 0;
    (void)conversionShift(s->inType, s->outType, &shift);
    if (shift == 0)
        std::memcpy(out, data, bytes);
    else if (s->inType == LVR_DT_UINT16)
        hostScale<uint16_t, uint8_t>(data, out, n, shift);
    else if (s->outType == LVR_DT_UINT8)
        hostScale<uint32_t, uint8_t>(data, out, n, shift);
    else
        hostScale<uint32_t, uint16_t>(data, out, n, shift);
    return LVR_OK;
}

int lvr_pyramid_volume_info(const lvr_raw_source* src, uint32_t block, uint32_t overlap,
                            lvr_volume_info* info)
{
    if (!src || !info || block == 0)
        return LVR_EINVAL;
    std::memset(info, 0, sizeof *info);
    for (int a = 0; a < 3; ++a)
    {
        info->voxels[a] = src->dims[a];
        info->max_block[a] = block + 2 * overlap;
        info->overlap[a] = overlap;
    }
    info->data_type = src->outType;
    return lvr_fill_regular_volume_info(info);
}

int lvr_pyramid_load(lvr_ctx* ctx, const lvr_raw_source* src, const uint64_t* node_ids,
                     int n)
{
    using namespace lvr;
    if (!ctx || !src || n < 0 || (n > 0 && !node_ids))
        return LVR_EINVAL;
    const lvr_volume_info& vi = ctx->vi;
    lvr_volume_info want;
    if (lvr_pyramid_volume_info(src, vi.max_block[0] - 2 * vi.overlap[0], vi.overlap[0],
                                &want) != LVR_OK ||
        std::memcmp(&want, &vi, sizeof vi) != 0)
        return fail(ctx, LVR_EINVAL, "context volume information is not this source's pyramid");
    for (int a = 0; a < 3; ++a)
        if (vi.root_blocks[a] != vi.root_blocks[0] || vi.max_block[a] != vi.max_block[0])
            return fail(ctx, LVR_EINVAL, "pyramid needs a cubic brick grid");
    int rc = ensureDevice(ctx);
    if (rc)
        return rc;
    const size_t bpv = bytesPerVoxel(vi.data_type);
    const uint32_t block = vi.max_block[0] - 2 * vi.overlap[0];

    if (!ctx->pyramid || ctx->pyramid->src != src)
    {
        rc = buildPyramid(ctx, src);
        if (rc)
        {
            destroyPyramid(ctx);
            return rc;
        }
    }

    // bricks into slots, in chunks (as the procedural sources do)
    RawPyramid& py = *ctx->pyramid;
    const int chunk = 4096;
    const size_t recBytes = (sizeof(PyramidBrick) * chunk + 255) & ~size_t(255);
    uint8_t* scratch =
        static_cast<uint8_t*>(genScratch(ctx, recBytes + sizeof(int32_t) * chunk));
    if (!scratch)
        return LVR_ENOMEM;
    PyramidBrick* d_bricks = reinterpret_cast<PyramidBrick*>(scratch);
    int32_t* d_slots = reinterpret_cast<int32_t*>(scratch + recBytes);
    rc = waitPoolReaders(ctx, ctx->stream);
    if (rc)
        return rc;
    std::vector<PyramidBrick> recs;
    std::vector<int32_t> slots;
    for (int base = 0; base < n && rc == LVR_OK; base += chunk)
    {
        const int cnt = std::min(chunk, n - base);
        recs.resize(size_t(cnt));
        slots.resize(size_t(cnt));
        int taken = 0;
        for (int i = 0; i < cnt && rc == LVR_OK; ++i)
        {
            const NodeKey k = decodeNode(node_ids[base + i]);
            if (k.level >= vi.depth)
            {
                rc = fail(ctx, LVR_EINVAL, "node id outside the octree");
                break;
            }
            const uint32_t pos[3] = {k.x, k.y, k.z};
            for (int a = 0; a < 3; ++a)
                recs[size_t(i)].origin[a] = int32_t(pos[a] * block) - int32_t(vi.overlap[a]);
            recs[size_t(i)].level = k.level;
            rc = acquireSlotFor(ctx, node_ids[base + i], &slots[size_t(i)]);
            taken += rc == LVR_OK;
        }
        if (rc == LVR_OK &&
            (hipMemcpyAsync(d_bricks, recs.data(), sizeof(PyramidBrick) * cnt,
                            hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
             hipMemcpyAsync(d_slots, slots.data(), sizeof(int32_t) * cnt,
                            hipMemcpyHostToDevice, ctx->stream) != hipSuccess))
            rc = fail(ctx, LVR_EHIP, "pyramid brick upload");
        if (rc)
        {
            const std::string msg = ctx->error;
            for (int i = 0; i < taken; ++i)
                (void)releaseSlot(ctx, node_ids[base + i]);
            ctx->error = msg;
            break;
        }
        const dim3 grid(32, unsigned(cnt)), blk(256);
        switch (bpv)
        {
        case 1:
            hipLaunchKernelGGL(cutBricksKernel<uint8_t>, grid, blk, 0, ctx->stream,
                               static_cast<const uint8_t* const*>(py.d_levelPtrs),
                               py.d_levelDims, d_bricks, d_slots, ctx->d_pool,
                               ctx->slotStride, ctx->layout);
            break;
        case 2:
            hipLaunchKernelGGL(cutBricksKernel<uint16_t>, grid, blk, 0, ctx->stream,
                               static_cast<const uint16_t* const*>(py.d_levelPtrs),
                               py.d_levelDims, d_bricks, d_slots, ctx->d_pool,
                               ctx->slotStride, ctx->layout);
            break;
        default:
            hipLaunchKernelGGL(cutBricksKernel<uint32_t>, grid, blk, 0, ctx->stream,
                               static_cast<const uint32_t* const*>(py.d_levelPtrs),
                               py.d_levelDims, d_bricks, d_slots, ctx->d_pool,
                               ctx->slotStride, ctx->layout);
            break;
        }
        if (hipGetLastError() != hipSuccess)
            rc = fail(ctx, LVR_EHIP, "pyramid bricks");
        // each brick's histogram into the cache as it lands in its slot
        if (rc == LVR_OK)
            rc = cacheBrickHistograms(ctx, node_ids + base, slots.data(), cnt, ctx->stream);
        // the host vectors and the scratch are reused by the next chunk
        if (rc == LVR_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = fail(ctx, LVR_EHIP, "pyramid bricks");
    }
    return rc;
}

int lvr_pyramid_release(lvr_ctx* ctx)
{
    if (!ctx)
        return LVR_EINVAL;
    if (ctx->pyramid)
        LVR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    lvr::destroyPyramid(ctx);
    return LVR_OK;
}
}
