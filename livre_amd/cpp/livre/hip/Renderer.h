/*
 * C++ face of the MI355X render path, in namespace livre::hip so that it can
 * sit in a translation unit beside Livre's own livre:: classes. It mirrors
 * the reference's interfaces for callers outside Livre (LivreIntegration.h
 * binds the same library to Livre's own livre::Renderer, DataSource and
 * DataSourcePlugin):
 *   livre::NodeId            livre/data/NodeId.h:31-160
 *   livre::VolumeInformation livre/data/VolumeInformation.h:42-111
 *   livre::Frustum           livre/data/Frustum.h:37-111
 *   livre::Renderer          livre/core/render/Renderer.h:38-109
 *   livre::DataSourcePlugin  livre/data/DataSourcePlugin.h:55-109 (getData)
 *   livre::RawDataSource     livre/data/RawDataSource.cpp:62-219
 *   livre::FrameGrabber      livre/eq/FrameGrabber.cpp:35-110
 *   livre::RayCastRenderer   livre/eq/render/RayCastRenderer.h:31-74
 *   RenderPipeline::renderSync livre/lib/pipeline/RenderPipeline.cpp:92-158
 * Header-only over the C ABI (include/livre_amd.h); link with liblivre_amd.so.
 * vmmlib types are replaced by plain arrays (column-major float[16]).
 * Errors throw std::runtime_error, as the reference's constructors do
 * (RayCastRenderer.cpp:117-138).
 */
#pragma once

#include <livre_amd.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace livre
{
namespace hip
{
typedef uint64_t Identifier;

/** 64-bit octree node key (level 4 bits, x/y/z 14 bits, time step 18). */
class NodeId
{
public:
    NodeId() : _id(~Identifier(0)) {}
    explicit NodeId(Identifier id) : _id(id) {}
    NodeId(uint32_t level, const std::array<uint32_t, 3>& pos, uint32_t timeStep = 0)
        : _id(Identifier(level & 0xFu) | (Identifier(pos[0] & 0x3FFFu) << 4) |
              (Identifier(pos[1] & 0x3FFFu) << 18) | (Identifier(pos[2] & 0x3FFFu) << 32) |
              (Identifier(timeStep & 0x3FFFFu) << 46))
    {
    }
    uint32_t getLevel() const { return uint32_t(_id & 0xFu); }
    uint32_t getTimeStep() const { return uint32_t(_id >> 46); }
    std::array<uint32_t, 3> getPosition() const
    {
        return {uint32_t((_id >> 4) & 0x3FFFu), uint32_t((_id >> 18) & 0x3FFFu),
                uint32_t((_id >> 32) & 0x3FFFu)};
    }
    bool isValid() const { return getLevel() != 0xFu; }
    bool isRoot() const { return getLevel() == 0; }
    NodeId getParent() const
    {
        if (!isValid() || isRoot())
            return NodeId();
        const auto p = getPosition();
        return NodeId(getLevel() - 1, {p[0] / 2, p[1] / 2, p[2] / 2}, getTimeStep());
    }
    /** Children in the reference order: x outer, y, z inner (NodeId.cpp:92-113). */
    std::vector<NodeId> getChildren() const
    {
        std::vector<NodeId> out;
        if (!isValid())
            return out;
        const auto p = getPosition();
        for (uint32_t x = 0; x < 2; ++x)
            for (uint32_t y = 0; y < 2; ++y)
                for (uint32_t z = 0; z < 2; ++z)
                    out.emplace_back(getLevel() + 1,
                                     std::array<uint32_t, 3>{2 * p[0] + x, 2 * p[1] + y,
                                                             2 * p[2] + z},
                                     getTimeStep());
        return out;
    }
    Identifier getId() const { return _id; }
    bool operator==(const NodeId& o) const { return _id == o._id; }
    bool operator!=(const NodeId& o) const { return _id != o._id; }
    bool operator<(const NodeId& o) const { return _id < o._id; }

private:
    Identifier _id;
};
typedef std::vector<NodeId> NodeIds;

enum DataType
{
    DT_FLOAT = LVR_DT_FLOAT,
    DT_UINT8 = LVR_DT_UINT8,
    DT_UINT16 = LVR_DT_UINT16,
    DT_UINT32 = LVR_DT_UINT32,
    DT_INT8 = LVR_DT_INT8,
    DT_INT16 = LVR_DT_INT16,
    DT_INT32 = LVR_DT_INT32
};

inline void check(int rc, const char* what, const lvr_ctx* ctx = nullptr)
{
    if (rc != LVR_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) +
                                 ")" + (ctx ? std::string(": ") + lvr_last_error(ctx) : ""));
}

/** VolumeInformation with fillRegularVolumeInfo applied. */
struct VolumeInformation
{
    lvr_volume_info info{};
    VolumeInformation() = default;
    VolumeInformation(const std::array<uint32_t, 3>& voxels, uint32_t block, uint32_t overlap,
                      DataType type)
    {
        for (int i = 0; i < 3; ++i)
        {
            info.voxels[i] = voxels[i];
            info.max_block[i] = block + 2 * overlap;
            info.overlap[i] = overlap;
        }
        info.data_type = type;
        check(lvr_fill_regular_volume_info(&info), "fillRegularVolumeInfo");
    }
    uint32_t depth() const { return info.depth; }
    size_t brickBytes() const
    {
        static const size_t bpv[] = {4, 1, 2, 4, 1, 2, 4};
        return size_t(info.max_block[0]) * info.max_block[1] * info.max_block[2] *
               bpv[info.data_type];
    }
};

struct PixelViewport
{
    int32_t x = 0, y = 0, w = 0, h = 0;
};

struct ClipPlanes
{
    std::vector<std::array<float, 4>> planes; // normal xyz, d
};

/** Frustum from modelview and projection (column-major). */
class Frustum
{
public:
    Frustum(const float* mv, const float* proj) { check(lvr_frustum_init(&_f, mv, proj), "Frustum"); }
    const lvr_frustum& raw() const { return _f; }
    const float* getMVMatrix() const { return _f.mv; }
    const float* getProjMatrix() const { return _f.proj; }
    const float* getInvMVMatrix() const { return _f.inv_mv; }
    const float* getInvProjMatrix() const { return _f.inv_proj; }
    std::array<float, 3> getEyePos() const { return {_f.eye[0], _f.eye[1], _f.eye[2]}; }
    std::array<float, 3> getViewDir() const
    {
        return {_f.view_dir[0], _f.view_dir[1], _f.view_dir[2]};
    }
    float nearPlane() const { return _f.near_plane; }
    float farPlane() const { return _f.far_plane; }
    float top() const { return _f.top; }
    float bottom() const { return _f.bottom; }

private:
    lvr_frustum _f;
};

/** DataSourcePlugin::getData counterpart: bricks of maximumBlockSize^3 voxels. */
class DataSourcePlugin
{
public:
    virtual ~DataSourcePlugin() {}
    const VolumeInformation& getVolumeInfo() const { return _volumeInfo; }
    /** Fill `out` (getVolumeInfo().brickBytes() bytes); false if unavailable. */
    virtual bool getData(const NodeId& node, void* out) = 0;

protected:
    VolumeInformation _volumeInfo;
};

/** RawDataSource (livre/data/RawDataSource.cpp:62-219): a raw or NRRD volume
 * served as one brick, from the reference's raw:// URIs. pyramidInfo() and
 * BrickPool::loadPyramid() give it levels of detail instead. */
class RawDataSource : public DataSourcePlugin
{
public:
    explicit RawDataSource(const std::string& uri)
    {
        char err[512] = {0};
        const int rc = lvr_raw_open(uri.c_str(), &_src, err, sizeof err);
        if (rc != LVR_OK)
            throw std::runtime_error(std::string("RawDataSource: ") + err);
        lvr_raw_info ri;
        lvr_raw_get_info(_src, &ri);
        _volumeInfo.info = ri.info;
    }
    ~RawDataSource() override { lvr_raw_close(_src); }
    RawDataSource(const RawDataSource&) = delete;
    RawDataSource& operator=(const RawDataSource&) = delete;
    bool getData(const NodeId& node, void* out) override
    {
        return lvr_raw_brick(_src, node.getId(), out, _volumeInfo.brickBytes()) == LVR_OK;
    }
    VolumeInformation pyramidInfo(uint32_t block, uint32_t overlap) const
    {
        VolumeInformation vi;
        check(lvr_pyramid_volume_info(_src, block, overlap, &vi.info), "pyramidInfo");
        return vi;
    }
    const lvr_raw_source* handle() const { return _src; }

private:
    lvr_raw_source* _src = nullptr;
};

/** The procedural volumes as a DataSource URI (no reference counterpart; the
 * reference's mem:// source, MemoryDataSource.cpp:63-98, is the template):
 *   procedural://noise?seed=3&octaves=5&cells=8&gain=1.5#2048,2048,2048,64,uint16
 *   procedural://blobs?seed=1&count=64&noise=0.1#256,256,256,32,uint8
 *   procedural://sines?seed=7&count=8#1536,1536,1536,48,float
 * The fragment gives the voxels per axis, the brick size (without the
 * 1-voxel overlap) and the type (uint8, uint16 or float). */
struct ProceduralSpec
{
    enum Kind
    {
        NOISE,
        BLOBS,
        SINES
    } kind = NOISE;
    std::array<uint32_t, 3> voxels{{0, 0, 0}};
    uint32_t block = 0;
    DataType type = DT_UINT8;
    uint64_t seed = 0;
    int octaves = 5, cells = 8, count = 64;
    float gain = 1.5f, noise = 0.0f;
    lvr_noise_params noiseParams{}; // the remap table of a noise source

    /** From the URI's host ("noise", "blobs", "sines"), query (key, value)
     * pairs and fragment; throws std::runtime_error on a malformed spec. */
    static ProceduralSpec parse(const std::string& host,
                                const std::vector<std::pair<std::string, std::string>>& query,
                                const std::string& fragment)
    {
        ProceduralSpec p;
        if (host == "noise")
            p.kind = NOISE;
        else if (host == "blobs")
            p.kind = BLOBS;
        else if (host == "sines")
            p.kind = SINES;
        else
            throw std::runtime_error("procedural source: unknown kind '" + host + "'");
        std::vector<std::string> f;
        size_t b = 0;
        for (size_t e; (e = fragment.find(',', b)) != std::string::npos; b = e + 1)
            f.push_back(fragment.substr(b, e - b));
        f.push_back(fragment.substr(b));
        if (f.size() != 5)
            throw std::runtime_error("procedural source: fragment must be nx,ny,nz,block,type");
        try
        {
            for (int i = 0; i < 3; ++i)
                p.voxels[size_t(i)] = uint32_t(std::stoul(f[size_t(i)]));
            p.block = uint32_t(std::stoul(f[3]));
            for (const auto& kv : query)
            {
                if (kv.first == "seed")
                    p.seed = std::stoull(kv.second);
                else if (kv.first == "octaves")
                    p.octaves = std::stoi(kv.second);
                else if (kv.first == "cells")
                    p.cells = std::stoi(kv.second);
                else if (kv.first == "count")
                    p.count = std::stoi(kv.second);
                else if (kv.first == "gain")
                    p.gain = std::stof(kv.second);
                else if (kv.first == "noise")
                    p.noise = std::stof(kv.second);
                else
                    throw std::runtime_error("procedural source: unknown parameter '" +
                                             kv.first + "'");
            }
        }
        catch (const std::logic_error&)
        {
            throw std::runtime_error("procedural source: malformed number");
        }
        if (f[4] == "uint8")
            p.type = DT_UINT8;
        else if (f[4] == "uint16")
            p.type = DT_UINT16;
        else if (f[4] == "float")
            p.type = DT_FLOAT;
        else
            throw std::runtime_error("procedural source: type must be uint8, uint16 or float");
        if (p.block == 0 || p.voxels[0] == 0 || p.voxels[1] == 0 || p.voxels[2] == 0)
            throw std::runtime_error("procedural source: empty volume or brick");
        if (p.kind == NOISE &&
            lvr_noise_params_init(&p.noiseParams, p.seed, p.octaves, p.cells, p.gain) != LVR_OK)
            throw std::runtime_error("procedural source: invalid noise parameters");
        return p;
    }

    VolumeInformation volumeInfo() const { return VolumeInformation(voxels, block, 1, type); }

    /** One brick, x fastest (DataSourcePlugin::getData): `out` holds
     * volumeInfo().brickBytes() bytes. */
    int brick(const VolumeInformation& vi, uint64_t nodeId, void* out) const
    {
        if (kind == NOISE)
            return lvr_noise_brick_host(&vi.info, &noiseParams, nodeId, out, vi.brickBytes());
        lvr_field_params fp{};
        fp.kind = kind == BLOBS ? LVR_FIELD_BLOBS : LVR_FIELD_SINES;
        fp.count = count;
        fp.seed = seed;
        fp.noise = noise;
        return lvr_field_brick_host(&vi.info, &fp, nodeId, out, vi.brickBytes());
    }
};

/** The HBM brick pool of one volume: TexturePool + the texture cache. */
class BrickPool
{
public:
    BrickPool(const VolumeInformation& vi, size_t bytes, int device = 0) : _vi(vi)
    {
        check(lvr_ctx_create(&vi.info, bytes, device, &_ctx), "BrickPool");
    }
    ~BrickPool() { lvr_ctx_destroy(_ctx); }
    BrickPool(const BrickPool&) = delete;
    BrickPool& operator=(const BrickPool&) = delete;
    lvr_ctx* ctx() const { return _ctx; }
    const VolumeInformation& volumeInfo() const { return _vi; }
    bool isResident(const NodeId& n) const { return lvr_brick_resident(_ctx, n.getId()) == 1; }
    void upload(const NodeId& n, const void* voxels)
    {
        check(lvr_brick_upload(_ctx, n.getId(), voxels, _vi.brickBytes()), "upload", _ctx);
    }
    void evict(const NodeId& n) { check(lvr_brick_evict(_ctx, n.getId()), "evict", _ctx); }
    /** Bricks of a raw source's LOD pyramid (the pool's volume information
     * must be source.pyramidInfo(block, overlap)). */
    void loadPyramid(const RawDataSource& source, const std::vector<uint64_t>& ids)
    {
        check(lvr_pyramid_load(_ctx, source.handle(), ids.data(), int(ids.size())),
              "loadPyramid", _ctx);
    }
    uint64_t capacity() const
    {
        uint64_t s = 0;
        lvr_pool_stats(_ctx, &s, nullptr, nullptr);
        return s;
    }

private:
    VolumeInformation _vi;
    lvr_ctx* _ctx = nullptr;
};

static const uint32_t RENDER_BEGIN = 1u;
static const uint32_t RENDER_FRAME = 1u << 1;
static const uint32_t RENDER_END = 1u << 2;
static const uint32_t RENDER_ALL = RENDER_BEGIN | RENDER_FRAME | RENDER_END;

/** livre::Renderer: template method order -> begin -> render -> end. */
class Renderer
{
public:
    void render(const Frustum& frustum, const ClipPlanes& planes, const PixelViewport& view,
                const NodeIds& bricks, uint32_t renderStages = RENDER_ALL)
    {
        const NodeIds ordered = order(bricks, frustum);
        if (renderStages & RENDER_BEGIN)
            _onFrameStart(frustum, planes, view, ordered);
        if (renderStages & RENDER_FRAME)
            _onFrameRender(frustum, planes, view, ordered);
        if (renderStages & RENDER_END)
            _onFrameEnd(frustum, planes, view, ordered);
    }
    virtual NodeIds order(const NodeIds& bricks, const Frustum& frustum) const = 0;
    virtual ~Renderer() {}

protected:
    virtual void _onFrameStart(const Frustum&, const ClipPlanes&, const PixelViewport&,
                               const NodeIds&)
    {
    }
    virtual void _onFrameRender(const Frustum&, const ClipPlanes&, const PixelViewport&,
                                const NodeIds&)
    {
    }
    virtual void _onFrameEnd(const Frustum&, const ClipPlanes&, const PixelViewport&,
                             const NodeIds&)
    {
    }
};

/** Drop-in for RayCastRenderer: same constructor meaning (samples per ray,
 * 0 = derived from the data as onFrameStart does), same order(), the frame
 * rendered by the HIP kernels into an RGBA8 framebuffer on the device. */
/** FrameGrabber (livre/eq/FrameGrabber.cpp:35-110): frames as the JPEG files
 * _encodeJpeg writes through tjCompress2 (4:4:4, quality 100, rows taken
 * bottom-up), encoded on the device. */
class FrameGrabber
{
public:
    explicit FrameGrabber(int device = 0)
    {
        if (lvr_jpeg_encoder_create(device, &_enc) != LVR_OK)
            throw std::runtime_error(std::string("FrameGrabber: ") + lvr_jpeg_error(nullptr));
    }
    ~FrameGrabber() { lvr_jpeg_encoder_destroy(_enc); }
    FrameGrabber(const FrameGrabber&) = delete;
    FrameGrabber& operator=(const FrameGrabber&) = delete;

    /** _encodeJpeg(width, height, rawData, dataSize): BGRA host pixels of a
     * read-back image, row 0 at the bottom. */
    std::vector<uint8_t> encodeJpeg(uint32_t width, uint32_t height, const uint8_t* rawData,
                                    int quality = 100)
    {
        std::vector<uint8_t> out(lvr_jpeg_max_size(int(width), int(height)));
        size_t size = 0;
        if (lvr_jpeg_encode_host(_enc, rawData, int(width), int(height), int(4 * width),
                                 LVR_PIXEL_BGRA, 1, quality, out.data(), out.size(),
                                 &size) != LVR_OK)
            throw std::runtime_error(std::string("encodeJpeg: ") + lvr_jpeg_error(_enc));
        out.resize(size);
        return out;
    }

    /** The pool's last rendered frame, encoded where it lies in HBM. */
    static std::vector<uint8_t> grab(const BrickPool& pool, int quality = 100)
    {
        void* ptr = nullptr;
        int w = 0, h = 0;
        check(lvr_frame_device_ptr(pool.ctx(), &ptr, &w, &h), "grab", pool.ctx());
        std::vector<uint8_t> out(lvr_jpeg_max_size(w, h));
        size_t size = 0;
        check(lvr_frame_grab_jpeg(pool.ctx(), quality, out.data(), out.size(), &size), "grab",
              pool.ctx());
        out.resize(size);
        return out;
    }

private:
    lvr_jpeg_encoder* _enc = nullptr;
};

class HipRayCastRenderer : public Renderer
{
public:
    HipRayCastRenderer(BrickPool& pool, uint32_t samplesPerRay)
        : _pool(pool), _samplesPerRay(samplesPerRay)
    {
    }

    /** RayCastRenderer::update: TF (RGBA float, n entries, data range),
     * samples per ray, linear filtering. */
    void update(const std::vector<float>& rgba, std::array<float, 2> range,
                uint32_t samplesPerRay, bool linearFiltering)
    {
        check(lvr_tf_set(_pool.ctx(), rgba.data(), int(rgba.size() / 4), range[0], range[1]),
              "tf", _pool.ctx());
        _samplesPerRay = samplesPerRay;
        _linear = linearFiltering;
    }
    void updateLUT(const std::vector<uint8_t>& lut, std::array<float, 2> range)
    {
        check(lvr_tf_set_rgba8(_pool.ctx(), lut.data(), int(lut.size() / 4), range[0],
                               range[1]),
              "tf", _pool.ctx());
    }
    void setEarlyExit(float a) { _earlyExit = a; }
    /** 0 = derived from the rendered bricks (RayCastRenderer.cpp:216-234). */
    void setSamplesPerRay(uint32_t n) { _samplesPerRay = n; }
    void setLinearFiltering(bool on) { _linear = on; }

    /** The three frame stages on an already ordered list, for a caller whose
     * own Renderer::render ran order() (LivreIntegration.h). */
    void frameStart(const Frustum& f, const ClipPlanes& c, const PixelViewport& v,
                    const NodeIds& ordered)
    {
        _onFrameStart(f, c, v, ordered);
    }
    void frameRender(const Frustum& f, const ClipPlanes& c, const PixelViewport& v,
                     const NodeIds& ordered)
    {
        _onFrameRender(f, c, v, ordered);
    }
    void frameEnd(const Frustum& f, const ClipPlanes& c, const PixelViewport& v,
                  const NodeIds& ordered)
    {
        _onFrameEnd(f, c, v, ordered);
    }

    /** Sort-last (SelectVisibles.cpp:112-131, Channel.cpp:464-603): render
     * the following frames' bricks, a range of the global order, into the
     * RGBA32F layer `layer` instead of the framebuffer; -1 (the default)
     * renders whole frames. */
    void setPartialLayer(int layer) { _layer = layer; }
    /** Device pointer of a layer, for the exchange between GPUs. */
    void* layerDevicePtr(int layer) const
    {
        void* p = nullptr;
        check(lvr_partial_device_ptr(_pool.ctx(), layer, &p), "layer", _pool.ctx());
        return p;
    }
    /** Depth-ordered composite of layers (front first) into the framebuffer;
     * readFrame() then returns it. */
    void composite(const std::vector<const void*>& layers, const PixelViewport& view)
    {
        check(lvr_composite_layers(_pool.ctx(), layers.data(), int(layers.size()), view.w,
                                   view.h, 1, nullptr),
              "composite", _pool.ctx());
        _vp = view;
    }

    NodeIds order(const NodeIds& bricks, const Frustum& frustum) const final
    {
        std::vector<uint64_t> in(bricks.size()), out(bricks.size());
        for (size_t i = 0; i < bricks.size(); ++i)
            in[i] = bricks[i].getId();
        check(lvr_order(&_pool.volumeInfo().info, &frustum.raw(), in.data(), int(in.size()),
                        out.data()),
              "order");
        NodeIds r;
        r.reserve(out.size());
        for (uint64_t id : out)
            r.emplace_back(id);
        return r;
    }

    /** Bricks rendered by the last frame (RayCastRenderer::getVisibleNodes). */
    const NodeIds& getVisibleNodes() const { return _visible; }
    /** The frame of the last render, RGBA8, row 0 = bottom. */
    std::vector<uint8_t> readFrame(uint64_t* samples = nullptr) const
    {
        std::vector<uint8_t> img(size_t(_vp.w) * _vp.h * 4);
        check(lvr_frame_read(_pool.ctx(), img.data(), img.size(), samples), "frame", _pool.ctx());
        return img;
    }

protected:
    void _onFrameStart(const Frustum&, const ClipPlanes& planes, const PixelViewport& view,
                       const NodeIds& bricks) override
    {
        _vp = view;
        // RayCastRenderer.cpp:216-234: 0 derives it from the finest level
        std::vector<uint64_t> ids(bricks.size());
        for (size_t i = 0; i < bricks.size(); ++i)
            ids[i] = bricks[i].getId();
        int32_t computed = 0;
        check(lvr_samples_per_ray(&_pool.volumeInfo().info, ids.data(), int(ids.size()),
                                  int32_t(_samplesPerRay), &computed),
              "samples per ray");
        _computed = uint32_t(computed);
        std::memset(&_params, 0, sizeof _params);
        _params.n_samples_per_ray = int32_t(_computed);
        _params.max_samples_per_ray = 32;
        _params.early_exit = _earlyExit;
        _params.linear_filter = _linear ? 1 : 0;
        _params.n_clip_planes = int32_t(std::min<size_t>(6, planes.planes.size()));
        for (int i = 0; i < _params.n_clip_planes; ++i)
            for (int j = 0; j < 4; ++j)
                _params.clip_planes[4 * i + j] = planes.planes[size_t(i)][size_t(j)];
    }

    void _onFrameRender(const Frustum& frustum, const ClipPlanes&, const PixelViewport& view,
                        const NodeIds& ordered) final
    {
        std::vector<uint64_t> ids(ordered.size());
        _visible.clear();
        for (size_t i = 0; i < ordered.size(); ++i)
        {
            ids[i] = ordered[i].getId();
            if (_pool.isResident(ordered[i]))
                _visible.push_back(ordered[i]);
        }
        const int vp[4] = {view.x, view.y, view.w, view.h};
        if (_layer >= 0)
            check(lvr_render_partial(_pool.ctx(), &frustum.raw(), vp, ids.data(),
                                     int(ids.size()), &_params, _layer),
                  "render partial", _pool.ctx());
        else
            check(lvr_render(_pool.ctx(), &frustum.raw(), vp, ids.data(), int(ids.size()),
                             &_params),
                  "render", _pool.ctx());
    }

    void _onFrameEnd(const Frustum&, const ClipPlanes&, const PixelViewport&,
                     const NodeIds&) final
    {
        check(lvr_ctx_sync(_pool.ctx()), "sync", _pool.ctx());
    }

private:
    BrickPool& _pool;
    uint32_t _samplesPerRay;
    uint32_t _computed = 0;
    bool _linear = false;
    int _layer = -1;
    float _earlyExit = 0.999f;
    PixelViewport _vp;
    lvr_render_params _params{};
    NodeIds _visible;
};

/** RenderPipeline::renderSync for one pipe: visible set on the device, order,
 * load every missing brick from the data source (DataUploadFilter, sync
 * mode), render, then the histogram of the rendered set. */
class RenderPipeline
{
public:
    RenderPipeline(DataSourcePlugin& source, BrickPool& pool) : _source(source), _pool(pool) {}

    struct Params
    {
        float sse = 4.0f; // volumeRendererParameters.fbs default
        uint32_t minLod = 0, maxLod = 14;
    };

    void render(HipRayCastRenderer& renderer, const Frustum& frustum, const PixelViewport& view,
                const Params& p, std::vector<uint64_t>* histogram = nullptr)
    {
        const lvr_volume_info& vi = _pool.volumeInfo().info;
        if (_vis.empty())
        {
            // every level is the root block grid refined 2x per axis
            size_t nodes = 0;
            for (uint32_t l = 0; l < vi.depth; ++l)
                nodes += size_t(vi.root_blocks[0] << l) * (vi.root_blocks[1] << l) *
                         (vi.root_blocks[2] << l);
            _vis.resize(nodes);
        }
        std::vector<uint64_t>& vis = _vis;
        int n = 0;
        check(lvr_select_visibles(_pool.ctx(), &frustum.raw(), uint32_t(view.h), p.sse, p.minLod,
                                  std::min(p.maxLod, vi.depth - 1), vis.data(), int(vis.size()),
                                  &n),
              "select", _pool.ctx());
        NodeIds nodes;
        nodes.reserve(size_t(n));
        std::vector<uint8_t> buf(_pool.volumeInfo().brickBytes());
        for (int i = 0; i < n; ++i)
        {
            const NodeId node(vis[size_t(i)]);
            if (!_pool.isResident(node) && _source.getData(node, buf.data()))
                _pool.upload(node, buf.data());
            nodes.push_back(node);
        }
        renderer.render(frustum, ClipPlanes(), view, nodes);
        if (histogram)
        {
            const int bins = vi.data_type == LVR_DT_UINT16 ? 65536 : 256;
            histogram->assign(size_t(bins), 0);
            std::vector<uint64_t> ids(vis.begin(), vis.begin() + n);
            check(lvr_histogram(_pool.ctx(), ids.data(), n, bins), "histogram", _pool.ctx());
            check(lvr_histogram_read(_pool.ctx(), histogram->data(), bins), "histogram",
                  _pool.ctx());
        }
    }

private:
    DataSourcePlugin& _source;
    BrickPool& _pool;
    std::vector<uint64_t> _vis; // visible-set buffer, one entry per octree node
};
} // namespace hip
} // namespace livre
