/*
 * The MI355X render path behind Livre's own interfaces, for a Livre build.
 *
 * Include it in a Livre translation unit after the Livre headers that
 * declare the classes it binds to (livre/core/render/Renderer.h,
 * livre/core/render/TransferFunction1D.h, livre/core/cache/Cache.h,
 * livre/data/DataSource.h, livre/data/DataSourcePlugin.h, livre/data/LODNode.h,
 * livre/data/MemoryUnit.h, livre/data/Frustum.h, livre/eq/FrameData.h,
 * livre/eq/settings/RenderSettings.h) and link liblivre_amd.so. Nothing here
 * redefines a livre:: name: everything lives in livre::hip.
 *
 *   livre::hip::RayCastRenderer       in place of livre::RayCastRenderer
 *                                     (livre/eq/render/RayCastRenderer.h:31-74):
 *                                     same constructor, update(), order(),
 *                                     getVisibleNodes() and frame stages
 *   livre::hip::ProceduralDataSource  a livre::DataSourcePlugin
 *                                     (livre/data/DataSourcePlugin.h:55-109) for
 *                                     the procedural:// volumes, with the
 *                                     handles()/getDescription() statics the
 *                                     plugin factory needs (DataSource.cpp:29-63)
 *   livre::hip::dataSourceBrick       livre::DataSource::getData as the
 *                                     library's brick source callback, for
 *                                     asynchronous streaming (lvr_stream_create)
 */
#pragma once

#include <livre/hip/Renderer.h>

namespace livre
{
namespace hip
{
/** The library's volume information from Livre's (the fields the render path
 * reads; fillRegularVolumeInfo recomputes the derived ones, as Livre does). */
inline VolumeInformation fromLivre(const ::livre::VolumeInformation& vi)
{
    VolumeInformation out;
    for (int i = 0; i < 3; ++i)
    {
        out.info.voxels[i] = vi.voxels[i];
        out.info.max_block[i] = vi.maximumBlockSize[i];
        out.info.overlap[i] = vi.overlap[i];
    }
    out.info.data_type = int32_t(vi.dataType); // the same DataType values
    check(lvr_fill_regular_volume_info(&out.info), "fillRegularVolumeInfo");
    return out;
}

/** lvr_brick_source_fn over a livre::DataSource (user = the DataSource):
 * DataSource::getData(NodeId) calls the plugin's getData(const LODNode&) and
 * returns its MemoryUnit (DataSource.cpp:101-111, DataSourcePlugin.h:81). A
 * missing or short brick is reported as LVR_ENOENT, as getData's empty
 * pointer is. */
inline int dataSourceBrick(void* user, uint64_t nodeId, void* out, size_t bytes)
{
    const ::livre::DataSource& source = *static_cast<const ::livre::DataSource*>(user);
    try
    {
        const ::livre::ConstMemoryUnitPtr mem = source.getData(::livre::NodeId(nodeId));
        if (!mem || mem->getAllocSize() < bytes)
            return LVR_ENOENT;
        std::memcpy(out, mem->getData<uint8_t>(), bytes);
        return LVR_OK;
    }
    catch (const std::exception&)
    {
        return LVR_ENOENT;
    }
}

/** A DataSourcePlugin serving the procedural volumes (ProceduralSpec) on the
 * host, bit-identical to the bricks the library computes on the device.
 * Register it in the plugin's library as Livre's sources do:
 *   lunchbox::PluginRegisterer<livre::hip::ProceduralDataSource> registerer; */
class ProceduralDataSource : public ::livre::DataSourcePlugin
{
public:
    explicit ProceduralDataSource(const ::livre::DataSourcePluginData& initData)
    {
        const auto& uri = initData.getURI();
        std::vector<std::pair<std::string, std::string>> query;
        for (auto it = uri.queryBegin(); it != uri.queryEnd(); ++it)
            query.emplace_back(it->first, it->second);
        _spec = ProceduralSpec::parse(uri.getHost(), query, uri.getFragment());
        _vi = _spec.volumeInfo();
        _volumeInfo.voxels =
            ::livre::Vector3ui(_spec.voxels[0], _spec.voxels[1], _spec.voxels[2]);
        _volumeInfo.overlap = ::livre::Vector3ui(1u, 1u, 1u);
        const uint32_t m = _spec.block + 2u;
        _volumeInfo.maximumBlockSize = ::livre::Vector3ui(m, m, m);
        _volumeInfo.dataType = ::livre::DataType(int(_spec.type));
        _volumeInfo.compCount = 1;
        if (!::livre::fillRegularVolumeInfo(_volumeInfo))
            throw std::runtime_error("ProceduralDataSource: cannot set up the regular tree");
    }

    /** DataSourcePlugin::getData: the brick of the node, (block + 2)^3
     * voxels, x fastest; an empty pointer on failure. */
    ::livre::MemoryUnitPtr getData(const ::livre::LODNode& node) final
    {
        auto mem = std::make_shared<::livre::AllocMemoryUnit>(_vi.brickBytes());
        if (_spec.brick(_vi, node.getNodeId().getId(), mem->getData<uint8_t>()) != LVR_OK)
            return ::livre::MemoryUnitPtr();
        return mem;
    }

    static bool handles(const ::livre::DataSourcePluginData& initData)
    {
        return initData.getURI().getScheme() == "procedural";
    }

    static std::string getDescription()
    {
        return "Procedural volumes computed per brick (value noise, Gaussian blobs, "
               "plane waves):\n"
               "  procedural://noise?seed=S&octaves=O&cells=C&gain=G#nx,ny,nz,block,type\n"
               "  procedural://blobs?seed=S&count=N&noise=A#nx,ny,nz,block,type\n"
               "  procedural://sines?seed=S&count=N#nx,ny,nz,block,type\n"
               "  type: uint8, uint16 or float; 1-voxel overlap";
    }

    const ProceduralSpec& spec() const { return _spec; }

private:
    ProceduralSpec _spec;
    VolumeInformation _vi;
};

/**
 * livre::RayCastRenderer's drop-in. The texture cache's budget becomes the
 * HBM brick pool (Window.cpp:62-67 sizes it from gpu-cache-mem); bricks the
 * frame needs and the pool lacks are read through the DataSource (the sync
 * path of DataUploadFilter, DataUploadFilter.cpp:45-103) before the frame is
 * ray-cast. The RGBA8 frame stays in HBM (readFrame(), frameDevicePtr()).
 */
class RayCastRenderer : public ::livre::Renderer
{
public:
    RayCastRenderer(const ::livre::DataSource& dataSource, const ::livre::Cache& textureCache,
                    uint32_t samplesPerRay, int device = 0)
        : _dataSource(dataSource)
        , _pool(fromLivre(dataSource.getVolumeInfo()),
                textureCache.getStatistics().getMaximumMemory(), device)
        , _impl(_pool, samplesPerRay)
    {
    }

    /** RayCastRenderer::update (RayCastRenderer.cpp:160-173): the transfer
     * function's LUT and range, samples per ray, linear filtering. */
    void update(const ::livre::FrameData& frameData)
    {
        const auto& tf = frameData.getRenderSettings().getTransferFunction();
        const auto lut = tf.getLUT();
        std::vector<uint8_t> bytes(lut.size() * 4);
        for (size_t i = 0; i < lut.size(); ++i)
            for (size_t c = 0; c < 4; ++c)
                bytes[4 * i + c] = uint8_t(lut[i][c]);
        const auto& range = tf.getRange();
        _impl.updateLUT(bytes, {{float(range[0]), float(range[1])}});
        _impl.setSamplesPerRay(frameData.getVRParameters().getSamplesPerRay());
        _impl.setLinearFiltering(frameData.getVRParameters().getLinearFiltering());
    }

    /** The bricks the last frame rendered (resident when it was drawn). */
    const ::livre::NodeIds& getVisibleNodes() const { return _visible; }

    ::livre::NodeIds order(const ::livre::NodeIds& bricks,
                           const ::livre::Frustum& frustum) const final
    {
        const NodeIds ordered = _impl.order(toHip(bricks), toHip(frustum));
        return toLivre(ordered);
    }

    BrickPool& pool() { return _pool; }
    HipRayCastRenderer& impl() { return _impl; }
    std::vector<uint8_t> readFrame(uint64_t* samples = nullptr) const
    {
        return _impl.readFrame(samples);
    }

protected:
    void _onFrameStart(const ::livre::Frustum& frustum, const ::livre::ClipPlanes& planes,
                       const ::livre::PixelViewport& view,
                       const ::livre::NodeIds& renderBricks) override
    {
        ClipPlanes clip;
        for (const auto& plane : planes.getPlanes())
        {
            const float* n = plane.getNormal();
            clip.planes.push_back({{n[0], n[1], n[2], plane.getD()}});
        }
        _clip = clip;
        _impl.frameStart(toHip(frustum), clip, toHip(view), toHip(renderBricks));
    }

    void _onFrameRender(const ::livre::Frustum& frustum, const ::livre::ClipPlanes&,
                        const ::livre::PixelViewport& view,
                        const ::livre::NodeIds& orderedBricks) final
    {
        // DataUploadFilter's sync path: every brick of the frame resident
        std::vector<uint8_t> buf(_pool.volumeInfo().brickBytes());
        for (const ::livre::NodeId& node : orderedBricks)
        {
            const NodeId id(node.getId());
            if (!_pool.isResident(id) &&
                dataSourceBrick(const_cast<::livre::DataSource*>(&_dataSource), id.getId(),
                                buf.data(), buf.size()) == LVR_OK)
                _pool.upload(id, buf.data());
        }
        _impl.frameRender(toHip(frustum), _clip, toHip(view), toHip(orderedBricks));
        _visible = toLivre(_impl.getVisibleNodes());
    }

    void _onFrameEnd(const ::livre::Frustum& frustum, const ::livre::ClipPlanes&,
                     const ::livre::PixelViewport& view,
                     const ::livre::NodeIds& renderBricks) final
    {
        _impl.frameEnd(toHip(frustum), _clip, toHip(view), toHip(renderBricks));
    }

private:
    static NodeIds toHip(const ::livre::NodeIds& in)
    {
        NodeIds out;
        out.reserve(in.size());
        for (const ::livre::NodeId& n : in)
            out.emplace_back(n.getId());
        return out;
    }
    static ::livre::NodeIds toLivre(const NodeIds& in)
    {
        ::livre::NodeIds out;
        out.reserve(in.size());
        for (const NodeId& n : in)
            out.emplace_back(n.getId());
        return out;
    }
    static Frustum toHip(const ::livre::Frustum& f)
    {
        return Frustum(f.getMVMatrix().array, f.getProjMatrix().array);
    }
    /** Livre's PixelViewport is (x, y, w, h) (Channel.cpp:308). */
    static PixelViewport toHip(const ::livre::PixelViewport& v)
    {
        PixelViewport p;
        p.x = v[0];
        p.y = v[1];
        p.w = v[2];
        p.h = v[3];
        return p;
    }

    const ::livre::DataSource& _dataSource;
    BrickPool _pool;
    HipRayCastRenderer _impl;
    ClipPlanes _clip;
    ::livre::NodeIds _visible;
};
} // namespace hip
} // namespace livre
