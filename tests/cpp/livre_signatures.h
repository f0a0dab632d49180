/*
 * Declarations, no definitions: the signatures of the Livre classes (and the
 * vmmlib, servus and Lexis types in them) that livre/hip/LivreIntegration.h
 * binds to, as the reference declares them. Used only by the compile check
 * tests/cpp/livre_adapter_check.cpp, which includes these beside the
 * library's headers to show that nothing is redefined and that the adapter
 * matches the reference's interfaces. Each class names the reference header
 * it restates.
 */
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace vmml
{
// vmmlib vector / matrix: public element arrays (RayCastRenderer.cpp:303-318
// passes Matrix4f::array to glUniformMatrix4fv)
template <size_t M, typename T>
struct vector
{
    T array[M];
    vector();
    explicit vector(T all);
    vector(T x, T y, T z);
    vector(T x, T y, T z, T w);
    T& operator[](size_t i);
    const T& operator[](size_t i) const;
};
template <size_t M, size_t N, typename T>
struct Matrix
{
    T array[M * N];
};
} // namespace vmml

namespace servus
{
// servus::URI as Livre's data sources read it (MemoryDataSource.cpp:67-79,
// RawDataSource.cpp:74-102, 243-252)
class URI
{
public:
    typedef std::map<std::string, std::string> KVMap;
    typedef KVMap::const_iterator ConstKVIter;
    const std::string& getScheme() const;
    const std::string& getHost() const;
    const std::string& getPath() const;
    const std::string& getFragment() const;
    ConstKVIter queryBegin() const;
    ConstKVIter queryEnd() const;
    ConstKVIter findQuery(const std::string& key) const;
};
} // namespace servus

namespace lexis
{
namespace render
{
// clip planes as RayCastRenderer.cpp:352-365 reads them
class Plane
{
public:
    const float* getNormal() const;
    float getD() const;
};
class ClipPlanes
{
public:
    const std::vector<Plane>& getPlanes() const;
};
} // namespace render
} // namespace lexis

namespace livre
{
typedef vmml::vector<3, uint32_t> Vector3ui; // core/types.h:210
typedef vmml::vector<3, float> Vector3f;
typedef vmml::vector<2, float> Vector2f;      // core/types.h:213
typedef vmml::vector<4, uint8_t> Vector4ub;
typedef vmml::Matrix<4, 4, float> Matrix4f;   // core/types.h:199
typedef vmml::vector<4, int> PixelViewport;   // core/types.h:227
using ::lexis::render::ClipPlanes;            // data/types.h:64
typedef uint64_t Identifier;

enum DataType // data/VolumeInformation.h:29-39
{
    DT_FLOAT,
    DT_UINT8,
    DT_UINT16,
    DT_UINT32,
    DT_INT8,
    DT_INT16,
    DT_INT32,
    DT_UNDEFINED
};

struct VolumeInformation // data/VolumeInformation.h:42-111 (fields used)
{
    VolumeInformation();
    bool bigEndian;
    uint32_t compCount;
    DataType dataType;
    Vector3ui overlap;
    Vector3ui maximumBlockSize;
    Vector3ui voxels;
    Vector3f worldSize;
    float worldSpacePerVoxel;
};
bool fillRegularVolumeInfo(VolumeInformation& info); // DataSourcePlugin.h:114

class NodeId // data/NodeId.h:54-68
{
public:
    NodeId();
    explicit NodeId(const Identifier& identifier);
    NodeId(const uint32_t level, const Vector3ui& position, const uint32_t timeStep = 0);
    Identifier getId() const;
    uint32_t getLevel() const;
};
typedef std::vector<NodeId> NodeIds;

class LODNode // data/LODNode.h:32-110
{
public:
    LODNode();
    NodeId getNodeId() const;
    uint32_t getRefLevel() const;
};

class MemoryUnit // data/MemoryUnit.h:30-60
{
public:
    MemoryUnit();
    virtual ~MemoryUnit();
    template <class T>
    const T* getData() const
    {
        return reinterpret_cast<const T*>(_getData());
    }
    template <class T>
    T* getData()
    {
        return reinterpret_cast<T*>(_getData());
    }
    virtual size_t getAllocSize() const = 0;

protected:
    virtual const uint8_t* _getData() const = 0;
    virtual uint8_t* _getData() = 0;
};
class AllocMemoryUnit : public MemoryUnit // data/MemoryUnit.h:116-160
{
public:
    explicit AllocMemoryUnit(const size_t size);
    ~AllocMemoryUnit();
    size_t getAllocSize() const final;

private:
    const uint8_t* _getData() const final;
    uint8_t* _getData() final;
};
typedef std::shared_ptr<MemoryUnit> MemoryUnitPtr;
typedef std::shared_ptr<const MemoryUnit> ConstMemoryUnitPtr;

class DataSourcePluginData // data/DataSourcePlugin.h:30-47
{
public:
    const servus::URI& getURI() const;
};

class DataSourcePlugin // data/DataSourcePlugin.h:55-109
{
public:
    DataSourcePlugin();
    typedef DataSourcePlugin InterfaceT;
    typedef DataSourcePluginData InitDataT;
    virtual ~DataSourcePlugin() {}
    const VolumeInformation& getVolumeInfo() const;
    virtual bool initializeGL() { return true; }
    virtual void finishGL() {}
    virtual MemoryUnitPtr getData(const LODNode& node) = 0;
    virtual LODNode internalNodeToLODNode(const NodeId& nodeId) const;
    virtual bool update() { return false; }
    LODNode getNode(const NodeId& nodeId) const;

protected:
    DataSourcePlugin(const DataSourcePlugin&) = delete;
    DataSourcePlugin& operator=(const DataSourcePlugin&) = delete;
    VolumeInformation _volumeInfo;
};

class DataSource // data/DataSource.h:38-140
{
public:
    const VolumeInformation& getVolumeInfo() const;
    MemoryUnitPtr getData(const NodeId& nodeId);
    ConstMemoryUnitPtr getData(const NodeId& nodeId) const;
    LODNode getNode(const NodeId& nodeId) const;
};

class CacheStatistics // core/cache/CacheStatistics.h:43-58
{
public:
    size_t getMaximumMemory() const;
};
class Cache // core/cache/Cache.h:45-175
{
public:
    virtual ~Cache();
    const CacheStatistics& getStatistics() const;
};

class Frustum // data/Frustum.h:37-111
{
public:
    Frustum(const Matrix4f& modelViewMatrix, const Matrix4f& projectionMatrix);
    const Matrix4f& getMVMatrix() const;
    const Matrix4f& getProjMatrix() const;
    const Matrix4f& getInvMVMatrix() const;
    const Matrix4f& getInvProjMatrix() const;
};

class TransferFunction1D // core/render/TransferFunction1D.h:30-70, Lexis MaterialLUT
{
public:
    std::vector<Vector4ub> getLUT() const;
    const Vector2f& getRange() const;
};
class RenderSettings // eq/settings/RenderSettings.h:54-60
{
public:
    const TransferFunction1D& getTransferFunction() const;
};
class VolumeRendererParameters // the generated parameters read at RayCastRenderer.cpp:165-168
{
public:
    uint32_t getSamplesPerRay() const;
    bool getLinearFiltering() const;
    bool getShowAxes() const;
};
class FrameData // eq/FrameData.h:31-150
{
public:
    const RenderSettings& getRenderSettings() const;
    const VolumeRendererParameters& getVRParameters() const;
};

static const uint32_t RENDER_BEGIN = 1u; // core/render/Renderer.h:30-33
static const uint32_t RENDER_FRAME = 1u << 1;
static const uint32_t RENDER_END = 1u << 2;
static const uint32_t RENDER_ALL = RENDER_BEGIN | RENDER_FRAME | RENDER_END;

class Renderer // core/render/Renderer.h:38-109
{
public:
    void render(const Frustum& frustum, const ClipPlanes& planes, const PixelViewport& view,
                const NodeIds& bricks, const uint32_t renderStages = RENDER_ALL);
    virtual NodeIds order(const NodeIds& bricks, const Frustum& frustum) const = 0;

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
    virtual ~Renderer();
};
} // namespace livre
