/*
 * Compile-only check (g++ -fsyntax-only): the library's C++ headers in one
 * translation unit with the declarations of Livre's own classes
 * (livre_signatures.h). Nothing may be redefined, and the drop-in classes
 * must be usable exactly where Livre uses RayCastRenderer and its data
 * source plugins: constructed as Channel::Impl::initializeRenderer does
 * (Channel.cpp:130-144), driven through livre::Renderer::render
 * (RenderFilter.cpp:57-58), registered as a DataSourcePlugin
 * (DataSource.cpp:29-63).
 */
#include "livre_signatures.h"

#include <livre/hip/LivreIntegration.h>
#include <livre/hip/Renderer.h>

#include <type_traits>

namespace
{
void drive(const livre::DataSource& dataSource, const livre::Cache& textureCache,
           const livre::FrameData& frameData, const livre::Frustum& frustum,
           const livre::ClipPlanes& planes, const livre::PixelViewport& pvp,
           const livre::NodeIds& bricks)
{
    // Livre owns the concrete renderer (Renderer's destructor is protected)
    // and calls it through the base class
    std::unique_ptr<livre::hip::RayCastRenderer> hipRenderer(
        new livre::hip::RayCastRenderer(dataSource, textureCache, 0));
    hipRenderer->update(frameData);
    livre::Renderer& renderer = *hipRenderer;
    renderer.render(frustum, planes, pvp, bricks, livre::RENDER_ALL);
    const livre::NodeIds& visible = hipRenderer->getVisibleNodes();
    const livre::NodeIds ordered = renderer.order(visible, frustum);
    (void)ordered;
}

static_assert(std::is_base_of<livre::Renderer, livre::hip::RayCastRenderer>::value,
              "drop-in for livre::RayCastRenderer");
static_assert(std::is_base_of<livre::DataSourcePlugin, livre::hip::ProceduralDataSource>::value,
              "a livre::DataSourcePlugin");
static_assert(std::is_constructible<livre::hip::ProceduralDataSource,
                                    const livre::DataSourcePluginData&>::value,
              "the plugin factory's constructor");
bool (*const handles)(const livre::DataSourcePluginData&) =
    &livre::hip::ProceduralDataSource::handles;
std::string (*const describe)() = &livre::hip::ProceduralDataSource::getDescription;
const lvr_brick_source_fn streamSource = &livre::hip::dataSourceBrick;
// the library's own classes keep their names beside Livre's
static_assert(!std::is_same<livre::NodeId, livre::hip::NodeId>::value, "distinct types");
} // namespace

void livreAdapterCheck(const livre::DataSource& ds, const livre::Cache& cache,
                       const livre::FrameData& fd, const livre::Frustum& f,
                       const livre::ClipPlanes& cp, const livre::PixelViewport& vp,
                       const livre::NodeIds& ids)
{
    drive(ds, cache, fd, f, cp, vp, ids);
    (void)handles;
    (void)describe;
    (void)streamSource;
}
