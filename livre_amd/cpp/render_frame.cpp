/*
 * render_frame: renders one frame of a procedural volume through the C++
 * drop-in classes (livre::hip::HipRayCastRenderer, livre::hip::RenderPipeline), the
 * way Livre's Channel drives RayCastRenderer, and writes the RGBA8 frame and
 * the histogram to files.
 *
 *   render_frame OUT_PREFIX [n voxels] [block] [width] [height]
 *   render_frame --host-check [nucleon.raw]  (host-side API only, no GPU)
 *
 * OUT_PREFIX.rgba (w*h*4 bytes, row 0 = bottom), OUT_PREFIX.jpg (the frame grab),
 * OUT_PREFIX.hist (u64 bins),
 * OUT_PREFIX.txt (bricks, samples, camera matrices).
 */
#include <livre/hip/Renderer.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>

namespace
{
class NoiseSource : public livre::hip::DataSourcePlugin
{
public:
    NoiseSource(uint32_t n, uint32_t block, uint64_t seed)
    {
        _volumeInfo = livre::hip::VolumeInformation({n, n, n}, block, 1, livre::hip::DT_UINT16);
        livre::hip::check(lvr_noise_params_init(&_params, seed, 4, 4, 1.5f), "noise");
    }
    bool getData(const livre::hip::NodeId& node, void* out) override
    {
        return lvr_noise_brick_host(&_volumeInfo.info, &_params, node.getId(), out,
                                    _volumeInfo.brickBytes()) == LVR_OK;
    }

private:
    lvr_noise_params _params;
};

void lookAt(const float eye[3], float* mv)
{
    float f[3] = {-eye[0], -eye[1], -eye[2]};
    float l = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (float& v : f)
        v /= l;
    const float up[3] = {0, 1, 0};
    float s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2],
                  f[0] * up[1] - f[1] * up[0]};
    l = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (float& v : s)
        v /= l;
    const float u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2],
                        s[0] * f[1] - s[1] * f[0]};
    const float rows[3][3] = {{s[0], s[1], s[2]}, {u[0], u[1], u[2]}, {-f[0], -f[1], -f[2]}};
    std::memset(mv, 0, 16 * sizeof(float));
    for (int r = 0; r < 3; ++r)
    {
        for (int c = 0; c < 3; ++c)
            mv[c * 4 + r] = rows[r][c];
        mv[12 + r] = -(rows[r][0] * eye[0] + rows[r][1] * eye[1] + rows[r][2] * eye[2]);
    }
    mv[15] = 1;
}

void perspective(float fovy, float aspect, float n, float f, float* p)
{
    const float t = 1.0f / std::tan(fovy * 3.14159265358979f / 360.0f);
    std::memset(p, 0, 16 * sizeof(float));
    p[0] = t / aspect;
    p[5] = t;
    p[10] = (f + n) / (n - f);
    p[11] = -1;
    p[14] = 2 * f * n / (n - f);
}

int hostCheck(const std::string& rawPath)
{
    const livre::hip::NodeId root(0, {0, 0, 0});
    const auto kids = root.getChildren();
    if (kids.size() != 8 || kids[1].getId() != livre::hip::NodeId(1, {0, 0, 1}).getId() ||
        kids[4].getId() != livre::hip::NodeId(1, {1, 0, 0}).getId() ||
        kids[7].getParent() != root || livre::hip::NodeId().isValid())
        return 1;
    if (livre::hip::NodeId(1, {1, 0, 0}).getId() != 17u) // lodSelection.cpp:78
        return 2;
    const livre::hip::VolumeInformation vi({2048, 2048, 2048}, 64, 0, livre::hip::DT_UINT8);
    if (vi.depth() != 6) // volumeInformation.cpp:63
        return 3;
    float mv[16], proj[16];
    const float eye[3] = {0.4f, 0.3f, 1.9f};
    lookAt(eye, mv);
    perspective(60.f, 1.f, 0.9f, 3.0f, proj);
    const livre::hip::Frustum fr(mv, proj);
    if (std::fabs(fr.getEyePos()[2] - 1.9f) > 1e-5f)
        return 4;
    if (!rawPath.empty())
    {
        // The reference's raw test volume (41^3 uint8, one brick), through
        // RawDataSource and against the file's own bytes.
        livre::hip::RawDataSource raw(rawPath + "#41,41,41");
        const livre::hip::VolumeInformation& ri = raw.getVolumeInfo();
        if (ri.info.voxels[0] != 41 || ri.info.data_type != LVR_DT_UINT8 || ri.depth() != 1)
            return 5;
        std::vector<uint8_t> brick(ri.brickBytes()), file(ri.brickBytes());
        std::ifstream in(rawPath, std::ios::binary);
        in.read(reinterpret_cast<char*>(file.data()), std::streamsize(file.size()));
        if (!in || !raw.getData(livre::hip::NodeId(0, {0, 0, 0}), brick.data()) || brick != file)
            return 6;
        if (raw.pyramidInfo(8, 0).depth() != 4) // 41 -> 64 = 8 * 2^3
            return 7;
        bool threw = false;
        try
        {
            livre::hip::RawDataSource bad(rawPath + "#41,41");
        }
        catch (const std::runtime_error&)
        {
            threw = true;
        }
        if (!threw)
            return 8;
    }
    {
        // procedural:// URIs: the parsed source serves the bricks of the
        // library's host sources with the same parameters
        using Spec = livre::hip::ProceduralSpec;
        const Spec noise = Spec::parse("noise", {{"seed", "3"}, {"octaves", "4"}, {"cells", "4"},
                                                 {"gain", "1.5"}},
                                       "40,40,40,8,uint16");
        const livre::hip::VolumeInformation nvi = noise.volumeInfo();
        if (nvi.info.voxels[0] != 40 || nvi.info.max_block[0] != 10 ||
            nvi.info.data_type != LVR_DT_UINT16)
            return 9;
        lvr_noise_params np;
        livre::hip::check(lvr_noise_params_init(&np, 3, 4, 4, 1.5f), "noise");
        const uint64_t id = livre::hip::NodeId(2, {1, 2, 3}).getId();
        std::vector<uint8_t> a(nvi.brickBytes()), b(nvi.brickBytes());
        if (noise.brick(nvi, id, a.data()) != LVR_OK ||
            lvr_noise_brick_host(&nvi.info, &np, id, b.data(), b.size()) != LVR_OK || a != b)
            return 10;
        const Spec blobs =
            Spec::parse("blobs", {{"seed", "1"}, {"count", "16"}, {"noise", "0.1"}},
                        "32,32,32,16,uint8");
        lvr_field_params fp{};
        fp.kind = LVR_FIELD_BLOBS;
        fp.count = 16;
        fp.seed = 1;
        fp.noise = 0.1f;
        const livre::hip::VolumeInformation bvi = blobs.volumeInfo();
        std::vector<uint8_t> c(bvi.brickBytes()), d(bvi.brickBytes());
        if (blobs.brick(bvi, 0, c.data()) != LVR_OK ||
            lvr_field_brick_host(&bvi.info, &fp, 0, d.data(), d.size()) != LVR_OK || c != d)
            return 11;
        const char* bad[][2] = {{"noise", "40,40,40,8"}, {"fog", "40,40,40,8,uint8"},
                                {"noise", "40,40,40,8,int64"}, {"sines", "0,40,40,8,float"},
                                {"noise", "40,x,40,8,uint8"}};
        for (const auto& bd : bad)
        {
            bool threw = false;
            try
            {
                Spec::parse(bd[0], {}, bd[1]);
            }
            catch (const std::runtime_error&)
            {
                threw = true;
            }
            if (!threw)
                return 12;
        }
        bool threw = false;
        try
        {
            Spec::parse("noise", {{"colour", "red"}}, "40,40,40,8,uint8");
        }
        catch (const std::runtime_error&)
        {
            threw = true;
        }
        if (!threw)
            return 13;
    }
    std::cout << "host check ok" << std::endl;
    return 0;
}
} // namespace

int main(int argc, char** argv)
{
    if (argc > 1 && std::string(argv[1]) == "--host-check")
        return hostCheck(argc > 2 ? argv[2] : "");
    if (argc < 2)
    {
        std::cerr << "usage: render_frame OUT_PREFIX [n] [block] [w] [h]" << std::endl;
        return 2;
    }
    const std::string out = argv[1];
    const uint32_t n = argc > 2 ? uint32_t(std::stoul(argv[2])) : 64;
    const uint32_t block = argc > 3 ? uint32_t(std::stoul(argv[3])) : 16;
    const int w = argc > 4 ? std::stoi(argv[4]) : 96;
    const int h = argc > 5 ? std::stoi(argv[5]) : 72;
    try
    {
        NoiseSource source(n, block, 7);
        const livre::hip::VolumeInformation& vi = source.getVolumeInfo();
        uint64_t slot = 0;
        livre::hip::check(lvr_slot_bytes(&vi.info, &slot), "slot bytes");
        livre::hip::BrickPool pool(vi, size_t(slot) * 4096);
        livre::hip::HipRayCastRenderer renderer(pool, n);
        std::vector<float> tf(256 * 4);
        for (int i = 0; i < 256; ++i)
        {
            const float x = float(i) / 255.0f;
            tf[size_t(4 * i + 0)] = x;
            tf[size_t(4 * i + 1)] = 1.0f - x;
            tf[size_t(4 * i + 2)] = 0.5f;
            tf[size_t(4 * i + 3)] = 0.4f * x * x;
        }
        renderer.update(tf, {0.0f, 65535.0f}, n, false);
        float mv[16], proj[16];
        const float eye[3] = {1.3f, 0.6f, 1.2f};
        lookAt(eye, mv);
        const float d = std::sqrt(eye[0] * eye[0] + eye[1] * eye[1] + eye[2] * eye[2]);
        perspective(60.f, float(w) / float(h), std::max(1e-4f, d - 1.f), d + 1.f, proj);
        const livre::hip::Frustum frustum(mv, proj);
        livre::hip::RenderPipeline pipeline(source, pool);
        livre::hip::RenderPipeline::Params params;
        params.sse = 1.0f;
        livre::hip::PixelViewport vp;
        vp.w = w;
        vp.h = h;
        std::vector<uint64_t> hist;
        pipeline.render(renderer, frustum, vp, params, &hist);
        uint64_t samples = 0;
        const std::vector<uint8_t> img = renderer.readFrame(&samples);
        std::ofstream(out + ".rgba", std::ios::binary)
            .write(reinterpret_cast<const char*>(img.data()), std::streamsize(img.size()));
        // frame grab: from HBM, and from a BGRA host copy as Equalizer hands it over
        const std::vector<uint8_t> jpg = livre::hip::FrameGrabber::grab(pool);
        std::vector<uint8_t> bgra(img);
        for (size_t i = 0; i < bgra.size(); i += 4)
            std::swap(bgra[i], bgra[i + 2]);
        livre::hip::FrameGrabber grabber;
        if (grabber.encodeJpeg(uint32_t(w), uint32_t(h), bgra.data()) != jpg)
            throw std::runtime_error("frame grab: host and device encodes differ");
        std::ofstream(out + ".jpg", std::ios::binary)
            .write(reinterpret_cast<const char*>(jpg.data()), std::streamsize(jpg.size()));
        // sort-last: the ordered visible set in two consecutive ranges, each
        // ray-cast into a layer, composited front first
        const livre::hip::NodeIds visible = renderer.getVisibleNodes();
        const auto mid = visible.begin() + std::ptrdiff_t(visible.size() / 2);
        const livre::hip::ClipPlanes noPlanes;
        renderer.setPartialLayer(0);
        renderer.render(frustum, noPlanes, vp, livre::hip::NodeIds(visible.begin(), mid));
        renderer.setPartialLayer(1);
        renderer.render(frustum, noPlanes, vp, livre::hip::NodeIds(mid, visible.end()));
        renderer.setPartialLayer(-1);
        renderer.composite({renderer.layerDevicePtr(0), renderer.layerDevicePtr(1)}, vp);
        const std::vector<uint8_t> sortLast = renderer.readFrame();
        int sortLastMaxDiff = 0;
        for (size_t i = 0; i < img.size(); ++i)
            sortLastMaxDiff = std::max(sortLastMaxDiff, std::abs(int(sortLast[i]) - int(img[i])));
        std::ofstream(out + ".hist", std::ios::binary)
            .write(reinterpret_cast<const char*>(hist.data()),
                   std::streamsize(hist.size() * sizeof(uint64_t)));
        std::ofstream txt(out + ".txt");
        txt.precision(9);
        txt << "bricks " << visible.size() << "\nsamples " << samples
            << "\nwidth " << w << "\nheight " << h << "\nmv";
        for (float v : mv)
            txt << ' ' << v;
        txt << "\nproj";
        for (float v : proj)
            txt << ' ' << v;
        txt << "\nsortlast_maxdiff " << sortLastMaxDiff;
        txt << "\nvisible";
        for (const auto& node : visible)
            txt << ' ' << node.getId();
        txt << "\n";
        std::cout << "rendered " << visible.size() << " bricks, " << samples
                  << " samples" << std::endl;
    }
    catch (const std::exception& e)
    {
        std::cerr << "render_frame: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
