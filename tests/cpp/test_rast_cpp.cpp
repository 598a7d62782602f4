// test_rast_cpp.cpp -- the C++ classes of rast.hpp driven the way the
// reference's app drives CUDARenderer (CCR/main.cpp:440-459,493,527-529),
// with app-side types shaped like the reference's (Mesh, SceneGraph,
// SceneManager, Camera, WindowHelper), on the 640x480 UV-sphere frame.
// The frame is checked against the CPU oracle: visibility and depth bit for
// bit, RGBA8 within 1/255 per channel; a texture loaded from the PNG fixture
// tests/golden/uv_grid.png (argv[1]: the repository root) is drawn and
// checked texel by texel.  Exit status 0 = pass.
#define RAST_REFERENCE_NAMES
#include "rast.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "rast_oracle.h"
#include "rast_scene.h"

namespace app {

struct FPoint3 { float x, y, z; };
struct FVector3 { float x, y, z; };
struct IPoint2 { int x, y; };
struct FMatrix4 { float data[4][4]; };
struct IVertex { float p[3], uv[2], n[3], tan[3], c[3]; };
enum class CullingMode { BackFace, FrontFace, NoCulling };
enum class SampleState { Point, Linear };
enum class VisualisationState { PBR, Depth, Normal };
enum class PrimitiveTopology : unsigned char { TriangleList, TriangleStrip };

class Mesh {
public:
    Mesh(std::vector<IVertex> v, std::vector<unsigned int> i) : v_(std::move(v)), i_(std::move(i))
    {
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) world_.data[c][r] = c == r ? 1.f : 0.f;
    }
    IVertex *GetVertexBuffer() const { return const_cast<IVertex *>(v_.data()); }
    unsigned int GetNumVertices() const { return (unsigned int)v_.size(); }
    unsigned int *GetIndexBuffer() const { return const_cast<unsigned int *>(i_.data()); }
    unsigned int GetNumIndices() const { return (unsigned int)i_.size(); }
    PrimitiveTopology GetTopology() const { return PrimitiveTopology::TriangleList; }
    const FMatrix4 &GetWorldConst() const { return world_; }
    const int *GetTextureIds() const { return tex_; }
    int tex_[4] = { -1, -1, -1, -1 };

private:
    std::vector<IVertex> v_;
    std::vector<unsigned int> i_;
    FMatrix4 world_;
};

class SceneGraph {
public:
    const std::vector<Mesh *> &GetMeshes() const { return meshes_; }
    std::vector<Mesh *> meshes_;
};

class SceneManager {
public:
    CullingMode GetCullingMode() const { return CullingMode::BackFace; }
    SampleState GetSampleState() const { return SampleState::Point; }
    VisualisationState GetVisualisationState() const { return VisualisationState::PBR; }
};

class Camera {
public:
    const FPoint3 &GetPosition() const { return pos_; }
    const FVector3 &GetForward() const { return fwd_; }
    float GetFov() const { return 45.f; }
    void GetNearFarValues(float &n, float &f) { n = 0.1f; f = 100.f; }

private:
    FPoint3 pos_{ 0.f, 0.f, 3.f };
    FVector3 fwd_{ 0.f, 0.f, 1.f };
};

struct Resolution { unsigned int Width, Height; };
struct WindowHelper {
    unsigned int *pBackBufferPixels = nullptr;
    app::Resolution Resolution;  // named as in CCR/WindowHelper.h:12
};

}  // namespace app

static uint32_t rgba(int r, int g, int b) { return (uint32_t)r | (uint32_t)g << 8 | (uint32_t)b << 16 | 0xFF000000u; }

#define EXPECT(c, ...)                           \
    do {                                         \
        if (!(c)) {                              \
            std::fprintf(stderr, __VA_ARGS__);   \
            std::fprintf(stderr, "\n");          \
            return 1;                            \
        }                                        \
    } while (0)

int main(int argc, char **argv)
{
    const std::string root = argc > 1 ? argv[1] : ".";  // the repository, for the PNG fixture
    const uint32_t W = 640, H = 480;
    uint32_t nv = 0, ni = 0;
    rsc_uv_sphere(64, 32, 1.f, nullptr, &nv, nullptr, &ni);
    std::vector<app::IVertex> verts(nv);
    std::vector<unsigned int> idx(ni);
    EXPECT(rsc_uv_sphere(64, 32, 1.f, &verts[0].p[0], &nv, idx.data(), &ni) == 0, "uv sphere");
    std::vector<uint32_t> checker(256 * 256);
    rsc_texture_checker(256, 256, 32, rgba(235, 235, 235), rgba(200, 60, 40), checker.data());

    app::Mesh mesh(verts, idx);
    mesh.tex_[0] = 0;
    app::SceneGraph sg;
    sg.meshes_.push_back(&mesh);
    app::SceneManager sm;
    app::Camera cam;
    std::vector<uint32_t> back(W * H, 0);
    app::WindowHelper wh;
    wh.pBackBufferPixels = back.data();
    wh.Resolution = { W, H };

    std::vector<uint32_t> vis(W * H);
    std::vector<float> depth(W * H);
    std::vector<uint32_t> drawn(W * H), rendered(W * H);
    // the reference app's order (CCR/main.cpp:418,440,529): the texture
    // manager lives on the stack and outlives the renderer it took its
    // textures from, so its textures are destroyed after the renderer
    CUDATextureManager tm;
    {
        CUDARenderer *pRenderer = new CUDARenderer(wh);
        CUDARenderer &renderer = *pRenderer;
        renderer.DisplayGPUSpecs(0);
        EXPECT(tm.AddCUDATexture(renderer.CreateTexture(checker.data(), 256, 256)) == 0, "texture id");
        renderer.LoadScene(&sg, tm);
        renderer.SetupBins(app::IPoint2{ 40, 30 }, app::IPoint2{ 16, 16 }, 256);
        renderer.KernelWarmUp();
        EXPECT(renderer.EnterValidRenderingState() == 0, "rendering state");
        renderer.StartTimer();
        renderer.RenderAuto(sm, tm, &cam);
        const float ms = renderer.StopTimer();
        rendered = back;  // Present() filled the back buffer
        renderer.Readback(vis.data(), depth.data(), nullptr);
        EXPECT(renderer.GetTotalNumTriangles() == 4032, "triangles %u", renderer.GetTotalNumTriangles());
        std::printf("rendered in %.3f ms, %u of %u triangles visible\n", ms, renderer.GetTotalNumVisibleTriangles(),
                    renderer.GetTotalNumTriangles());
        bool threw = false;
        try {
            renderer.SetupBins(app::IPoint2{ 1, 1 }, app::IPoint2{ 16, 8 }, 256);
        } catch (const rast::Error &e) {
            threw = e.code() == RAST_E_INVALID_ARG;
        }
        EXPECT(threw, "non-square bins must be rejected");
        // the benchmark marker times on the renderer's stream
        renderer.GetBenchMarker().StartTimer();
        renderer.Render(sm, tm, &cam);
        EXPECT(renderer.GetBenchMarker().StopTimer() >= 0.f, "bench marker");
        // DrawTextureGlobal: the checker stretched over the window (point sampling)
        renderer.DrawTextureGlobal(*tm.GetCUDATexture(0), true, app::SampleState::Point);
        drawn = back;
        // a texture from a PNG file, as the reference's CUDATexture(path)
        // loads its maps: the uv grid fixture is the 32-pixel checker
        const std::string grid = root + "/tests/golden/uv_grid.png";
        EXPECT(tm.AddCUDATexture(renderer.CreateTexture(grid)) == 1, "png texture id");
        const CUDATexture *png = tm.GetCUDATexture(1);
        EXPECT(png && png->GetWidth() == 512 && png->GetHeight() == 512, "png texture size");
        EXPECT(renderer.CreateTexture(root + "/tests/golden/no_such.png") == nullptr,
               "a missing file gives no texture");
        renderer.DrawTextureGlobal(*png, false, app::SampleState::Point);
        // the grid's texels by its rule (white where the 32-pixel cells'
        // indices sum to an even number, black elsewhere), sampled by the
        // oracle at uv = (x, y) / 512 (not stretched; magenta past u = 1)
        std::vector<uint32_t> rule(512 * 512);
        for (uint32_t y = 0; y < 512; ++y)
            for (uint32_t x = 0; x < 512; ++x)
                rule[(size_t)y * 512 + x] = ((x / 32 + y / 32) % 2) == 0 ? 0xFFFFFFFFu : 0xFF000000u;
        const orc_texture grid_tex{ rule.data(), 512, 512 };
        for (uint32_t y = 0; y < H; y += 9)
            for (uint32_t x = 0; x < W; x += 7) {
                const float uv[2] = { (float)x / 512.f, (float)y / 512.f };
                float c[3];
                orc_sample(&grid_tex, uv, ORC_SAMPLE_POINT, c);
                const uint32_t want = orc_pack_rgba(c), got = back[(size_t)y * W + x];
                EXPECT(got == want, "png grid at (%u, %u): %08x, want %08x", x, y, got, want);
            }
        // the reference's path form of the hook gives the same frame
        const std::vector<uint32_t> from_texture = back;
        renderer.DrawTextureGlobal(grid.c_str(), false, app::SampleState::Point);
        EXPECT(back == from_texture, "DrawTextureGlobal(path) differs from DrawTextureGlobal(texture)");
        renderer.DrawTexture(*png);
        const std::vector<uint32_t> point_fetch = back;
        renderer.DrawTexture(grid.c_str());
        EXPECT(back == point_fetch, "DrawTexture(path) differs from DrawTexture(texture)");
        EXPECT(point_fetch[0] == 0xFFFFFFFFu && point_fetch[50] == 0xFF000000u, "DrawTexture: %08x %08x",
               point_fetch[0], point_fetch[50]);
        threw = false;
        try {
            renderer.DrawTexture((root + "/tests/golden/no_such.png").c_str());
        } catch (const rast::Error &e) {
            threw = e.code() == RAST_E_INVALID_ARG;
        }
        EXPECT(threw, "DrawTexture of a missing file throws");
        delete pRenderer;
    }
    // the texture handles outlived their renderer: destroying them now is safe
    tm.RemoveCUDATexture(0);
    EXPECT(tm.GetCUDATexture(0) == nullptr, "texture removed");
    tm.RemoveCUDATexture(1);
    for (uint32_t y = 0; y < H; y += 7)
        for (uint32_t x = 0; x < W; x += 5) {
            const orc_texture ot{ checker.data(), 256, 256 };
            const float uv[2] = { (float)x / (float)W, (float)y / (float)H };
            float c[3];
            orc_sample(&ot, uv, ORC_SAMPLE_POINT, c);
            const uint32_t want = orc_pack_rgba(c), got = drawn[(size_t)y * W + x];
            EXPECT(got == want, "DrawTextureGlobal at (%u, %u): %08x, want %08x", x, y, got, want);
        }

    // the oracle on the same inputs
    orc_mesh om{};
    om.verts = reinterpret_cast<const orc_ivertex *>(verts.data());
    om.num_verts = nv;
    om.indices = idx.data();
    om.num_indices = ni;
    om.topology = ORC_TOPO_LIST;
    for (int i = 0; i < 16; ++i) om.world[i] = (i % 5 == 0) ? 1.f : 0.f;
    om.tex[0] = { checker.data(), 256, 256 };
    orc_camera oc{ { 0.f, 0.f, 3.f }, { 0.f, 0.f, 1.f }, 45.f, 0.1f, 100.f };
    orc_opts oo{ ORC_CULL_BACK, ORC_SAMPLE_POINT, ORC_VIS_PBR, rgba(63, 63, 63), rgba(128, 128, 255),
                 rgba(128, 128, 128), rgba(128, 128, 128) };
    std::vector<uint32_t> ovis(W * H), orgba(W * H);
    std::vector<float> odepth(W * H);
    orc_stats st{};
    EXPECT(orc_render(&om, 1, &oc, &oo, W, H, 4, ovis.data(), odepth.data(), orgba.data(), &st) == 0, "oracle");
    size_t bad_vis = 0, bad_depth = 0, bad_rgba = 0;
    int maxd = 0;
    for (size_t p = 0; p < (size_t)W * H; ++p) {
        bad_vis += vis[p] != ovis[p];
        uint32_t a, b;
        std::memcpy(&a, &depth[p], 4);
        std::memcpy(&b, &odepth[p], 4);
        bad_depth += a != b;
        for (int c = 0; c < 4; ++c) {
            const int d = std::abs((int)((rendered[p] >> (8 * c)) & 255u) - (int)((orgba[p] >> (8 * c)) & 255u));
            maxd = d > maxd ? d : maxd;
            bad_rgba += d > 1;
        }
    }
    std::printf("vis mismatches %zu, depth mismatches %zu, rgba channels off by > 1: %zu (max %d)\n", bad_vis,
                bad_depth, bad_rgba, maxd);
    EXPECT(bad_vis == 0 && bad_depth == 0 && bad_rgba == 0, "frame differs from the oracle");
    std::printf("PASS\n");
    return 0;
}
