// rast.hpp -- C++ classes named after the reference's GPU-side classes,
// implemented over the C ABI of rast.h (header only, C++17).
//
// The reference's public classes (kdedeurw/Custom-GPGPU-Rasterizer, CCR =
// CustomCUDARasterizer/):
//   CUDARenderer        CCR/CUDARenderer.h:34-74
//   CUDAMesh            CCR/CUDAMesh.h:19-48
//   CUDATexture         CCR/CUDATexture.h:4-28
//   CUDATextureManager  CCR/CUDATextureManager.h:6-19
// keep their member names and meaning here.  The app-side types they take
// (WindowHelper, SceneGraph, SceneManager, Mesh, Camera, IPoint2) stay the
// caller's: the members that read them are templates that use only the
// getters the reference calls (GetMeshes, GetVertexBuffer, GetWorldConst,
// GetPosition, GetForward, ...), so a reference caller keeps its own types.
// Differences: errors throw rast::Error (the reference printed and went on,
// CCR/PCH.h:46-56); textures are created from a PNG file (decoded by
// librast_io, include/rast_io.h: link -lrast_io when using the path forms)
// or from RGBA8 texels; Present() copies the frame into
// WindowHelper::pBackBufferPixels instead of an SDL surface.
//
// The classes live in namespace rast; define RAST_REFERENCE_NAMES before
// including this header to also make them visible in the global namespace
// under the same names, as the reference declares them.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rast.h"
#include "rast_io.h"

namespace rast {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc, const char *what)
{
    if (rc == RAST_OK) return;
    char buf[512];
    rast_last_error(buf, sizeof buf);
    throw Error(rc, std::string(what) + ": " + buf);
}

// Values of the reference enums CullingMode, SampleState, VisualisationState
// and PrimitiveTopology (CCR/CullingMode.h:3-8, CCR/SampleState.h:3-7,
// CCR/VisualisationState.h:3-8, CCR/PrimitiveTopology.h:3-7) map one to one
// onto the ABI's integers.
template <class E>
inline int enum_value(E e)
{
    return static_cast<int>(e);
}

class CUDARenderer;

// CUDABenchMarker (CCR/CUDABenchMarker.h:4-18): StartTimer / StopTimer over
// HIP events on the renderer's stream; StopTimer waits for the stop event and
// returns the milliseconds in between.
class CUDABenchMarker {
public:
    explicit CUDABenchMarker(rast_renderer *r = nullptr) : r_(r) {}
    virtual ~CUDABenchMarker() = default;
    virtual void StartTimer() { check(rast_timer_start(r_), "rast_timer_start"); }
    virtual float StopTimer()
    {
        float ms = 0.f;
        check(rast_timer_stop(r_, &ms), "rast_timer_stop");
        return ms;
    }

private:
    rast_renderer *r_;
};

// CUDATexture (CCR/CUDATexture.h:4-28): one RGBA8 texture in HBM.
class CUDATexture {
public:
    CUDATexture() = default;
    CUDATexture(rast_renderer *r, const uint32_t *rgba8, uint32_t width, uint32_t height) { Create(r, rgba8, width, height); }
    // CUDATexture(const std::string& path) (CCR/CUDATexture.cpp:13-17)
    CUDATexture(rast_renderer *r, const std::string &path) { Create(r, path.c_str()); }
    virtual ~CUDATexture() noexcept { Destroy(); }
    CUDATexture(const CUDATexture &) = delete;
    CUDATexture &operator=(const CUDATexture &) = delete;

    bool IsAllocated() const { return tex_ != nullptr; }
    void Create(rast_renderer *r, const uint32_t *rgba8, uint32_t width, uint32_t height)
    {
        Destroy();
        check(rast_texture_create(r, rgba8, width, height, &tex_), "rast_texture_create");
        width_ = width;
        height_ = height;
    }
    // Create(const char* path) (CCR/CUDATexture.cpp:24-45): IMG_Load of a
    // PNG, then the upload; false (nothing allocated) if the file cannot be
    // read or decoded, as the reference returns false when IMG_Load fails.
    bool Create(rast_renderer *r, const char *path)
    {
        uint32_t *texels = nullptr, w = 0, h = 0;
        if (rio_png_load(path, &texels, &w, &h) != RAST_OK) return false;
        Destroy();
        const int rc = rast_texture_create(r, texels, w, h, &tex_);
        rio_free(texels);
        if (rc != RAST_OK) {
            tex_ = nullptr;
            return false;
        }
        width_ = w;
        height_ = h;
        return true;
    }
    virtual void Destroy()
    {
        if (tex_) rast_texture_destroy(tex_);
        tex_ = nullptr;
        width_ = height_ = 0;
    }
    unsigned int GetBytesPerPixel() const { return 4; }
    unsigned int GetWidth() const { return width_; }
    unsigned int GetHeight() const { return height_; }
    rast_texture *GetHandle() const { return tex_; }

private:
    rast_texture *tex_ = nullptr;
    uint32_t width_ = 0, height_ = 0;
};

// CUDATextureManager (CCR/CUDATextureManager.h:6-19): owns textures by id.
class CUDATextureManager final {
public:
    CUDATextureManager() = default;
    ~CUDATextureManager()
    {
        for (CUDATexture *t : textures_) delete t;
    }
    CUDATextureManager(const CUDATextureManager &) = delete;
    CUDATextureManager &operator=(const CUDATextureManager &) = delete;

    CUDATexture *GetCUDATexture(int id) const
    {
        return id >= 0 && (size_t)id < textures_.size() ? textures_[(size_t)id] : nullptr;
    }
    // Takes ownership; returns the id.  A null texture is an error
    // (the reference throws too, CCR/CUDATextureManager.cpp:29-32).
    int AddCUDATexture(CUDATexture *pTex)
    {
        if (!pTex) throw Error(RAST_E_INVALID_ARG, "AddCUDATexture: null texture");
        textures_.push_back(pTex);
        return (int)textures_.size() - 1;
    }
    // An unknown id is ignored, as in CCR/CUDATextureManager.cpp:39-46.
    void RemoveCUDATexture(int id)
    {
        if (id < 0 || (size_t)id >= textures_.size()) return;
        delete textures_[(size_t)id];
        textures_[(size_t)id] = nullptr;
    }

private:
    std::vector<CUDATexture *> textures_;
};

// CUDAMesh (CCR/CUDAMesh.h:19-48): one mesh of a renderer, uploaded once.
class CUDAMesh {
public:
    CUDAMesh(rast_renderer *r, const rast_ivertex *vertices, uint32_t numVertices, const uint32_t *indices,
             uint32_t numIndices, int topology, const float world[16] = nullptr)
        : num_vertices_(numVertices), num_indices_(numIndices), topology_(topology)
    {
        check(rast_mesh_create(r, vertices, numVertices, indices, numIndices, topology, world, &mesh_), "rast_mesh_create");
        if (world) std::memcpy(world_, world, sizeof world_);
        else for (int i = 0; i < 16; ++i) world_[i] = (i % 5 == 0) ? 1.f : 0.f;
    }
    // From a reference-style Mesh: GetVertexBuffer (56-byte IVertex records),
    // GetNumVertices, GetIndexBuffer, GetNumIndices, GetTopology, GetWorldConst
    // (data[col][row]) (CCR/Mesh.h:23-37, CCR/CUDAMesh.cpp:4-25).
    template <class MeshT>
    CUDAMesh(rast_renderer *r, const MeshT *pMesh)
        : CUDAMesh(r, reinterpret_cast<const rast_ivertex *>(pMesh->GetVertexBuffer()), pMesh->GetNumVertices(),
                   pMesh->GetIndexBuffer(), pMesh->GetNumIndices(), enum_value(pMesh->GetTopology()),
                   world_of(pMesh->GetWorldConst()).data())
    {
        static_assert(sizeof(*pMesh->GetVertexBuffer()) == sizeof(rast_ivertex), "IVertex must be 56 bytes");
    }
    virtual ~CUDAMesh()
    {
        if (mesh_) rast_mesh_destroy(mesh_);
    }
    CUDAMesh(const CUDAMesh &) = delete;
    CUDAMesh &operator=(const CUDAMesh &) = delete;

    const float *GetWorldConst() const { return world_; }
    void SetWorld(const float world[16])
    {
        std::memcpy(world_, world, sizeof world_);
        check(rast_mesh_set_world(mesh_, world_), "rast_mesh_set_world");
    }
    int GetTopology() const { return topology_; }
    unsigned int GetNumVertices() const { return num_vertices_; }
    unsigned int GetNumIndices() const { return num_indices_; }
    unsigned int GetTotalNumTriangles() const
    {
        uint32_t n = 0;
        check(rast_mesh_num_triangles(mesh_, &n), "rast_mesh_num_triangles");
        return n;
    }
    // SetTextures(CUDATexturesCompact) (CCR/CUDAMesh.h:42): diffuse, normal,
    // specular, glossiness; null = the 1x1 stand-in.
    void SetTextures(const CUDATexture *d, const CUDATexture *n = nullptr, const CUDATexture *s = nullptr,
                     const CUDATexture *g = nullptr)
    {
        auto h = [](const CUDATexture *t) { return t ? t->GetHandle() : nullptr; };
        check(rast_mesh_set_textures(mesh_, h(d), h(n), h(s), h(g)), "rast_mesh_set_textures");
    }
    rast_mesh *GetHandle() const { return mesh_; }

    template <class M4>
    static std::vector<float> world_of(const M4 &m)
    {
        std::vector<float> w(16);
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) w[(size_t)c * 4 + r] = m.data[c][r];
        return w;
    }

private:
    rast_mesh *mesh_ = nullptr;
    uint32_t num_vertices_ = 0, num_indices_ = 0;
    int topology_ = RAST_TOPOLOGY_LIST;
    float world_[16];
};

// CUDARenderer (CCR/CUDARenderer.h:31-74): one framebuffer on one GPU.
class CUDARenderer final {
public:
    CUDARenderer(uint32_t width, uint32_t height, int device = 0)
    {
        check(rast_renderer_create(device, width, height, &r_), "rast_renderer_create");
        width_ = width;
        height_ = height;
        bench_ = CUDABenchMarker(r_);
    }
    // CUDARenderer(const WindowHelper&): the resolution, and the back buffer
    // that Present() fills (RGBA8, width x height).
    template <class WindowHelperT, class = decltype(std::declval<const WindowHelperT &>().Resolution.Width)>
    explicit CUDARenderer(const WindowHelperT &wh) : CUDARenderer(wh.Resolution.Width, wh.Resolution.Height)
    {
        back_buffer_ = wh.pBackBufferPixels;
    }
    ~CUDARenderer() noexcept
    {
        meshes_.clear();
        if (r_) rast_renderer_destroy(r_);
    }
    CUDARenderer(const CUDARenderer &) = delete;
    CUDARenderer(CUDARenderer &&) noexcept = delete;
    CUDARenderer &operator=(const CUDARenderer &) = delete;
    CUDARenderer &operator=(CUDARenderer &&) noexcept = delete;

    // LoadScene(const SceneGraph*, const CUDATextureManager&)
    // (CCR/CUDARenderer.cu:178-198): uploads every mesh of the scene graph and
    // binds its texture ids (Mesh::GetTextureIds, -1 = none).
    template <class SceneGraphT>
    void LoadScene(const SceneGraphT *pSceneGraph, const CUDATextureManager &tm)
    {
        if (!pSceneGraph) throw Error(RAST_E_INVALID_ARG, "LoadScene: null scene graph");
        meshes_.clear();
        for (const auto *pMesh : pSceneGraph->GetMeshes()) {
            CUDAMesh &m = AddMesh(pMesh);
            const int *ids = pMesh->GetTextureIds();
            m.SetTextures(tm.GetCUDATexture(ids[0]), tm.GetCUDATexture(ids[1]), tm.GetCUDATexture(ids[2]),
                          tm.GetCUDATexture(ids[3]));
        }
    }
    template <class MeshT>
    CUDAMesh &AddMesh(const MeshT *pMesh)
    {
        meshes_.emplace_back(new CUDAMesh(r_, pMesh));
        return *meshes_.back();
    }
    CUDAMesh &AddMesh(const rast_ivertex *v, uint32_t nv, const uint32_t *idx, uint32_t ni, int topology,
                      const float world[16] = nullptr)
    {
        meshes_.emplace_back(new CUDAMesh(r_, v, nv, idx, ni, topology, world));
        return *meshes_.back();
    }
    CUDATexture *CreateTexture(const uint32_t *rgba8, uint32_t width, uint32_t height)
    {
        return new CUDATexture(r_, rgba8, width, height);
    }
    // A texture from a PNG file (the reference's CUDATexture(path)); null if
    // the file cannot be read or decoded.
    CUDATexture *CreateTexture(const std::string &path)
    {
        std::unique_ptr<CUDATexture> t(new CUDATexture());
        if (!t->Create(r_, path.c_str())) return nullptr;
        return t.release();
    }

    // SetupBins(numBins, binDim, binQueueMaxSize) (CCR/CUDARenderer.cu:200-205):
    // square tiles of binDim (8, 16 or 32 pixels); the queue depth is sized
    // exactly per frame here, so binQueueMaxSize is the screen-list span.
    template <class IPoint2T>
    void SetupBins(const IPoint2T &numBins, const IPoint2T &binDim, unsigned int binQueueMaxSize)
    {
        (void)numBins;
        if (binDim.x != binDim.y) throw Error(RAST_E_INVALID_ARG, "SetupBins: tiles are square");
        check(rast_setup_bins(r_, (uint32_t)binDim.x, binQueueMaxSize), "rast_setup_bins");
    }
    void SetupBins(uint32_t tile, uint32_t bigThreshold = 0)
    {
        check(rast_setup_bins(r_, tile, bigThreshold), "rast_setup_bins");
    }
    void SetTuning(uint32_t microLimit, uint32_t smallLimit)
    {
        check(rast_set_tuning(r_, microLimit, smallLimit), "rast_set_tuning");
    }

    // EnterValidRenderingState (CCR/CUDARenderer.cu:374-382): nothing to lock.
    int EnterValidRenderingState() { return 0; }

    // Render(const SceneManager&, const CUDATextureManager&, const Camera*)
    // (CCR/CUDARenderer.cu:207-355): culling, sampling and visualisation from
    // the scene manager, position / forward / fov / near-far from the camera.
    template <class SceneManagerT, class CameraT>
    void Render(const SceneManagerT &sm, const CUDATextureManager &tm, const CameraT *pCamera)
    {
        (void)tm;
        rast_frame_opts o = DefaultOptions();
        o.cull_mode = enum_value(sm.GetCullingMode());
        o.sample_state = enum_value(sm.GetSampleState());
        o.visualisation = enum_value(sm.GetVisualisationState());
        Render(CameraOf(*pCamera), o);
    }
    void Render(const rast_camera &cam, const rast_frame_opts &opts)
    {
        check(rast_render(r_, &cam, &opts), "rast_render");
    }
    // Present (CCR/CUDARenderer.cu:384-396): the frame's RGBA8 into the back buffer.
    void Present()
    {
        if (back_buffer_) Readback(nullptr, nullptr, back_buffer_);
        else check(rast_sync(r_), "rast_sync");
    }
    // RenderAuto (CCR/CUDARenderer.cu:357-372)
    template <class SceneManagerT, class CameraT>
    void RenderAuto(const SceneManagerT &sm, const CUDATextureManager &tm, const CameraT *pCamera)
    {
        Render(sm, tm, pCamera);
        Present();
    }
    void Readback(uint32_t *vis, float *depth, uint32_t *rgba)
    {
        check(rast_readback(r_, vis, depth, rgba), "rast_readback");
    }
    rast_stats Stats()
    {
        rast_stats s;
        check(rast_get_stats(r_, &s), "rast_get_stats");
        return s;
    }
    unsigned int GetTotalNumVisibleTriangles() { return (unsigned int)Stats().num_visible; }
    unsigned int GetTotalNumTriangles() { return (unsigned int)Stats().num_triangles; }

    // DisplayGPUSpecs (CCR/CUDARenderer.cu:60-176)
    void DisplayGPUSpecs(int deviceId = 0)
    {
        rast_device_info d;
        check(rast_device_query(deviceId, &d), "rast_device_query");
        std::printf("%s (%s): %d CUs, %d MHz, %.1f GiB, wave%d, %d B LDS per workgroup\n", d.name, d.arch,
                    d.compute_units, d.clock_khz / 1000, d.global_mem_bytes / 1073741824.0, d.wavefront_size,
                    d.lds_bytes_per_block);
    }
    // KernelWarmUp (CCR/CUDARenderer.cu:2015-2025): kernels are loaded with
    // the library; nothing to warm.
    void KernelWarmUp() {}
    // StartTimer / StopTimer / GetBenchMarker (CCR/CUDARenderer.cu:398-412,
    // CCR/CUDARenderer.h:71-73): HIP events on the renderer's stream.
    void StartTimer() { bench_.StartTimer(); }
    float StopTimer() { return bench_.StopTimer(); }
    CUDABenchMarker &GetBenchMarker() { return bench_; }

    // Test hooks DrawTexture / DrawTextureGlobal (CCR/CUDARenderer.cu:1969-2012):
    // the frame becomes the texture, then Present().  The reference takes an
    // image path (the char* overloads below: the PNG is loaded, drawn and
    // released); the CUDATexture overloads draw a texture already made.
    void DrawTexture(const CUDATexture &tex)
    {
        check(rast_draw_texture(r_, tex.GetHandle(), RAST_DRAW_TEXTURE, 1, RAST_SAMPLE_POINT), "rast_draw_texture");
        Present();
    }
    template <class SampleStateT = int>
    void DrawTextureGlobal(const CUDATexture &tex, bool isStretchedToWindow = true, SampleStateT sampleState = SampleStateT(0))
    {
        check(rast_draw_texture(r_, tex.GetHandle(), RAST_DRAW_TEXTURE_GLOBAL, isStretchedToWindow ? 1 : 0,
                                enum_value(sampleState)),
              "rast_draw_texture");
        Present();
    }
    void DrawTexture(const char *tp) { DrawTexture(*LoadForDraw(tp)); }
    template <class SampleStateT = int>
    void DrawTextureGlobal(const char *tp, bool isStretchedToWindow = true, SampleStateT sampleState = SampleStateT(0))
    {
        DrawTextureGlobal(*LoadForDraw(tp), isStretchedToWindow, sampleState);
    }

    rast_renderer *GetHandle() const { return r_; }
    uint32_t GetWidth() const { return width_; }
    uint32_t GetHeight() const { return height_; }

    static rast_frame_opts DefaultOptions()
    {
        rast_frame_opts o{};
        o.cull_mode = RAST_CULL_BACK;
        o.sample_state = RAST_SAMPLE_POINT;
        o.visualisation = RAST_VIS_PBR;
        o.bin_policy = RAST_BIN_EXACT;
        o.clear_rgba = 0xFF3F3F3Fu;  // RGBColor(0.25, 0.25, 0.25) packed by (u8)(c * 255)
        o.default_normal = 0xFFFF8080u;
        o.default_specular = 0xFF808080u;
        o.default_gloss = 0xFF808080u;
        o.flags = 0;
        return o;
    }
    // Camera (CCR/Camera.h:4-60): GetPosition, GetForward (.x .y .z), GetFov,
    // GetNearFarValues(float&, float&).
    template <class CameraT>
    static rast_camera CameraOf(const CameraT &c)
    {
        rast_camera cam{};
        const auto &p = c.GetPosition();
        const auto &f = c.GetForward();
        cam.position[0] = p.x;
        cam.position[1] = p.y;
        cam.position[2] = p.z;
        cam.forward[0] = f.x;
        cam.forward[1] = f.y;
        cam.forward[2] = f.z;
        cam.fov_deg = c.GetFov();
        float n = 0.1f, fa = 100.f;
        const_cast<CameraT &>(c).GetNearFarValues(n, fa);
        cam.near_z = n;
        cam.far_z = fa;
        return cam;
    }

private:
    // the char* draw hooks' texture: the reference drew whatever Create left
    // (nothing, for a missing file); here a file that does not load throws
    std::unique_ptr<CUDATexture> LoadForDraw(const char *tp)
    {
        std::unique_ptr<CUDATexture> t(new CUDATexture());
        if (!tp || !t->Create(r_, tp))
            throw Error(RAST_E_INVALID_ARG, std::string("cannot load texture ") + (tp ? tp : "(null)"));
        return t;
    }

    rast_renderer *r_ = nullptr;
    uint32_t width_ = 0, height_ = 0;
    uint32_t *back_buffer_ = nullptr;
    std::vector<std::unique_ptr<CUDAMesh>> meshes_;
    CUDABenchMarker bench_;
};

}  // namespace rast

#ifdef RAST_REFERENCE_NAMES
using rast::CUDABenchMarker;
using rast::CUDAMesh;
using rast::CUDARenderer;
using rast::CUDATexture;
using rast::CUDATextureManager;
#endif
