// pt_tracer.hpp -- C++ host surface over the C-ABI (pt_api.h) with the
// reference engine's class and method names, for code written against
// CUDAPathTracer (source/application/cuda/cuda_path_tracer.h:13-40),
// CUDA::Scene (Scene.h:19-50), CUDA::HostSceneDescription
// (HostSceneDescription.h:10-17) and CUDA::BVHBuilder (BVHBuilder.h:12-19).
// Header only; link against libptamd.
//
// Differences a caller sees:
//   * vectors are ptamd::vec3 (three floats, the layout of glm::vec3);
//   * the screen is a caller-owned device RGBA32F buffer (setScreenBuffer)
//     instead of a CUDA array bound to a surface object;
//   * the per-pass seed comes from a std::mt19937 seeded in the constructor
//     (the reference's Random::getUInt draws from an mt19937 seeded from
//     std::random_device, so its runs are not reproducible; pass a seed
//     from std::random_device for the same behaviour);
//   * errors follow the reference's CUDA_ERROR_CHECK (cuda_common.h:4-12):
//     they are logged to stderr and the call returns; lastStatus() holds the
//     status of the last library call.
#ifndef PT_TRACER_HPP
#define PT_TRACER_HPP

#include <cstdint>
#include <cstdio>
#include <memory>
#include <random>
#include <vector>

#include "pt_api.h"

namespace ptamd {

struct vec3 {
    float x, y, z;
};

// CUDA::Ray (geometry.h): origin + direction.
struct Ray {
    vec3 origin, direction;
};

// CUDA::HostSceneDescription plus the surface detail the reference keeps in
// its TriangleInfo vertices and texture arrays.
struct HostSceneDescription {
    std::vector<pt_triangle> triangles;
    std::vector<pt_triangle_info> triangleInfo;  // indexed by triangle id
    bool hasEnclosingBBox = false;               // false: the union of the triangle boxes
    vec3 enclosingMin{0, 0, 0}, enclosingMax{0, 0, 0};
    std::vector<pt_dir_light> dirLights;
    std::vector<pt_triangle_detail> detail;      // empty: no surface detail
    std::vector<pt_texture> textures;            // texels owned by the caller
};

namespace detail {
inline int report(pt_context* ctx, int rc, const char* what) {
    if (rc != PT_OK) {
        char msg[256] = {0};
        if (ctx) pt_last_error(ctx, msg, sizeof msg);
        std::fprintf(stderr, "ptamd: %s failed (%d): %s\n", what, rc, msg);
    }
    return rc;
}
}  // namespace detail

class PathTracer;

// CUDA::Scene: the uploaded geometry, its LBVH and shading data.
class Scene {
public:
    // Scene::upload (Scene.cu:28-55): copies the description and builds the LBVH.
    void upload(const HostSceneDescription& desc, const pt_settings& settings) {
        const float* bmin = desc.hasEnclosingBBox ? &desc.enclosingMin.x : nullptr;
        const float* bmax = desc.hasEnclosingBBox ? &desc.enclosingMax.x : nullptr;
        m_status = detail::report(
            m_ctx,
            pt_upload_scene(m_ctx, desc.triangles.data(), static_cast<uint32_t>(desc.triangles.size()),
                            desc.triangleInfo.data(), static_cast<uint32_t>(desc.triangleInfo.size()),
                            desc.dirLights.data(), static_cast<uint32_t>(desc.dirLights.size()), bmin, bmax,
                            &settings),
            "Scene::upload");
        m_ready = m_status == PT_OK;
        if (m_ready && !desc.detail.empty()) upload(desc.detail, desc.textures);
    }
    // Scene::upload(textureArrays) (Scene.cu:57-79), with the vertex data the
    // textures are sampled through.
    void upload(const std::vector<pt_triangle_detail>& detail, const std::vector<pt_texture>& textures) {
        m_status = detail::report(m_ctx,
                                  pt_upload_detail(m_ctx, detail.data(), static_cast<uint32_t>(detail.size()),
                                                   textures.data(), static_cast<uint32_t>(textures.size())),
                                  "Scene::upload(textures)");
    }
    bool isReady() const { return m_ready; }
    int lastStatus() const { return m_status; }

private:
    friend class PathTracer;
    explicit Scene(pt_context* ctx) : m_ctx(ctx) {}
    pt_context* m_ctx;
    bool m_ready = false;
    int m_status = PT_OK;
};

// CUDA::BVHBuilder: rebuilds the LBVH of the context's uploaded scene(s).
class BVHBuilder {
public:
    explicit BVHBuilder(pt_context* ctx) : m_ctx(ctx) {}
    // BVHBuilder::build (BVHBuilder.cu:207-269); returns the build time in
    // ms (HIP events), or a negative value on error.
    float build() {
        float ms = -1.0f;
        if (detail::report(m_ctx, pt_build_bvh(m_ctx, &ms), "BVHBuilder::build") != PT_OK) return -1.0f;
        return ms;
    }

private:
    pt_context* m_ctx;
};

// CUDAPathTracer: progressive 1-spp passes into the FP64 accumulator.
class PathTracer {
public:
    explicit PathTracer(int device = 0, uint32_t seed = 5489u) : m_rng(seed) {
        if (detail::report(nullptr, pt_create(&m_ctx, device), "pt_create") != PT_OK) m_ctx = nullptr;
        m_scene.reset(new Scene(m_ctx));
    }
    ~PathTracer() {
        if (m_ctx) pt_destroy(m_ctx);
    }
    PathTracer(const PathTracer&) = delete;
    PathTracer& operator=(const PathTracer&) = delete;

    // setScreenTexture (cuda_path_tracer.cu:343-364): a device RGBA32F
    // buffer of w x h texels, rows `pitch` bytes apart, written every pass.
    void setScreenBuffer(void* d_rgba, uint32_t w, uint32_t h, size_t pitch) {
        m_status = detail::report(m_ctx, pt_set_screen_buffer(m_ctx, d_rgba, w, h, pitch), "setScreenBuffer");
    }
    // setPrimaryRays(ray00, ray10, ray11, ray01): corner order (0,0), (1,0), (1,1), (0,1).
    void setPrimaryRays(const Ray& ray00, const Ray& ray10, const Ray& ray11, const Ray& ray01) {
        const Ray* r[4] = {&ray00, &ray10, &ray11, &ray01};
        for (int k = 0; k < 4; ++k) {
            m_dirs[3 * k] = r[k]->direction.x;
            m_dirs[3 * k + 1] = r[k]->direction.y;
            m_dirs[3 * k + 2] = r[k]->direction.z;
        }
    }
    void setEyePos(const vec3& eye) { m_eye = eye; }

    // raytrace (cuda_path_tracer.cu:423-464): one pass when the scene is
    // ready, the seed drawn from the tracer's generator.
    void raytrace(int screenWidth, int screenHeight) {
        if (!m_scene->isReady()) return;
        m_lastSeed = m_rng();
        m_status = detail::report(m_ctx, pt_set_primary_rays(m_ctx, m_dirs, &m_eye.x), "setPrimaryRays");
        if (m_status != PT_OK) return;
        m_status = detail::report(m_ctx, pt_render_pass(m_ctx, screenWidth, screenHeight, m_lastSeed), "raytrace");
    }

    // upload (cuda_path_tracer.cu:466-470): scene upload, frame counter reset.
    void upload(const HostSceneDescription& desc, const pt_settings& settings) {
        m_scene->upload(desc, settings);
        m_status = m_scene->lastStatus();
    }
    Scene* scene() const { return m_scene.get(); }
    void resetFrames() { m_status = detail::report(m_ctx, pt_reset_frames(m_ctx), "resetFrames"); }
    // PathTracerSettings::DEMO.stopAtFrame (0: never).
    void setStopAtFrame(uint32_t frames) {
        m_status = detail::report(m_ctx, pt_set_stop_frame(m_ctx, frames), "setStopAtFrame");
    }

    // Readback: the running-mean image (rgba, w * h * 4 floats) and the FP64
    // sums with the pass count.
    void readImage(std::vector<float>& rgba, int w, int h) {
        rgba.resize(static_cast<size_t>(w) * h * 4);
        m_status = detail::report(m_ctx, pt_read_image(m_ctx, rgba.data()), "readImage");
    }
    uint32_t readAccumulator(std::vector<double>& sums, int w, int h) {
        sums.resize(static_cast<size_t>(w) * h * 3);
        uint32_t spp = 0;
        m_status = detail::report(m_ctx, pt_read_accumulator(m_ctx, sums.data(), &spp), "readAccumulator");
        return spp;
    }

    pt_context* context() const { return m_ctx; }
    uint64_t lastSeed() const { return m_lastSeed; }
    int lastStatus() const { return m_status; }

private:
    pt_context* m_ctx = nullptr;
    std::unique_ptr<Scene> m_scene;
    std::mt19937 m_rng;
    float m_dirs[12] = {0};
    vec3 m_eye{0, 0, 0};
    uint64_t m_lastSeed = 0;
    int m_status = PT_OK;
};

}  // namespace ptamd

#endif  // PT_TRACER_HPP
