// Exercises the C++ host surface (include/pt_tracer.hpp) the way the
// reference application drives CUDAPathTracer (PathTracingApp.cpp:150-290):
// upload a scene, set the corner rays and eye, raytrace passes, read back.
// Checks, on device 0:
//   * raytrace before an upload does nothing (CUDAPathTracer::raytrace
//     returns when the scene is not ready);
//   * three passes through the facade equal three pt_render_pass calls with
//     the same seeds on a second context, bit for bit;
//   * BVHBuilder::build rebuilds and reports its time; an invalid call is
//     logged and returns (the reference's CUDA_ERROR_CHECK behaviour);
//   * the stop frame holds the pass count.
// Prints one line starting with "OK" and exits 0, or prints the failure and
// exits 1.
#include <cstdio>
#include <cstring>
#include <vector>

#include "pt_scenes.h"
#include "pt_tracer.hpp"

namespace {
int fail(const char* what) {
    std::printf("FAIL %s\n", what);
    return 1;
}
}  // namespace

int main() {
    const int W = 48, H = 32;
    const uint32_t n = 8192;
    ptamd::HostSceneDescription desc;
    desc.triangles.resize(n + 2);
    desc.triangleInfo.resize(n + 2);
    const int count = pt_gen_soup(7, n, 1, 0, desc.triangles.data(), desc.triangleInfo.data());
    if (count <= 0) return fail("pt_gen_soup");
    desc.triangles.resize(static_cast<size_t>(count));
    desc.triangleInfo.resize(static_cast<size_t>(count));
    const pt_settings gi{3u, PT_INVALID_INDEX, 1.0f, 1u};

    const float eye[3] = {0.5f, 0.5f, -1.5f}, target[3] = {0.5f, 0.5f, 0.0f}, up[3] = {0.0f, 1.0f, 0.0f};
    float dirs[12];
    if (pt_camera_corners(eye, target, up, 60.0f, float(W) / H, 0.0f, 0.0f, dirs) != 4)
        return fail("pt_camera_corners");
    ptamd::Ray r[4];
    for (int k = 0; k < 4; ++k)
        r[k] = ptamd::Ray{{eye[0], eye[1], eye[2]}, {dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2]}};

    ptamd::PathTracer tracer(0, 2024u);
    if (!tracer.context()) return fail("pt_create");
    tracer.setPrimaryRays(r[0], r[1], r[2], r[3]);
    tracer.setEyePos({eye[0], eye[1], eye[2]});
    tracer.raytrace(W, H);  // no scene yet: nothing happens
    if (tracer.scene()->isReady() || tracer.lastSeed() != 0) return fail("raytrace before upload");

    tracer.upload(desc, gi);
    if (!tracer.scene()->isReady()) return fail("upload");
    uint64_t seeds[3];
    for (int p = 0; p < 3; ++p) {
        tracer.raytrace(W, H);
        if (tracer.lastStatus() != PT_OK) return fail("raytrace");
        seeds[p] = tracer.lastSeed();
    }
    std::vector<double> a, b;
    if (tracer.readAccumulator(a, W, H) != 3) return fail("pass count");

    pt_context* ctx = nullptr;
    if (pt_create(&ctx, 0) != PT_OK) return fail("second context");
    if (pt_upload_scene(ctx, desc.triangles.data(), uint32_t(desc.triangles.size()), desc.triangleInfo.data(),
                        uint32_t(desc.triangleInfo.size()), nullptr, 0, nullptr, nullptr, &gi) != PT_OK ||
        pt_set_primary_rays(ctx, dirs, eye) != PT_OK)
        return fail("direct upload");
    for (int p = 0; p < 3; ++p)
        if (pt_render_pass(ctx, W, H, seeds[p]) != PT_OK) return fail("direct render");
    b.resize(a.size());
    uint32_t spp = 0;
    if (pt_read_accumulator(ctx, b.data(), &spp) != PT_OK || spp != 3) return fail("direct readback");
    pt_destroy(ctx);
    if (std::memcmp(a.data(), b.data(), a.size() * sizeof(double)) != 0) return fail("facade != C-ABI");
    double total = 0.0;
    for (double v : a) total += v;
    if (!(total > 0.0)) return fail("empty image");

    ptamd::BVHBuilder builder(tracer.context());
    const float ms = builder.build();
    if (!(ms > 0.0f)) return fail("BVHBuilder::build");

    tracer.setScreenBuffer(reinterpret_cast<void*>(16), W, H, 8);  // bad pitch: logged, returns
    if (tracer.lastStatus() != PT_E_INVALID_ARG) return fail("error status");

    tracer.setStopAtFrame(4);
    for (int p = 0; p < 3; ++p) tracer.raytrace(W, H);
    if (tracer.readAccumulator(a, W, H) != 4) return fail("stop frame");
    std::vector<float> img;
    tracer.readImage(img, W, H);
    if (img.size() != size_t(W) * H * 4 || img[3] != 1.0f) return fail("readImage");

    std::printf("OK facade passes=3 equal to C-ABI, build %.3f ms, stop frame 4, image sum %.6g\n", ms, total);
    return 0;
}
