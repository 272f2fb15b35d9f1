// C-ABI entry points of libptamd (declared in include/pt_api.h).  Each entry
// converts library exceptions into a status code plus a message retrievable
// with pt_last_error; nothing throws across the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "context.hpp"
#include "pt_api.h"
#include "pt_scenes.h"

namespace {

thread_local std::string g_last_error;

template <class F>
int guarded(pt_context* ctx, F&& f) {
    try {
        if (ctx) PT_HIP(hipSetDevice(ctx->device));
        f();
        return PT_OK;
    } catch (const pt::Error& e) {
        if (ctx) ctx->last_error = e.what();
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->last_error = e.what();
        g_last_error = e.what();
        return PT_E_INTERNAL;
    } catch (...) {
        if (ctx) ctx->last_error = "unknown error";
        g_last_error = "unknown error";
        return PT_E_INTERNAL;
    }
}

void require(bool cond, const char* msg) {
    if (!cond) throw pt::Error(PT_E_INVALID_ARG, msg);
}

hipStream_t pick(pt_context* ctx, void* stream) {
    return stream ? static_cast<hipStream_t>(stream) : ctx->stream;
}

// The ray queue block of closest-hit launches on stream `st` (allocated on
// the stream's first trace).
uint32_t* trace_queue(pt_context* ctx, hipStream_t st) {
    pt::DevBuf& q = ctx->trace_queues[st];
    if (q.capacity() == 0) q.reserve(pt::trace_queue_bytes());
    return q.as<uint32_t>();
}

// Uploads and builds run on the context stream; a trace on another stream
// waits for the last of them.  A stream under graph capture is left alone:
// work captured into a graph is ordered by whoever launches the graph.
void order_after_build(pt_context* ctx, hipStream_t st) {
    if (st == ctx->stream || !ctx->ready_ev) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    PT_HIP(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) return;
    PT_HIP(hipStreamWaitEvent(st, ctx->ready_ev, 0));
}

void mark_ready(pt_context* ctx) {
    if (!ctx->ready_ev) PT_HIP(hipEventCreateWithFlags(&ctx->ready_ev, hipEventDisableTiming));
    PT_HIP(hipEventRecord(ctx->ready_ev, ctx->stream));
}

// Reads and clears the scene's status word after the context stream's work.
uint32_t take_status(pt_context* ctx) {
    pt::DevBuf& sb = ctx->scene.status;
    if (sb.capacity() == 0) return 0u;
    uint32_t flags = 0;
    PT_HIP(hipMemcpyAsync(&flags, sb.as<void>(), sizeof flags, hipMemcpyDeviceToHost, ctx->stream));
    PT_HIP(hipStreamSynchronize(ctx->stream));
    if (flags) {
        PT_HIP(hipMemsetAsync(sb.as<void>(), 0, sizeof(uint32_t), ctx->stream));
        PT_HIP(hipStreamSynchronize(ctx->stream));
    }
    return flags;
}

// At a synchronisation point: a raised status flag becomes PT_E_INTERNAL.
void check_status(pt_context* ctx) {
    const uint32_t f = take_status(ctx);
    if (f == 0u) return;
    std::string m = "device status:";
    if (f & PT_STATUS_STACK_OVERFLOW) m += " traversal stack deeper than 64 entries;";
    if (f & PT_STATUS_REFIT_OVERFLOW) m += " refit level list overflow;";
    throw pt::Error(PT_E_INTERNAL, m + " results of the affected calls are incomplete");
}

}  // namespace

extern "C" {

int pt_create(pt_context** out, int device) {
    if (!out) return PT_E_INVALID_ARG;
    *out = nullptr;
    pt_context* ctx = nullptr;
    int rc = guarded(nullptr, [&] {
        int count = 0;
        hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess || count <= 0) throw pt::Error(PT_E_NO_DEVICE, "no HIP device visible");
        if (device < 0 || device >= count) throw pt::Error(PT_E_NO_DEVICE, "device index out of range");
        PT_HIP(hipSetDevice(device));
        pt::trace_scratch_setup(device);
        ctx = new pt_context();
        ctx->device = device;
        PT_HIP(hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking));
        ctx->stream = ctx->own_stream;
        PT_HIP(hipEventCreate(&ctx->ev0));
        PT_HIP(hipEventCreate(&ctx->ev1));
    });
    if (rc != PT_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return PT_OK;
}

int pt_destroy(pt_context* ctx) {
    if (!ctx) return PT_OK;
    (void)hipSetDevice(ctx->device);
    delete ctx;  // waits for its work, releases events, stream and buffers
    return PT_OK;
}

int pt_last_error(const pt_context* ctx, char* buf, size_t len) {
    if (!buf || len == 0) return PT_E_INVALID_ARG;
    const std::string& s = ctx ? ctx->last_error : g_last_error;
    size_t k = std::min(len - 1, s.size());
    std::memcpy(buf, s.data(), k);
    buf[k] = '\0';
    return PT_OK;
}

int pt_set_stream(pt_context* ctx, void* stream) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        PT_HIP(hipStreamSynchronize(ctx->stream));
        ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
    });
}

int pt_synchronize(pt_context* ctx) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        PT_HIP(hipStreamSynchronize(ctx->stream));
        check_status(ctx);
    });
}

int pt_get_status(pt_context* ctx, uint32_t* flags) {
    if (!ctx || !flags) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] { *flags = take_status(ctx); });
}

namespace {
// Scene::upload (Scene.cu:28-55) for one or more scenes: deep copy, per-scene
// enclosing boxes, then the LBVH build.  boxes: 6 floats per scene (min.xyz,
// max.xyz) or NULL for the union of each scene's triangle boxes.
void upload_scenes(pt_context* ctx, const pt_triangle* tris, const uint32_t* offs, uint32_t nscenes,
                   const pt_triangle_info* infos, uint32_t ninfos, const pt_dir_light* lights,
                   uint32_t nlights, const float* boxes, const pt_settings* settings) {
    require(nscenes >= 1 && offs != nullptr, "at least one scene");
    require(offs[0] == 0, "scene offsets must start at 0");
    for (uint32_t k = 0; k < nscenes; ++k) require(offs[k] <= offs[k + 1], "scene offsets must not decrease");
    const uint32_t n = offs[nscenes];
    require(n == 0 || tris != nullptr, "tris is NULL");
    require(ninfos == 0 || infos != nullptr, "infos is NULL");
    require(nlights == 0 || lights != nullptr, "lights is NULL");
    require(nlights <= 64, "at most 64 directional lights");
    require(n < (1u << 30), "at most 2^30 - 1 triangles");
    uint32_t max_id = 0;
    for (uint32_t i = 0; i < n; ++i) max_id = std::max(max_id, tris[i].id);
    require(n == 0 || (infos != nullptr && ninfos > max_id),
            "infos must hold a record for every triangle id");
    pt::DeviceScene& s = ctx->scene;
    hipStream_t st = ctx->stream;
    PT_HIP(hipStreamSynchronize(st));
    s.ready = false;
    s.n = n;
    s.ninfos = ninfos;
    s.max_id = max_id;
    s.ndetail = 0;  // surface detail belongs to the previous upload
    s.ntex = 0;
    s.off_host.assign(offs, offs + nscenes + 1);
    s.box_host.assign(static_cast<size_t>(nscenes) * 6, 0.0f);
    std::vector<float> mbox(static_cast<size_t>(nscenes) * 6);
    for (uint32_t k = 0; k < nscenes; ++k) {
        float* b = &s.box_host[6 * static_cast<size_t>(k)];
        if (boxes) {
            std::memcpy(b, boxes + 6 * static_cast<size_t>(k), 6 * sizeof(float));
        } else {
            pt_scene_bounds(tris + offs[k], offs[k + 1] - offs[k], b, b + 3);
        }
        // invScale = 1 / scale on the host (BVHBuilder.cu:15-17).
        for (int a = 0; a < 3; ++a) {
            mbox[6 * static_cast<size_t>(k) + a] = b[a];
            mbox[6 * static_cast<size_t>(k) + 3 + a] = 1.0f / (b[3 + a] - b[a]);
        }
    }
    if (s.status.capacity() == 0) s.status.reserve(256);
    PT_HIP(hipMemsetAsync(s.status.as<void>(), 0, sizeof(uint32_t), st));
    s.off.reserve((nscenes + 1) * sizeof(uint32_t));
    s.morton_box.reserve(mbox.size() * sizeof(float));
    PT_HIP(hipMemcpyAsync(s.off.as<void>(), offs, (nscenes + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    PT_HIP(hipMemcpyAsync(s.morton_box.as<void>(), mbox.data(), mbox.size() * sizeof(float),
                          hipMemcpyHostToDevice, st));
    if (n) {
        s.tris.reserve(sizeof(pt_triangle) * static_cast<size_t>(n));
        PT_HIP(hipMemcpyAsync(s.tris.as<void>(), tris, sizeof(pt_triangle) * static_cast<size_t>(n),
                              hipMemcpyHostToDevice, st));
    }
    if (ninfos) {
        s.infos.reserve(sizeof(pt_triangle_info) * static_cast<size_t>(ninfos));
        PT_HIP(hipMemcpyAsync(s.infos.as<void>(), infos, sizeof(pt_triangle_info) * static_cast<size_t>(ninfos),
                              hipMemcpyHostToDevice, st));
    }
    s.nlights = nlights;
    if (nlights) {
        s.lights.reserve(sizeof(pt_dir_light) * nlights);
        PT_HIP(hipMemcpyAsync(s.lights.as<void>(), lights, sizeof(pt_dir_light) * nlights, hipMemcpyHostToDevice,
                              st));
    }
    if (settings) ctx->renderer.settings = *settings;
    pt::build_lbvh(s, st);
    mark_ready(ctx);
    PT_HIP(hipStreamSynchronize(st));
    check_status(ctx);
    s.ready = true;
    // CUDAPathTracer::upload resets the frame counter (cuda_path_tracer.cu:469).
    pt::render_reset(ctx->renderer, st);
    PT_HIP(hipStreamSynchronize(st));
}
}  // namespace

int pt_upload_scene(pt_context* ctx, const pt_triangle* tris, uint32_t n,
                    const pt_triangle_info* infos, uint32_t ninfos, const pt_dir_light* lights,
                    uint32_t nlights, const float* bmin, const float* bmax,
                    const pt_settings* settings) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        require((bmin == nullptr) == (bmax == nullptr), "bmin and bmax must both be given or both NULL");
        const uint32_t offs[2] = {0, n};
        float box[6];
        if (bmin) {
            std::memcpy(box, bmin, 3 * sizeof(float));
            std::memcpy(box + 3, bmax, 3 * sizeof(float));
        }
        upload_scenes(ctx, tris, offs, 1, infos, ninfos, lights, nlights, bmin ? box : nullptr, settings);
    });
}

int pt_upload_scenes(pt_context* ctx, const pt_triangle* tris, const uint32_t* scene_offsets,
                     uint32_t nscenes, const pt_triangle_info* infos, uint32_t ninfos,
                     const float* boxes, const pt_settings* settings) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        upload_scenes(ctx, tris, scene_offsets, nscenes, infos, ninfos, nullptr, 0, boxes, settings);
    });
}

static_assert(sizeof(pt_triangle_detail) == 144, "pt_triangle_detail layout");

int pt_upload_detail(pt_context* ctx, const pt_triangle_detail* detail, uint32_t n,
                     const pt_texture* textures, uint32_t ntextures) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        pt::DeviceScene& s = ctx->scene;
        if (!s.ready) throw pt::Error(PT_E_NO_SCENE, "no scene uploaded");
        hipStream_t st = ctx->stream;
        PT_HIP(hipStreamSynchronize(st));
        if (n == 0) {
            s.ndetail = 0;
            s.ntex = 0;
            return;
        }
        require(detail != nullptr, "detail is NULL");
        require(s.n == 0 || n > s.max_id, "detail must hold a record for every triangle id");
        require(ntextures == 0 || textures != nullptr, "textures is NULL");
        std::vector<uint32_t> info(4 * static_cast<size_t>(ntextures));
        uint64_t total = 0;
        for (uint32_t k = 0; k < ntextures; ++k) {
            const pt_texture& t = textures[k];
            require(t.width >= 1 && t.width <= 16384 && t.height >= 1 && t.height <= 16384,
                    "texture sizes must be 1..16384");
            require(t.rgba != nullptr, "texture texels are NULL");
            info[4 * k] = static_cast<uint32_t>(total);
            info[4 * k + 1] = t.width;
            info[4 * k + 2] = t.height;
            info[4 * k + 3] = 0;
            total += static_cast<uint64_t>(t.width) * t.height;
            require(total < (1ull << 31), "at most 2^31 - 1 texels in all textures");
        }
        for (uint32_t i = 0; i < n; ++i) {
            require(detail[i].albedo_tex == PT_NO_TEXTURE || detail[i].albedo_tex < ntextures,
                    "albedo texture index out of range");
            require(detail[i].normal_tex == PT_NO_TEXTURE || detail[i].normal_tex < ntextures,
                    "normal map index out of range");
        }
        s.ndetail = 0;
        s.ntex = 0;
        s.detail.reserve(sizeof(pt_triangle_detail) * static_cast<size_t>(n));
        PT_HIP(hipMemcpyAsync(s.detail.as<void>(), detail, sizeof(pt_triangle_detail) * static_cast<size_t>(n),
                              hipMemcpyHostToDevice, st));
        if (ntextures) {
            s.texels.reserve(static_cast<size_t>(total) * 4);
            s.texinfo.reserve(info.size() * sizeof(uint32_t));
            for (uint32_t k = 0; k < ntextures; ++k) {
                PT_HIP(hipMemcpyAsync(s.texels.as<uint32_t>() + info[4 * k], textures[k].rgba,
                                      static_cast<size_t>(textures[k].width) * textures[k].height * 4,
                                      hipMemcpyHostToDevice, st));
            }
            PT_HIP(hipMemcpyAsync(s.texinfo.as<void>(), info.data(), info.size() * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, st));
        }
        PT_HIP(hipStreamSynchronize(st));
        s.ndetail = n;
        s.ntex = ntextures;
    });
}

int pt_set_settings(pt_context* ctx, const pt_settings* settings) {
    if (!ctx || !settings) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] { ctx->renderer.settings = *settings; });
}

int pt_build_bvh(pt_context* ctx, float* ms_out) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        if (!ctx->scene.ready) throw pt::Error(PT_E_NO_SCENE, "no scene uploaded");
        hipStream_t st = ctx->stream;
        PT_HIP(hipEventRecord(ctx->ev0, st));
        pt::build_lbvh(ctx->scene, st);
        PT_HIP(hipEventRecord(ctx->ev1, st));
        mark_ready(ctx);
        if (ms_out) {
            PT_HIP(hipEventSynchronize(ctx->ev1));
            PT_HIP(hipEventElapsedTime(ms_out, ctx->ev0, ctx->ev1));
            check_status(ctx);
        }
    });
}

int pt_get_bvh(pt_context* ctx, uint32_t* morton, uint32_t* perm, uint32_t* children,
               uint32_t* parents_leaf, uint32_t* parents_internal, float* bboxes) {
    return pt_get_bvh_scene(ctx, 0, morton, perm, children, parents_leaf, parents_internal, bboxes);
}

int pt_get_bvh_scene(pt_context* ctx, uint32_t scene, uint32_t* morton, uint32_t* perm,
                     uint32_t* children, uint32_t* parents_leaf, uint32_t* parents_internal,
                     float* bboxes) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        if (!ctx->scene.ready) throw pt::Error(PT_E_NO_SCENE, "no scene uploaded");
        pt::dump_bvh(ctx->scene, scene, ctx->stream, morton, perm, children, parents_leaf,
                     parents_internal, bboxes);
    });
}

int pt_get_scene_info(pt_context* ctx, uint32_t* ntris, float* bmin, float* bmax) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        const pt::DeviceScene& s = ctx->scene;
        if (ntris) *ntris = s.n;
        for (int k = 0; k < 3; ++k) {
            const bool have = s.box_host.size() >= 6;
            if (bmin) bmin[k] = have ? s.box_host[k] : 0.0f;
            if (bmax) bmax[k] = have ? s.box_host[3 + k] : 0.0f;
        }
    });
}

int pt_verify_bvh(pt_context* ctx, uint32_t* bad_scenes) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        const pt::DeviceScene& s = ctx->scene;
        if (!s.ready) throw pt::Error(PT_E_NO_SCENE, "no scene uploaded");
        const uint32_t ns = s.nscenes();
        std::vector<float> root(static_cast<size_t>(ns) * 6, 0.0f);
        if (s.n > 0) {
            PT_HIP(hipMemcpyAsync(root.data(), s.bvh.root_box.as<void>(), root.size() * sizeof(float),
                                  hipMemcpyDeviceToHost, ctx->stream));
        }
        PT_HIP(hipStreamSynchronize(ctx->stream));
        check_status(ctx);
        uint32_t bad = 0, first = 0;
        for (uint32_t k = 0; k < ns; ++k) {
            if (s.off_host[k + 1] == s.off_host[k]) continue;  // an empty scene has no root
            if (std::memcmp(&root[6 * static_cast<size_t>(k)], &s.box_host[6 * static_cast<size_t>(k)],
                            6 * sizeof(float)) != 0) {
                if (bad++ == 0) first = k;
            }
        }
        if (bad_scenes) *bad_scenes = bad;
        if (bad) {
            const float* r = &root[6 * static_cast<size_t>(first)];
            const float* e = &s.box_host[6 * static_cast<size_t>(first)];
            char msg[320];
            std::snprintf(msg, sizeof msg,
                          "%u scene(s) whose root box differs from the enclosing box; first: scene %u, root "
                          "(%g %g %g)-(%g %g %g), enclosing (%g %g %g)-(%g %g %g)",
                          bad, first, r[0], r[1], r[2], r[3], r[4], r[5], e[0], e[1], e[2], e[3], e[4], e[5]);
            throw pt::Error(PT_E_INTERNAL, msg);
        }
    });
}

int pt_trace_closest_scenes_device(pt_context* ctx, const float* d_rays, const uint32_t* d_scene,
                                   uint32_t n, uint32_t* d_id, float* d_t, float* d_uv, void* stream) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        if (!ctx->scene.ready) throw pt::Error(PT_E_NO_SCENE, "no scene uploaded");
        require(n == 0 || (d_rays != nullptr && d_scene != nullptr), "rays or scene ids NULL");
        require(n < (1u << 31), "at most 2^31 - 1 rays per call");
        hipStream_t st = pick(ctx, stream);
        order_after_build(ctx, st);
        pt::trace_closest_scenes(ctx->scene, trace_queue(ctx, st), d_rays, d_scene, n, d_id, d_t, d_uv, st);
    });
}

int pt_trace_closest_device(pt_context* ctx, const float* d_rays, uint32_t n, uint32_t* d_id,
                            float* d_t, float* d_uv, void* stream) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        if (!ctx->scene.ready) throw pt::Error(PT_E_NO_SCENE, "no scene uploaded");
        require(n == 0 || d_rays != nullptr, "rays is NULL");
        require(n < (1u << 31), "at most 2^31 - 1 rays per call");
        hipStream_t st = pick(ctx, stream);
        order_after_build(ctx, st);
        pt::trace_closest(ctx->scene, trace_queue(ctx, st), d_rays, n, d_id, d_t, d_uv, st);
    });
}

int pt_trace_closest_device_noalloc(pt_context* ctx, const float* d_rays, uint32_t n, uint32_t* d_id,
                                    float* d_t, float* d_uv, void* stream) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        if (!ctx->scene.ready) throw pt::Error(PT_E_NO_SCENE, "no scene uploaded");
        require(n == 0 || d_rays != nullptr, "rays is NULL");
        require(n <= PT_NOALLOC_MAX_RAYS, "at most 2^31 - 1 rays per call (32-bit hand-out counters)");
        uint32_t* queue = pt::trace_scratch(ctx->device);
        if (!queue) throw pt::Error(PT_E_INTERNAL, "the device's scratch block was not set up");
        hipStream_t st = pick(ctx, stream);
        order_after_build(ctx, st);
        pt::trace_closest(ctx->scene, queue, d_rays, n, d_id, d_t, d_uv, st);
    });
}

int pt_trace_closest(pt_context* ctx, const float* rays, uint32_t n, uint32_t* id, float* t,
                     float* uv) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        if (!ctx->scene.ready) throw pt::Error(PT_E_NO_SCENE, "no scene uploaded");
        require(n == 0 || rays != nullptr, "rays is NULL");
        require(n < (1u << 31), "at most 2^31 - 1 rays per call");
        if (n == 0) return;
        hipStream_t st = ctx->stream;
        const size_t nn = n;
        ctx->io_rays.reserve(nn * 24);
        ctx->io_id.reserve(nn * 4);
        ctx->io_t.reserve(nn * 4);
        ctx->io_uv.reserve(nn * 8);
        PT_HIP(hipMemcpyAsync(ctx->io_rays.as<void>(), rays, nn * 24, hipMemcpyHostToDevice, st));
        pt::trace_closest(ctx->scene, trace_queue(ctx, st), ctx->io_rays.as<float>(), n,
                          id ? ctx->io_id.as<uint32_t>() : nullptr, t ? ctx->io_t.as<float>() : nullptr,
                          uv ? ctx->io_uv.as<float>() : nullptr, st);
        if (id) PT_HIP(hipMemcpyAsync(id, ctx->io_id.as<void>(), nn * 4, hipMemcpyDeviceToHost, st));
        if (t) PT_HIP(hipMemcpyAsync(t, ctx->io_t.as<void>(), nn * 4, hipMemcpyDeviceToHost, st));
        if (uv) PT_HIP(hipMemcpyAsync(uv, ctx->io_uv.as<void>(), nn * 8, hipMemcpyDeviceToHost, st));
        PT_HIP(hipStreamSynchronize(st));
        check_status(ctx);
    });
}

int pt_trace_closest_counted(pt_context* ctx, const float* d_rays, uint32_t n, uint32_t* d_id,
                             float* d_t, float* d_uv, uint64_t* counters, void* stream) {
    if (!ctx || !counters) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        if (!ctx->scene.ready) throw pt::Error(PT_E_NO_SCENE, "no scene uploaded");
        require(n == 0 || d_rays != nullptr, "rays is NULL");
        require(n < (1u << 31), "at most 2^31 - 1 rays per call");
        hipStream_t st = pick(ctx, stream);
        order_after_build(ctx, st);
        pt::DevBuf& c = ctx->count_buf;
        c.reserve(4 * sizeof(unsigned long long));
        PT_HIP(hipMemsetAsync(c.as<void>(), 0, 4 * sizeof(unsigned long long), st));
        pt::trace_closest_counted(ctx->scene, trace_queue(ctx, st), d_rays, n, d_id, d_t, d_uv,
                                  c.as<unsigned long long>(), st);
        PT_HIP(hipMemcpyAsync(counters, c.as<void>(), 4 * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, st));
        PT_HIP(hipStreamSynchronize(st));
        check_status(ctx);
    });
}

namespace {
// Writes the running-mean image into the caller's screen buffer, when one is
// set, after every frame (the reference writes its surface in raytrace_kernel,
// cuda_path_tracer.cu:411-420).
void present(pt_context* ctx) {
    const pt::Renderer& r = ctx->renderer;
    if (r.screen) pt::render_display(r, r.screen, r.screen_w, r.screen_h, r.screen_pitch, ctx->stream);
}
}  // namespace

namespace {
// CUDAPathTracer::raytrace's stop at a frame count (cuda_path_tracer.cu:428-441):
// how many of `want` passes may still run; at the limit, records the time
// since the first pass after the last reset.
uint32_t frames_allowed(pt_context* ctx, uint32_t want) {
    pt::Renderer& r = ctx->renderer;
    if (r.spp == 0) r.t_start = std::chrono::steady_clock::now();
    if (r.stop_at == 0) return want;
    if (r.spp >= r.stop_at) {
        if (r.gen_seconds < 0.0) {
            PT_HIP(hipStreamSynchronize(ctx->stream));
            r.gen_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - r.t_start).count();
        }
        return 0;
    }
    return std::min(want, r.stop_at - r.spp);
}
}  // namespace

int pt_set_stop_frame(pt_context* ctx, uint32_t frames) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] { ctx->renderer.stop_at = frames; });
}

int pt_generation_time(pt_context* ctx, double* seconds) {
    if (!ctx || !seconds) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] { *seconds = ctx->renderer.gen_seconds; });
}

int pt_set_screen_buffer(pt_context* ctx, void* d_rgba, uint32_t w, uint32_t h, size_t row_pitch) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        pt::Renderer& r = ctx->renderer;
        if (!d_rgba) {
            r.screen = nullptr;
            r.screen_w = r.screen_h = 0;
            r.screen_pitch = 0;
            return;
        }
        require(w > 0 && h > 0, "screen size must be positive");
        require(row_pitch >= static_cast<size_t>(w) * 16 && row_pitch % 16 == 0,
                "row pitch must hold w float4 texels and be a multiple of 16 bytes");
        require(reinterpret_cast<uintptr_t>(d_rgba) % 16 == 0, "screen buffer must be 16-byte aligned");
        r.screen = d_rgba;
        r.screen_w = w;
        r.screen_h = h;
        r.screen_pitch = row_pitch;
    });
}

int pt_read_image_device(pt_context* ctx, void* d_rgba, uint32_t w, uint32_t h, size_t row_pitch,
                         void* stream) {
    if (!ctx || !d_rgba) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        require(row_pitch >= static_cast<size_t>(w) * 16 && row_pitch % 16 == 0,
                "row pitch must hold w float4 texels and be a multiple of 16 bytes");
        require(reinterpret_cast<uintptr_t>(d_rgba) % 16 == 0, "image buffer must be 16-byte aligned");
        pt::render_display(ctx->renderer, d_rgba, w, h, row_pitch, pick(ctx, stream));
    });
}

int pt_set_primary_rays(pt_context* ctx, const float* dirs, const float* eye) {
    if (!ctx || !dirs || !eye) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        std::memcpy(ctx->renderer.dirs, dirs, sizeof ctx->renderer.dirs);
        std::memcpy(ctx->renderer.eye, eye, sizeof ctx->renderer.eye);
    });
}

int pt_render_pass(pt_context* ctx, int w, int h, uint64_t pass_seed) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        if (!ctx->scene.ready) throw pt::Error(PT_E_NO_SCENE, "no scene uploaded");
        hipStream_t st = ctx->stream;
        pt::render_resize(ctx->renderer, w, h, st);
        if (frames_allowed(ctx, 1) == 0) return;
        PT_HIP(hipEventRecord(ctx->ev0, st));
        pt::render_passes(ctx->renderer, ctx->scene, 1, nullptr, &pass_seed, nullptr, st);
        PT_HIP(hipEventRecord(ctx->ev1, st));
        ctx->renderer.timed = true;
        present(ctx);
    });
}

int pt_render_passes(pt_context* ctx, int w, int h, uint32_t npasses, const float* dirs,
                     const uint64_t* seeds, const uint32_t* scenes) {
    if (!ctx || !seeds) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        if (!ctx->scene.ready) throw pt::Error(PT_E_NO_SCENE, "no scene uploaded");
        hipStream_t st = ctx->stream;
        pt::render_resize(ctx->renderer, w, h, st);
        const uint32_t run = npasses ? frames_allowed(ctx, npasses) : npasses;
        if (npasses && run == 0) return;
        PT_HIP(hipEventRecord(ctx->ev0, st));
        pt::render_passes(ctx->renderer, ctx->scene, run, dirs, seeds, scenes, st);
        PT_HIP(hipEventRecord(ctx->ev1, st));
        ctx->renderer.timed = true;
        present(ctx);
    });
}

int pt_last_frame_rays(pt_context* ctx, uint64_t* rays) {
    if (!ctx || !rays) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] { *rays = pt::render_last_rays(ctx->renderer, ctx->stream); });
}

int pt_reset_frames(pt_context* ctx) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] { pt::render_reset(ctx->renderer, ctx->stream); });
}

int pt_read_accumulator(pt_context* ctx, double* rgb_sum, uint32_t* spp) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        pt::Renderer& r = ctx->renderer;
        if (rgb_sum && r.w > 0) {
            PT_HIP(hipMemcpyAsync(rgb_sum, r.accum.as<void>(),
                                  static_cast<size_t>(r.w) * r.h * 3 * sizeof(double),
                                  hipMemcpyDeviceToHost, ctx->stream));
        }
        PT_HIP(hipStreamSynchronize(ctx->stream));
        check_status(ctx);
        if (spp) *spp = r.spp;
    });
}

int pt_write_accumulator(pt_context* ctx, int w, int h, const double* rgb_sum, uint32_t spp) {
    if (!ctx || !rgb_sum) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        pt::Renderer& r = ctx->renderer;
        pt::render_resize(r, w, h, ctx->stream);
        PT_HIP(hipMemcpyAsync(r.accum.as<void>(), rgb_sum,
                              static_cast<size_t>(w) * h * 3 * sizeof(double),
                              hipMemcpyHostToDevice, ctx->stream));
        PT_HIP(hipStreamSynchronize(ctx->stream));
        r.spp = spp;
    });
}

int pt_read_image(pt_context* ctx, float* rgba) {
    if (!ctx || !rgba) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        pt::Renderer& r = ctx->renderer;
        const size_t np = static_cast<size_t>(r.w) * r.h;
        std::vector<double> sum(np * 3);
        if (np) {
            PT_HIP(hipMemcpyAsync(sum.data(), r.accum.as<void>(), np * 3 * sizeof(double),
                                  hipMemcpyDeviceToHost, ctx->stream));
        }
        PT_HIP(hipStreamSynchronize(ctx->stream));
        const double inv = r.spp ? 1.0 / static_cast<double>(r.spp) : 0.0;
        for (size_t p = 0; p < np; ++p) {
            for (int c = 0; c < 3; ++c) rgba[4 * p + c] = static_cast<float>(sum[3 * p + c] * inv);
            rgba[4 * p + 3] = 1.0f;
        }
    });
}

int pt_last_render_ms(pt_context* ctx, float* ms) {
    if (!ctx || !ms) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        if (!ctx->renderer.timed) throw pt::Error(PT_E_INVALID_ARG, "no pass rendered");
        PT_HIP(hipEventSynchronize(ctx->ev1));
        PT_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    });
}

namespace {
uint32_t* tune_slot(const char* name) {
    if (!name) return nullptr;
    const std::string k(name);
    if (k == "build.staged") return &pt::build_tuning().staged;
    if (k == "trace.blocks") return &pt::trace_tuning().blocks;
    if (k == "trace.refill") return &pt::trace_tuning().refill;
    if (k == "trace.lds_stack") return &pt::trace_tuning().lds;
    if (k == "trace.pool") return &pt::trace_tuning().pool;
    if (k == "trace.filter") return &pt::trace_tuning().filter;
    if (k == "render.graphs") return &pt::render_tuning().graphs;
    if (k == "render.chains") return &pt::render_tuning().chains;
    if (k == "render.paths") return &pt::render_tuning().paths;
    if (k == "render.shade_min") return &pt::render_tuning().shade_min;
    return nullptr;
}

bool tune_valid(const uint32_t* slot, uint32_t v) {
    if (slot == &pt::render_tuning().graphs) return true;
    if (slot == &pt::trace_tuning().filter) return v <= 1;
    if (slot == &pt::build_tuning().staged) return v <= 1;
    if (slot == &pt::render_tuning().chains) return v >= 1 && v <= 8;
    if (slot == &pt::render_tuning().paths) return v <= 1;
    if (slot == &pt::render_tuning().shade_min) return v >= 1 && v <= 64;
    if (slot == &pt::trace_tuning().lds) return v == 8 || v == 16 || v == 32;
    if (slot == &pt::trace_tuning().pool) return v >= 64 && v <= 4096;
    return v != 0;
}
}  // namespace

int pt_device_compute_units(int device, uint32_t* cus) {
    if (!cus) return PT_E_INVALID_ARG;
    return guarded(nullptr, [&] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
            throw pt::Error(PT_E_NO_DEVICE, "no such HIP device");
        int v = 0;
        PT_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device));
        *cus = static_cast<uint32_t>(v);
    });
}

int pt_tune_set(const char* name, uint32_t value) {
    uint32_t* slot = tune_slot(name);
    if (!slot || !tune_valid(slot, value)) return PT_E_INVALID_ARG;
    *slot = value;
    return PT_OK;
}

int pt_tune_get(const char* name, uint32_t* value) {
    uint32_t* slot = tune_slot(name);
    if (!slot || !value) return PT_E_INVALID_ARG;
    *value = *slot;
    return PT_OK;
}

int pt_cache_param_count(void) { return pt::RadianceCache::param_count(); }

int pt_cache_init(pt_context* ctx, uint64_t seed, const float* bmin, const float* bmax) {
    if (!ctx || !bmin || !bmax) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        ctx->cache.set_box(bmin, bmax);
        ctx->cache.init(seed, ctx->stream);
    });
}

int pt_cache_set_params(pt_context* ctx, const float* params) {
    if (!ctx || !params) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] { ctx->cache.set_params(params, ctx->stream); });
}

int pt_cache_get_params(pt_context* ctx, float* params) {
    if (!ctx || !params) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] { ctx->cache.get_params(params, ctx->stream); });
}

int pt_cache_set_adam(pt_context* ctx, float lr, float beta1, float beta2, float eps) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        ctx->cache.lr = lr;
        ctx->cache.beta1 = beta1;
        ctx->cache.beta2 = beta2;
        ctx->cache.eps = eps;
    });
}

int pt_cache_infer_device(pt_context* ctx, const float* d_raw, uint32_t n, float* d_out, void* stream) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        require(n == 0 || (d_raw && d_out), "NULL buffer");
        ctx->cache.infer(d_raw, n, d_out, pick(ctx, stream));
    });
}

int pt_cache_train_device(pt_context* ctx, const float* d_raw, const float* d_target, uint32_t n,
                          float* d_loss, void* stream) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        require(n == 0 || (d_raw && d_target), "NULL buffer");
        ctx->cache.train(d_raw, d_target, n, d_loss, nullptr, pick(ctx, stream));
    });
}

int pt_cache_gradient_device(pt_context* ctx, const float* d_raw, const float* d_target, uint32_t n,
                             float* d_grad, void* stream) {
    if (!ctx || !d_grad) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        require(n > 0 && d_raw && d_target, "empty batch");
        ctx->cache.train(d_raw, d_target, n, nullptr, d_grad, pick(ctx, stream));
    });
}

int pt_render_cache_frame(pt_context* ctx, int w, int h, uint64_t pass_seed,
                          const pt_cache_settings* settings) {
    if (!ctx || !settings) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] {
        if (!ctx->scene.ready) throw pt::Error(PT_E_NO_SCENE, "no scene uploaded");
        hipStream_t st = ctx->stream;
        pt::render_resize(ctx->renderer, w, h, st);
        if (frames_allowed(ctx, 1) == 0) return;
        PT_HIP(hipEventRecord(ctx->ev0, st));
        pt::render_cache_frame(ctx->renderer, ctx->scene, ctx->cache, pass_seed, *settings, st);
        PT_HIP(hipEventRecord(ctx->ev1, st));
        ctx->renderer.timed = true;
        present(ctx);
    });
}

int pt_cache_last_batch(pt_context* ctx, float* raw, float* target, uint32_t n) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] { pt::render_last_batch(ctx->renderer, raw, target, n, ctx->stream); });
}

int pt_cache_last_queries(pt_context* ctx, float* raw, float* pred, uint32_t* valid) {
    if (!ctx) return PT_E_INVALID_ARG;
    return guarded(ctx, [&] { pt::render_last_queries(ctx->renderer, raw, pred, valid, ctx->stream); });
}

}  // extern "C"
