/*
 * pt_api.h -- C-ABI of the MI355X-native LBVH path tracer (libptamd.so).
 *
 * Drop-in boundary for the reference's CUDA path
 * (compix/CUDA-Path-Tracer, source/application/cuda):
 *   CUDA::BVHBuilder::build            BVHBuilder.h:15, BVHBuilder.cu:207-269
 *   CUDA::Scene::upload                Scene.h:24, Scene.cu:28-55
 *   CUDAPathTracer::{upload, setPrimaryRays, setEyePos, raytrace, resetFrames}
 *                                      cuda_path_tracer.h:18-28, cuda_path_tracer.cu:366-470
 *   CUDA::trace (closest hit, device)  trace.h:73-169
 *
 * Conventions (all entry points):
 *   - every function returns int status: PT_OK (0) or a negative PT_E* code;
 *     nothing throws across the ABI; pt_last_error() returns the message.
 *   - plain pointers and sizes only.  Functions without a `_device` suffix take
 *     HOST pointers and are host-synchronous.  `_device` variants take device
 *     pointers plus an optional hipStream_t (passed as void*, NULL = the
 *     context's own stream) and are asynchronous on that stream.
 *   - one context per host thread; a context owns one device, its scene,
 *     its BVH, its accumulator and its scratch.
 */
#ifndef PT_API_H
#define PT_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_OK 0
#define PT_E_INVALID_ARG -1
#define PT_E_HIP -2
#define PT_E_NO_DEVICE -3
#define PT_E_NO_SCENE -4
#define PT_E_UNSUPPORTED -5
#define PT_E_OOM -6
#define PT_E_INTERNAL -7

#define PT_INVALID_INDEX 0xFFFFFFFFu

/* Device status flags (pt_get_status).  Each marks a bound that the
 * library's arguments show cannot be exceeded; a kernel that finds one
 * exceeded skips the step it cannot store and raises the flag, and the next
 * synchronising entry point of the context (pt_synchronize, host-pointer
 * traces, pt_trace_closest_counted, timed pt_build_bvh, uploads,
 * pt_read_accumulator) returns PT_E_INTERNAL and clears it. */
#define PT_STATUS_STACK_OVERFLOW 1u /* deferred-node stack deeper than 64 (reference: */
                                    /* MAX_TRAVERSAL_STACK_SIZE 64, cuda_defs.h:21, unchecked) */
#define PT_STATUS_REFIT_OVERFLOW 2u /* refit level list over its capacity */

/* Triangle as the caller hands it over: three fp32 vertices plus the caller's
 * triangle id (reference: CUDA::Triangle, geometry.h:20-30 -- v0, v1, v2,
 * infoIdx; 40 bytes).  `id` indexes the shading records and is the "original
 * ID" reported by traces and used to break equal-distance ties. */
typedef struct pt_triangle {
    float v0[3];
    float v1[3];
    float v2[3];
    uint32_t id;
} pt_triangle;

/* Per-triangle shading record (reference: CUDA::TriangleInfo + Material,
 * TriangleInfo.h:10-75, reduced to what the bounce loop reads):
 * vertex normals (interpolated with the hit barycentrics and normalised,
 * TriangleInfo.h:44-47), albedo used as-is (the texture branch,
 * cuda_path_tracer.cu:127-131), emission (cuda_path_tracer.cu:194-199). */
typedef struct pt_triangle_info {
    float n0[3];
    float n1[3];
    float n2[3];
    float albedo[3];
    float emission[3];
    float specularity;
} pt_triangle_info;

/* Per-triangle surface detail (reference: the Vertex fields and Material
 * texture indices of CUDA::TriangleInfo, TriangleInfo.h:14-29, read by
 * extractInfo / unpackNormal, cuda_path_tracer.cu:85-161).  With a detail
 * record a hit's albedo is the albedo texture sampled at the interpolated uv
 * (rgb), or else normalize(interpolated vertex colour) (computeColor,
 * TriangleInfo.h:39-42); its normal is normalize(interpolated vertex normal),
 * and with a normal map n = 2 * texel - 1 gives
 * normalize(n.x * T + n.y * B + n.z * N) with T, B the normalised
 * interpolated tangent and bitangent.  Interpolation is the reference's
 * barycentricLerp, v0 * (1 - u - v) + v1 * u + v2 * v (math_util.h:19-25). */
#define PT_NO_TEXTURE 0xFFFFFFFFu
typedef struct pt_triangle_detail {
    float color[3][3];     /* vertex colours */
    float uv[3][2];        /* vertex texture coordinates */
    float tangent[3][3];
    float bitangent[3][3];
    uint32_t albedo_tex;   /* texture index, or PT_NO_TEXTURE: vertex colours */
    uint32_t normal_tex;   /* texture index, or PT_NO_TEXTURE: no normal map */
    uint32_t reserved;     /* zero */
} pt_triangle_detail;      /* 144 bytes */

/* An RGBA8 texture, texels row by row starting at t = 0 (the bottom row, as
 * the reference's GL upload stores images after its Invert_Y flip,
 * Texture2D.cpp:139-145).  Sampling (tex2D on the reference's texture
 * objects, Scene.cu:57-79: normalised coordinates, wrap, bilinear, level 0;
 * 8-bit texels read as c / 255): x = s * width - 0.5, y = t * height - 0.5,
 * weights the fractions of x and y in fp32 (the hardware's are 8-bit fixed
 * point, so the reference differs in the last bits of a filtered texel). */
typedef struct pt_texture {
    uint32_t width, height;
    const uint8_t* rgba;
} pt_texture;

/* Directional light (reference: CUDA::DirectionalLight, Lights.h:6-11);
 * direction points from the light into the scene; at most 64 per scene.
 * Primary hits get Blinn-Phong from every light, bounce hits its cosine
 * term, each behind a shadow ray (LightAlgorithms.h:12-20). */
typedef struct pt_dir_light {
    float color[3];
    float direction[3];
    float intensity;
} pt_dir_light;

/* Render settings (reference: PathTracerSettings::GI, PathTracingSettings.h:19-29,
 * read at upload in Scene.cu:43-54).  rr_start = PT_INVALID_INDEX disables
 * Russian roulette (the reference has none). */
typedef struct pt_settings {
    uint32_t bounces;            /* indirectBounceCount                       */
    uint32_t rr_start;           /* first bounce index with Russian roulette   */
    float indirect_intensity;    /* indirectIntensity                         */
    uint32_t stratified;         /* useStratifiedSampling (first bounce only)  */
} pt_settings;

typedef struct pt_context pt_context;

/* ---- context ------------------------------------------------------------ */
int pt_create(pt_context** out, int device);
int pt_destroy(pt_context* ctx);
/* Copies the last error message of ctx (or of the library when ctx is NULL). */
int pt_last_error(const pt_context* ctx, char* buf, size_t len);
/* Use the caller's hipStream_t for all further work (NULL restores the
 * context's own stream). */
int pt_set_stream(pt_context* ctx, void* stream);
int pt_synchronize(pt_context* ctx);
/* Waits for the context's stream, then returns and clears the device status
 * flags (PT_STATUS_*) raised since the last read; 0 = no bound exceeded. */
int pt_get_status(pt_context* ctx, uint32_t* flags);

/* ---- scene + LBVH --------------------------------------------------------- */
/* Reference: CUDAPathTracer::upload -> CUDA::Scene::upload -> BVHBuilder::build
 * (cuda_path_tracer.cu:466-470, Scene.cu:28-55, BVHBuilder.cu:207-269).
 * Copies triangles/infos to the device and builds the LBVH.  infos[k] belongs
 * to the triangle whose id == k (ninfos >= max id + 1).  bmin/bmax is the
 * enclosing box (HostSceneDescription::enclosingBBox, HostSceneDescription.h:14);
 * NULL computes it as the union of the triangle boxes.  settings may be NULL
 * (reference defaults: 1 bounce, no RR, intensity 3, stratified). */
int pt_upload_scene(pt_context* ctx, const pt_triangle* tris, uint32_t n,
                    const pt_triangle_info* infos, uint32_t ninfos,
                    const pt_dir_light* lights, uint32_t nlights,
                    const float* bmin, const float* bmax,
                    const pt_settings* settings);
/* Surface detail for the uploaded scene(s) (reference: CUDA::Scene::upload
 * of the texture arrays and the TriangleInfo vertex data, Scene.cu:57-79,
 * PathTracingApp.cpp:59-93): detail[k] belongs to the triangle whose id == k
 * (n >= max id + 1) and replaces that triangle's pt_triangle_info albedo and
 * normal in shading (emission and specularity stay in pt_triangle_info).
 * Texture indices must be below ntextures; every texture at most 16384 x
 * 16384.  The data is copied; n == 0 removes the detail (shading uses
 * pt_triangle_info alone again).  A new pt_upload_scene(s) also removes it. */
int pt_upload_detail(pt_context* ctx, const pt_triangle_detail* detail, uint32_t n,
                     const pt_texture* textures, uint32_t ntextures);
/* Batched scenes (no reference counterpart; the reference builds one scene
 * per BVHBuilder::build call): scene s owns triangles
 * [scene_offsets[s], scene_offsets[s+1]) and gets its own LBVH, identical to
 * the one pt_upload_scene would build from those triangles alone; all scenes
 * are built by one launch sequence (one segmented sort).  boxes: 6 floats per
 * scene (min.xyz, max.xyz) or NULL for each scene's union of triangle boxes.
 * Triangle ids index one shared infos array. */
int pt_upload_scenes(pt_context* ctx, const pt_triangle* tris, const uint32_t* scene_offsets,
                     uint32_t nscenes, const pt_triangle_info* infos, uint32_t ninfos,
                     const float* boxes, const pt_settings* settings);
int pt_set_settings(pt_context* ctx, const pt_settings* settings);

/* Rebuild the LBVH of the uploaded scene (BVHBuilder::build, BVHBuilder.cu:207).
 * Asynchronous on the context stream; ms_out (may be NULL) receives the
 * HIP-event time of the build kernels and forces a synchronisation. */
int pt_build_bvh(pt_context* ctx, float* ms_out);

/* Parity dump of the last build, in the reference's layout:
 *   morton[n]             sorted 30-bit Morton keys (BVHBuilder.cu:229-234)
 *   perm[n]               input position of the k-th sorted triangle
 *   children[n-1]         Node words (Node.h:8-30): (left<<2)|leftInternal|rightInternal<<1
 *   parents_leaf[n]       parents[0] (BVHBuilder.cu:180,192)
 *   parents_internal[n-1] parents[1], [0] = PT_INVALID_INDEX (BVHBuilder.cu:224)
 *   bboxes[6*(n-1)]       internal node boxes min.xyz,max.xyz (BVHBuilder.cu:86)
 * Any pointer may be NULL. */
int pt_get_bvh(pt_context* ctx, uint32_t* morton, uint32_t* perm, uint32_t* children,
               uint32_t* parents_leaf, uint32_t* parents_internal, float* bboxes);
/* pt_get_bvh for scene `scene` of a batch, in that scene's local indices. */
int pt_get_bvh_scene(pt_context* ctx, uint32_t scene, uint32_t* morton, uint32_t* perm,
                     uint32_t* children, uint32_t* parents_leaf, uint32_t* parents_internal,
                     float* bboxes);
/* Total triangles and the enclosing box of scene 0. */
int pt_get_scene_info(pt_context* ctx, uint32_t* ntris, float* bmin, float* bmax);
/* The build's self-check (BVHBuilder.cu:258-265 asserts that the root box's
 * volume equals the enclosing box's), made exact and returned as a status:
 * PT_OK when every non-empty scene's root box equals, bound for bound, the
 * enclosing box its Morton keys were computed in; PT_E_INTERNAL (message
 * naming the first such scene) otherwise.  bad_scenes (may be NULL) receives
 * the number of mismatched scenes.  Holds for boxes pt_upload_scene(s)
 * computed itself (bmin/bmax NULL, as the reference computes them); a
 * caller-given box larger than the geometry is reported as a mismatch.
 * Host-synchronous. */
int pt_verify_bvh(pt_context* ctx, uint32_t* bad_scenes);

/* ---- closest hit ---------------------------------------------------------- */
/* Batch form of CUDA::trace (trace.h:73-169): rays[i] = origin.xyz, dir.xyz.
 * Outputs (any may be NULL): id = original triangle id of the closest hit
 * (PT_INVALID_INDEX on miss, equal distances -> lowest id), t (FLT_MAX on
 * miss), uv = barycentrics of v1, v2 (math_util.h:19-25).
 * The `_device` forms on a stream other than the context's first wait for
 * the context's last upload or build (unless the stream is being captured
 * into a graph: then the caller orders it).  Each stream has its own ray
 * queue, allocated on its first trace, so traces on different streams may
 * overlap. */
int pt_trace_closest(pt_context* ctx, const float* rays, uint32_t n,
                     uint32_t* id, float* t, float* uv);
int pt_trace_closest_device(pt_context* ctx, const float* d_rays, uint32_t n,
                            uint32_t* d_id, float* d_t, float* d_uv, void* stream);
/* pt_trace_closest_device without any allocation (no reference
 * counterpart; for callers that cannot allocate: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The ray
 * queue block lives in a static device array of the library (16,640 bytes
 * per GPU, held until the process ends) instead of a per-stream buffer
 * allocated on first use.  n <= PT_NOALLOC_MAX_RAYS; larger batches are
 * rejected with PT_E_INVALID_ARG.  All calls on one GPU share the block:
 * calls that overlap on one GPU (different streams, or different contexts
 * on the same device) race and must be serialised by the caller. */
#define PT_NOALLOC_MAX_RAYS 0x7fffffffu
int pt_trace_closest_device_noalloc(pt_context* ctx, const float* d_rays, uint32_t n,
                                    uint32_t* d_id, float* d_t, float* d_uv, void* stream);
/* Closest hit in batched scenes: ray i is traced in scene d_scene[i]. */
int pt_trace_closest_scenes_device(pt_context* ctx, const float* d_rays, const uint32_t* d_scene,
                                   uint32_t n, uint32_t* d_id, float* d_t, float* d_uv,
                                   void* stream);
/* Measurement variant (no reference counterpart): the same traversal, also
 * counting inner-loop iterations, box tests and triangle tests into
 * counters[0..2] (host, uint64, totals) and the deepest deferred-node stack
 * any ray reached into counters[3].  Host-synchronous. */
int pt_trace_closest_counted(pt_context* ctx, const float* d_rays, uint32_t n,
                             uint32_t* d_id, float* d_t, float* d_uv,
                             uint64_t* counters, void* stream);

/* ---- progressive render --------------------------------------------------- */
/* CUDAPathTracer::setPrimaryRays + setEyePos (cuda_path_tracer.cu:366-372,
 * cuda_path_tracer.h:19).  dirs = 4 corner directions, 12 floats, in the
 * reference order bottom-left, bottom-right, top-right, top-left. */
int pt_set_primary_rays(pt_context* ctx, const float* dirs, const float* eye);
/* CUDAPathTracer::raytrace (cuda_path_tracer.cu:423-464): one 1-spp pass over
 * w x h pixels, added to the FP64 accumulator.  The accumulator is reset when
 * w/h change.  pass_seed replaces the per-pass host random seed. */
int pt_render_pass(pt_context* ctx, int w, int h, uint64_t pass_seed);
/* Several passes in one frame (no reference counterpart; the reference
 * launches one pass per frame): the accumulator ends up exactly as after
 * npasses calls of pt_render_pass with the same corner rays and seeds.
 * dirs = 12 floats per pass (NULL: the corner rays of pt_set_primary_rays for
 * every pass), seeds[npasses]; 1 <= npasses <= 1024.  Every pass starts at
 * the eye given to pt_set_primary_rays (set it before the first call).
 * scenes (may be NULL =
 * all scene 0) picks, per pass, the scene of a pt_upload_scenes batch that
 * the pass traces; all passes add into the one accumulator. */
int pt_render_passes(pt_context* ctx, int w, int h, uint32_t npasses, const float* dirs,
                     const uint64_t* seeds, const uint32_t* scenes);
/* Rays traced by the last frame (camera + bounce rays); synchronises. */
int pt_last_frame_rays(pt_context* ctx, uint64_t* rays);
/* CUDAPathTracer::resetFrames (cuda_path_tracer.h:28): zero the accumulator. */
int pt_reset_frames(pt_context* ctx);
/* Accumulator access: rgb_sum[3*w*h] FP64 sums (row 0 = bottom), spp passes. */
int pt_read_accumulator(pt_context* ctx, double* rgb_sum, uint32_t* spp);
int pt_write_accumulator(pt_context* ctx, int w, int h, const double* rgb_sum, uint32_t spp);
/* Running-mean image as the reference's RGBA32F surface holds it
 * (cuda_path_tracer.cu:413-420): rgba[4*w*h] = sum/spp, alpha 1.  Row 0 is
 * the bottom row of the image (pixel y = 0, the reference's tc.y = 0). */
int pt_read_image(pt_context* ctx, float* rgba);
/* The same image into device memory: the min(w, width) x min(h, height)
 * corner of it as float4 texels, rows row_pitch bytes apart (a multiple of
 * 16 and at least 16 w; buffer 16-byte aligned), on `stream` (NULL: the
 * context's stream).  Values equal pt_read_image's bit for bit. */
int pt_read_image_device(pt_context* ctx, void* d_rgba, uint32_t w, uint32_t h, size_t row_pitch,
                         void* stream);
/* Replaces CUDAPathTracer::setScreenTexture (cuda_path_tracer.h:18): after
 * every pt_render_pass, pt_render_passes and pt_render_cache_frame the
 * running-mean image is written into this caller-owned device RGBA32F buffer
 * (same layout and clipping as pt_read_image_device) on the context's stream,
 * outside the frame's timed region.  A graphics API maps the buffer as its
 * screen texture through its own interop.  d_rgba NULL detaches it. */
int pt_set_screen_buffer(pt_context* ctx, void* d_rgba, uint32_t w, uint32_t h, size_t row_pitch);
/* Stop at a frame count (reference: CUDAPathTracer::raytrace with
 * PathTracerSettings::DEMO.stopAtFrame, cuda_path_tracer.cu:428-441,
 * PathTracingSettings.h:15; the reference's -1 = never is 0 here, the
 * default): once `frames` passes have accumulated since the last reset or
 * upload, pt_render_pass, pt_render_passes and pt_render_cache_frame render
 * nothing (a multi-pass call renders only the passes up to the limit). */
int pt_set_stop_frame(pt_context* ctx, uint32_t frames);
/* The reference's "Picture generation time": seconds from the first pass
 * after the last reset to the render call that found the limit reached, or -1
 * while it has not been reached. */
int pt_generation_time(pt_context* ctx, double* seconds);
/* Wall time in ms of the last pt_render_pass (HIP events). */
int pt_last_render_ms(pt_context* ctx, float* ms);

/* ---- learned radiance cache (no reference counterpart) -----------------------
 * MLP: 96 inputs (12-frequency sin/cos of the position mapped into the scene
 * box [bmin, bmax] (72), normal, view direction, albedo, 15 zeros) ->
 * 4 x (64, ReLU) -> 3 outputs with exp.  Parameters (fp32 masters, flat):
 * W0[96][64] b0[64] W1[64][64] b1 W2 b2 W3 b3 W4[64][3] b4[3], W[in][out];
 * pt_cache_param_count() of them.  Raw samples are 12 floats: position.xyz,
 * normal.xyz, view.xyz, albedo.rgb.  Training minimises the relative L2 loss
 * (1/n) sum (y - t)^2 / (y^2 + 0.01) (y not differentiated in the
 * denominator) with Adam (default lr 1e-3, betas 0.9/0.99, eps 1e-8). */
int pt_cache_param_count(void);
int pt_cache_init(pt_context* ctx, uint64_t seed, const float* bmin, const float* bmax);
int pt_cache_set_params(pt_context* ctx, const float* params);
int pt_cache_get_params(pt_context* ctx, float* params);
int pt_cache_set_adam(pt_context* ctx, float lr, float beta1, float beta2, float eps);
/* d_out[3*i] = prediction for sample i. */
int pt_cache_infer_device(pt_context* ctx, const float* d_raw, uint32_t n, float* d_out,
                          void* stream);
/* One Adam step on a batch; d_loss (may be NULL) receives per-sample losses. */
int pt_cache_train_device(pt_context* ctx, const float* d_raw, const float* d_target,
                          uint32_t n, float* d_loss, void* stream);
/* Gradient of the batch loss (no update), d_grad[pt_cache_param_count()]. */
int pt_cache_gradient_device(pt_context* ctx, const float* d_raw, const float* d_target,
                             uint32_t n, float* d_grad, void* stream);

/* Cache frames: a 1-spp pass over w x h pixels (corner rays, eye and RR
 * settings as for pt_render_pass) whose paths end at their second vertex x2
 * (after the first bounce) and add throughput * cache(x2).  train_samples
 * training paths start on one pixel of every 4x4 block (the cell steps
 * through all 16 with the pass index), record x2 and extend extra_bounces
 * bounces past it; their suffix radiance is the cache's target at x2.  With
 * PT_CACHE_TRAIN the cache then takes train_steps Adam steps over the batch
 * (split evenly), so the next frame queries the updated network.  The whole
 * frame replays as one captured graph. */
#define PT_CACHE_TRAIN 1u
typedef struct pt_cache_settings {
    uint32_t train_samples; /* training paths per frame, <= w*h (65536) */
    uint32_t train_steps;   /* Adam steps per frame, 1..64 (4) */
    uint32_t extra_bounces; /* bounces of a training path past x2, <= 58 (8) */
    uint32_t flags;         /* PT_CACHE_TRAIN */
} pt_cache_settings;
int pt_render_cache_frame(pt_context* ctx, int w, int h, uint64_t pass_seed,
                          const pt_cache_settings* settings);
/* The first n (raw, target) samples of the last cache frame's training batch:
 * raw 12 floats, target 3 floats (NaN: the path never reached x2). */
int pt_cache_last_batch(pt_context* ctx, float* raw, float* target, uint32_t n);
/* The last cache frame's pixel queries (raw 12 floats per pixel), the
 * cache's predictions (3 per pixel) and whether the pixel's path reached x2
 * (valid[w*h]); any pointer may be NULL.  For inspecting the cache. */
int pt_cache_last_queries(pt_context* ctx, float* raw, float* pred, uint32_t* valid);

/* Compute units of HIP device `device` (no reference counterpart; launch
 * shapes and the autotuner's search space scale with it). */
int pt_device_compute_units(int device, uint32_t* cus);

/* ---- launch tuning (no reference counterpart) -------------------------------
 * Process-wide launch-shape knobs read by the kernels' host launchers; results
 * do not depend on them (tests check this).  Names:
 *   "build.staged"  1: the LBVH refit's upper levels (nodes whose leaves span
 *                   more than one 1024-leaf block) run with each block's
 *                   node list staged in LDS when it holds at most 1024
 *                   nodes; 0: always from global memory
 *   "trace.blocks"  persistent closest-hit workgroups (256 threads each)
 *   "trace.refill"  refill idle lanes when fewer than this many of 64 are busy
 *   "trace.lds_stack" traversal stack entries per lane kept in LDS (8, 16 or
 *                   32; deeper entries spill to scratch)
 *   "trace.pool"    rays a wave takes from the ray queue at a time on long
 *                   queues (64..4096)
 *   "trace.filter"  1: closest-hit batches run every ray's root visit in
 *                   converged 64-ray rounds and trace only the rays that go
 *                   deeper (fast when most rays miss the scene); 0: off
 *   "render.graphs" 1: replay each frame's kernels as one captured hipGraph,
 *                   0: launch them one by one
 *   "render.paths"  1: a plain frame traces every path from its primary
 *                   ray to its end in one persistent launch (no bounce
 *                   levels), and a cache frame its pixel paths from their
 *                   primary ray to the cache; 0: bounce levels (wavefront)
 *   "render.shade_min" 1..64: in path-persistent frames, the lanes of a
 *                   wave whose segments have ended shade together once this
 *                   many wait (or none is traversing)
 *   "render.chains" 1..8: with render.paths 0, a plain frame's paths split
 *                   into this many chains of bounce levels on streams of
 *                   their own (at most one per 16,384 paths), so one chain's
 *                   level tail overlaps the others' work
 * Unknown names and values outside a knob's range return PT_E_INVALID_ARG. */
int pt_tune_set(const char* name, uint32_t value);
int pt_tune_get(const char* name, uint32_t* value);

#ifdef __cplusplus
}
#endif

#endif /* PT_API_H */
