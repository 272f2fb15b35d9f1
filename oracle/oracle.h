/*
 * oracle.h -- CPU restatement of the reference's hot path.  TEST
 * INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load liboracle.so; the product (libptamd.so) never does.
 *
 * Every function follows the reference file:line named beside it.  Build
 * contract decisions where the reference is unspecified or undefined:
 *   - sort is stable (equal Morton keys keep ascending input order);
 *   - the float->int quantisation saturates (NaN -> 0), as CUDA's cvt.rzi does;
 *   - equal-distance hits resolve to the lowest original id, which needs
 *     non-strict culling (t <= best) where trace.h:119,134 use strict tests;
 *   - n == 0 builds an empty BVH (every ray misses); n == 1 has no internal
 *     node and the single triangle is tested directly.
 * Pinning of each function (which test checks it against what) is listed in
 * tests/test_oracle_*.py headers; every exported function has such a check.
 * Parity unpinned (no check can fix it, the reference's GPU results being
 * unavailable): thrust's tie order among equal keys, the fp32 values nvcc's
 * default FMA contraction gives, and cuRAND XORWOW streams (the sampling here
 * uses its own published counter-based hash instead).
 */
#ifndef PT_ORACLE_H
#define PT_ORACLE_H

#include <stdint.h>

#include "../include/pt_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* a1: Morton keys (BVHBuilder.cu:12-32, math_util.h:10-13, morton.h:84-102). */
void or_morton(const pt_triangle* tris, uint32_t n, const float* bmin, const float* bmax,
               uint32_t* keys);
uint32_t or_encode30(uint32_t x, uint32_t y, uint32_t z);
uint32_t or_quantize(float x, uint32_t levels);

/* a2: stable key sort (BVHBuilder.cu:234 with the stability contract). */
void or_sort(const uint32_t* keys, uint32_t n, uint32_t* sorted_keys, uint32_t* perm);

/* a3: Karras radix tree (BVHBuilder.cu:107-201, parents[1][0] from :224). */
void or_radix_tree(const uint32_t* sorted_keys, uint32_t n, uint32_t* children,
                   uint32_t* parents_leaf, uint32_t* parents_internal);

/* a4: bottom-up refit (BVHBuilder.cu:39-98, BBox.cpp:4-36). boxes[6*(n-1)]. */
void or_refit(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
              const uint32_t* parents_leaf, const uint32_t* parents_internal, uint32_t n,
              float* boxes);

/* a5: whole build (BVHBuilder.cu:207-269).  Returns 0. */
int or_build(const pt_triangle* tris, uint32_t n, const float* bmin, const float* bmax,
             uint32_t* sorted_keys, uint32_t* perm, uint32_t* children,
             uint32_t* parents_leaf, uint32_t* parents_internal, float* boxes);

/* a8 / a9 single tests, for known-answer tests. */
int or_ray_triangle(const float* o, const float* d, const float* p0, const float* p1,
                    const float* p2, float* u, float* v, float* t);
int or_ray_aabb(const float* o, const float* d, const float* bmin, const float* bmax,
                float* t);

/* a7: closest hit in the reference's traversal order (trace.h:73-169) with the
 * lowest-id tie rule.  Outputs may be NULL.  counters (may be NULL) receives
 * 4 uint64: totals of inner-loop iterations, box tests, triangle tests, and
 * the deepest deferred-node stack any ray reached. */
void or_trace(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
              const float* boxes, uint32_t n, const float* rays, uint32_t nrays,
              uint32_t* id, float* t, float* uv, uint64_t* counters, int threads);

/* Brute force over all triangles with the same Moller-Trumbore test and the
 * same tie rule (no BVH). */
void or_brute(const pt_triangle* tris, uint32_t n, const float* rays, uint32_t nrays,
              uint32_t* id, float* t, float* uv, int threads);

/* a11/a12/a15: one progressive 1-spp pass over w x h pixels
 * (cuda_path_tracer.cu:245-333, 385-421), added into accum[3*w*h] (FP64).
 * infos are indexed by original triangle id.  seg_counts (may be NULL)
 * receives the number of traced segments. */
void or_render_pass(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
                    const float* boxes, uint32_t n, const pt_triangle_info* infos,
                    const pt_settings* settings, int w, int h, const float* dirs,
                    const float* eye, uint32_t pass, uint64_t seed, double* accum,
                    uint64_t* seg_counts, int threads);

/* Radiance-cache MLP in FP64 (mlp_oracle.cpp): parameter count, encoding of
 * one raw sample (12 floats) into 96 inputs, forward (exp outputs), loss and
 * gradient of the relative L2 loss, and one Adam step. */
int or_mlp_param_count(void);
void or_mlp_encode(const float* raw, const float* bmin, const float* bmax, double* x);
void or_mlp_forward(const float* params, const float* raw, uint32_t n, const float* bmin,
                    const float* bmax, double* out);
double or_mlp_loss_grad(const float* params, const float* raw, const float* target, uint32_t n,
                        const float* bmin, const float* bmax, double* grad);
void or_adam_step(float* w, float* m, float* v, const double* g, uint32_t count, uint32_t t,
                  float lr, float b1, float b2, float eps);

/* The same pass with directional lights (a10: shadowRay trace.h:45-50,
 * computeL LightAlgorithms.h:12-20, primary Blinn-Phong
 * cuda_path_tracer.cu:163-214, bounce light term :314-321). */
void or_render_pass_lit(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
                        const float* boxes, uint32_t n, const pt_triangle_info* infos,
                        const pt_dir_light* lights, uint32_t nlights, const pt_settings* settings,
                        int w, int h, const float* dirs, const float* eye, uint32_t pass,
                        uint64_t seed, double* accum, uint64_t* seg_counts, int threads);
/* The same pass with surface detail (pt_triangle_detail per original id,
 * textures): extractInfo / unpackNormal, cuda_path_tracer.cu:85-161.
 * detail NULL: or_render_pass_lit. */
void or_render_pass_detail(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
                           const float* boxes, uint32_t n, const pt_triangle_info* infos,
                           const pt_triangle_detail* detail, const pt_texture* textures,
                           uint32_t ntextures, const pt_dir_light* lights, uint32_t nlights,
                           const pt_settings* settings, int w, int h, const float* dirs,
                           const float* eye, uint32_t pass, uint64_t seed, double* accum,
                           uint64_t* seg_counts, int threads);
/* or_render_pass with traversal statistics: stats[5] = traced segments,
 * node visits, box tests, triangle tests (reference-order traversal) and
 * segments that hit -- the algorithmic-byte counts of a render pass. */
void or_render_pass_stats(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
                          const float* boxes, uint32_t n, const pt_triangle_info* infos,
                          const pt_settings* settings, int w, int h, const float* dirs,
                          const float* eye, uint32_t pass, uint64_t seed, double* accum,
                          uint64_t* stats, int threads);
/* tex2D as pt_texture documents it (bilinear, wrap, level 0), rgb[3]. */
void or_tex_sample(const pt_texture* tex, float s, float t, float* rgb);
/* Shading normal and albedo of one hit (u, v barycentrics) on a triangle
 * with `info` and (if not NULL) detail[0]. */
void or_shade_detail(const pt_triangle_info* info, const pt_triangle_detail* detail,
                     const pt_texture* textures, uint32_t ntextures, float u, float v, float* normal,
                     float* albedo);
/* x^y for x >= 0, y > 0 as the specular power is computed (own function). */
float or_pow_pos(float x, float y);

/* Learned-cache frame (build addition, no reference counterpart): pixel
 * paths end at their second vertex x2 and add throughput * cache(x2) (FP64
 * MLP with `params`, box [cbmin, cbmax]) into accum; ntrain training paths
 * give train_raw[12 * ntrain] and train_target[3 * ntrain] (NaN: never
 * reached x2).  Follows or_render_pass for the shared path steps. */
void or_render_cache_pass(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
                          const float* boxes, uint32_t n, const pt_triangle_info* infos,
                          const pt_dir_light* lights, uint32_t nlights,
                          const pt_settings* settings, int w, int h, const float* dirs,
                          const float* eye, uint32_t pass, uint64_t seed, const float* params,
                          const float* cbmin, const float* cbmax, uint32_t ntrain,
                          uint32_t extra_bounces, double* accum, float* train_raw,
                          float* train_target, int threads);

/* The cache frame with surface detail (detail NULL: or_render_cache_pass). */
void or_render_cache_pass_detail(const pt_triangle* tris, const uint32_t* perm, const uint32_t* children,
                                 const float* boxes, uint32_t n, const pt_triangle_info* infos,
                                 const pt_triangle_detail* detail, const pt_texture* textures,
                                 uint32_t ntextures, const pt_dir_light* lights, uint32_t nlights,
                                 const pt_settings* settings, int w, int h, const float* dirs,
                                 const float* eye, uint32_t pass, uint64_t seed, const float* params,
                                 const float* cbmin, const float* cbmax, uint32_t ntrain,
                                 uint32_t extra_bounces, double* accum, float* train_raw,
                                 float* train_target, int threads);

/* a14: direction of PathTracingApp::screenToRay (PathTracingApp.cpp:217-242)
 * for screen point p[3] (x, y in pixels, y = 0 at the bottom; z a depth
 * between the clip planes).  rot[9] = the camera rotation's columns right,
 * up, look (rot[0..2] = right); pos = camera position; fovy as passed to
 * glm::perspective (radians); jx, jy = the NDC jitter translation;
 * left_handed: glm's perspectiveLH (1) or its default perspectiveRH (0). */
void or_screen_to_ray(const float* rot, const float* pos, float fovy, float screen_w, float screen_h,
                      float znear, float zfar, float jx, float jy, const float* p, int left_handed,
                      float* dir);

/* Sampling helpers shared by the render, exposed for tests. */
void or_pcg4d(const uint32_t* in, uint32_t* out);
void or_sincos2pi(float u, float* s, float* c);
void or_to_hemisphere(const float* v, const float* n, float* out);

#ifdef __cplusplus
}
#endif

#endif /* PT_ORACLE_H */
