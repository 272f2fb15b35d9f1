/*
 * pt_scenes.h -- seeded synthetic workloads for libptamd (host-only C code).
 *
 * These replace the reference application's scene extraction
 * (PathTracingApp::extractScene, PathTracingApp.cpp:31-124) and camera corner
 * rays (PathTracingApp::screenToRay, PathTracingApp.cpp:217-242) with
 * deterministic generators, so the CPU oracle and the GPU path receive
 * byte-identical fp32 inputs.  No function here touches the GPU.
 * All return the number of triangles/rays written, or a negative PT_E* code.
 */
#ifndef PT_SCENES_H
#define PT_SCENES_H

#include "pt_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Enclosing box = union of the triangle boxes, with the reference's glm
 * min/max semantics (BBox::unite, BBox.cpp:20-36; PathTracingApp.cpp:113). */
int pt_scene_bounds(const pt_triangle* tris, uint32_t n, float* bmin, float* bmax);

/* Triangle soup: n triangles, centroid ~ U[0,1]^3, vertex = centroid +
 * U[-0.01,0.01]^3, grey albedo 0.7; with_emitter appends a 0.4 x 0.4 quad at
 * y = 1.2 facing -y with emission (10,10,10).  Output capacity n + 2.
 * Ids are first_id + position. */
int pt_gen_soup(uint64_t seed, uint32_t n, int with_emitter, uint32_t first_id,
                pt_triangle* tris, pt_triangle_info* infos);

/* UV sphere of `bands` latitude x `bands` longitude cells (2 triangles per
 * cell; pole cells contain one zero-area triangle each), radius 1 displaced
 * radially by 0.15 x 6-octave fBm value noise, albedo U[0.3,0.9] per channel.
 * with_room appends a 4x4x4 box room (12 inward triangles, albedo 0.7) and a
 * 1 x 1 ceiling emitter (2 triangles, emission (10,10,10)).
 * Output capacity 2*bands*bands + 14. */
int pt_gen_sphere(uint64_t seed, uint32_t bands, int with_room,
                  pt_triangle* tris, pt_triangle_info* infos);

/* Hair ball: `strands` random-walk polylines of `steps` segments of length
 * 0.005, direction jitter N(0, 0.3 rad) per step, roots uniform on the unit
 * sphere; each segment is a 2-triangle ribbon of width 0.002.
 * with_light appends a 2 x 2 emitter quad at y = 2.5 facing -y.
 * Output capacity 2*strands*steps + 2. */
int pt_gen_hairball(uint64_t seed, uint32_t strands, uint32_t steps, int with_light,
                    pt_triangle* tris, pt_triangle_info* infos);

/* Closest-hit microbench rays: origins uniform on the surface of the box
 * [bmin,bmax] scaled by `enlarge` about its centre, directions uniform on the
 * unit sphere.  rays[6*i] = origin.xyz, dir.xyz. */
int pt_gen_shell_rays(uint64_t seed, const float* bmin, const float* bmax, float enlarge,
                      uint32_t n, float* rays);

/* Log-uniform scene sizes in [lo, hi] for the batched-build workload. */
int pt_gen_batch_counts(uint64_t seed, uint32_t nscenes, uint32_t lo, uint32_t hi,
                        uint32_t* counts);

/* The four corner directions PathTracingApp::screenToRay returns
 * (PathTracingApp.cpp:217-242: corners (0,0), (w,0), (w,h), (0,h) at the near
 * plane -- bottom-left, bottom-right, top-right, top-left), normalised, for a
 * camera whose rotation has columns right, up, look (CameraComponent::
 * updateViewMatrix, CameraComponent.cpp:97-135).  NDC x runs along right,
 * NDC y along up; with left_handed = 1 (glm's perspectiveLH, the projection
 * the app's scene presets, its left-handed asset import and Transform::lookAt
 * are set up for, GLM_FORCE_LEFT_HANDED at CMakeLists.txt:109) the rays leave
 * along +look, with 0 (glm 0.9.8's default perspectiveRH, what the shipped
 * build compiles to) along -look.  fovy is the value handed to
 * glm::perspective, in radians (the app's setPerspective(45.0f),
 * PathTracingApp.cpp:441, is 45 radians: tan(22.5) = 0.557, a 58.3 degree
 * vertical field of view).  jx, jy is the NDC jitter translation; the corners
 * move by (-jx, -jy).  Returns the number of corners written (4). */
int pt_camera_corners_basis(const float* right, const float* up, const float* look, float fovy,
                            float aspect, float jx, float jy, int left_handed, float* dirs);

/* Pinhole camera at `eye` viewing `target`: pt_camera_corners_basis of the
 * camera that Transform::lookAt(target, up) orients (Transform.cpp:78-87:
 * look = normalize(target - eye), right = normalize(up x look),
 * up' = look x right), left-handed projection, vertical field of view
 * fovy_deg in degrees.  up must not be parallel to target - eye.  Returns 4. */
int pt_camera_corners(const float* eye, const float* target, const float* up,
                      float fovy_deg, float aspect, float jx, float jy, float* dirs);

/* Per-pass camera jitter in NDC, uniform in [-2/w, 2/w] x [-2/h, 2/h]
 * (PathTracingApp.cpp:225-232), derived from (seed, pass). */
int pt_pass_jitter(uint64_t seed, uint32_t pass, int w, int h, float* jx, float* jy);

#ifdef __cplusplus
}
#endif

#endif /* PT_SCENES_H */
