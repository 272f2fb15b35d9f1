/*
 * pt_assets.h -- Wavefront OBJ/MTL import into pt_triangle / pt_triangle_info
 * and PNG image output (host only, no GPU).  Reference: AssetImporter::import
 * (source/engine/util/AssetImporter.cpp:12-25, assimp with
 * aiProcessPreset_TargetRealtime_Fast | MakeLeftHanded | FlipWindingOrder),
 * material defaults (:98-102) and PathTracingApp::extractScene
 * (source/application/CUDAPathTracer/PathTracingApp.cpp:31-124): triangles in
 * file order with ids 0.., vertex normals from the file or the face normal,
 * albedo from Kd (normalised to unit length by computeColor,
 * TriangleInfo.h:39-42), emission from Ke, specularity from Ns.
 */
#ifndef PT_ASSETS_H
#define PT_ASSETS_H

#include <stdint.h>

#include "pt_api.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pt_obj_options {
    /* 1: the reference's import flags -- z mirrored, triangle winding
     * reversed (MakeLeftHanded | FlipWindingOrder); 0: the file's frame. */
    uint32_t left_handed;
    /* 1: albedo = Kd / |Kd| as the reference's computeColor shades it;
     * 0: albedo = Kd. */
    uint32_t normalize_albedo;
} pt_obj_options;

/* Loads `path` (and the MTL files it names, relative to its folder).
 * Polygons are split as fans (0, i, i+1); faces without normals for all
 * their corners get the normalised face normal of the (converted) triangle.
 * Two calls: with tris == NULL, *count receives the number of triangles;
 * then tris and infos (pt_triangle_info indexed by id) must hold *count
 * entries.  Missing materials take Kd = 1, Ke = 0, Ns = 0 (the reference's
 * defaults).  Returns PT_OK or PT_E_INVALID_ARG (message from
 * pt_asset_error). */
int pt_load_obj(const char* path, const pt_obj_options* options, pt_triangle* tris,
                pt_triangle_info* infos, uint32_t* count);

/* Full asset import with surface detail (PathTracingApp.cpp:59-93): a handle
 * owning the triangles, infos, pt_triangle_detail records (vertex colour from
 * a `v x y z r g b` line or else the material's Kd; uv from vt, else 0;
 * a per-face tangent frame from the uv derivatives, T along +u and B along
 * +v, or normalize(edge 1) and cross(N, T) where the uvs are degenerate) and
 * the textures the materials name: map_Kd as the albedo texture
 * (aiTextureType_DIFFUSE, AssetImporter.cpp:92), map_bump / bump as the
 * normal map (aiTextureType_HEIGHT, :93).  Textures are TGA files (types 2,
 * 3, 10, 11; 8, 24 or 32 bits; the format of the reference's Sponza
 * textures), stored bottom row first as pt_texture describes; other formats
 * fail the load.  Texture paths are relative to the MTL file, '\' read as
 * '/'.  pt_asset_textures' texel pointers live as long as the handle. */
typedef struct pt_asset pt_asset;
int pt_asset_load_obj(const char* path, const pt_obj_options* options, pt_asset** out);
int pt_asset_info(const pt_asset* asset, uint32_t* ntris, uint32_t* ntextures);
/* Any output may be NULL; each non-NULL one holds ntris records. */
int pt_asset_copy(const pt_asset* asset, pt_triangle* tris, pt_triangle_info* infos,
                  pt_triangle_detail* detail);
int pt_asset_textures(const pt_asset* asset, pt_texture* textures);
int pt_asset_free(pt_asset* asset);

/* One TGA file as RGBA8, bottom row first.  Two calls: rgba NULL fills *w,
 * *h; then rgba holds 4 * w * h bytes (w, h as returned). */
int pt_load_tga(const char* path, uint32_t* w, uint32_t* h, uint8_t* rgba);

/* Image output (host only).  Replaces the reference's on-screen display of
 * its RGBA32F screen texture (PathTracingApp.cpp:282-290) for runs without a
 * window: writes an 8-bit RGB PNG of a running-mean image such as
 * pt_read_image returns (rgba[4*w*h], alpha ignored).  Values are clamped to
 * [0, 1] and rounded to 8 bits, linear as the reference's framebuffer shows
 * them, or through the sRGB transfer curve with PT_PNG_SRGB.  Row 0 of rgba
 * is the bottom row (pt_read_image's order) unless PT_PNG_TOP_DOWN.
 * Returns PT_OK or PT_E_INVALID_ARG (message from pt_asset_error). */
#define PT_PNG_SRGB 1u
#define PT_PNG_TOP_DOWN 2u
int pt_write_png(const char* path, const float* rgba, uint32_t w, uint32_t h, uint32_t flags);

/* The calling thread's last pt_load_obj / pt_asset_load_obj / pt_load_tga /
 * pt_write_png error message. */
int pt_asset_error(char* buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* PT_ASSETS_H */
