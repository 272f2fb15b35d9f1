/*
 * pt_assets.h -- Wavefront OBJ/MTL import into pt_triangle / pt_triangle_info
 * (host only, no GPU).  Reference: AssetImporter::import
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

/* The calling thread's last pt_load_obj error message. */
int pt_asset_error(char* buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* PT_ASSETS_H */
