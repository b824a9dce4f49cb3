/*
 * slimtracin_seam.h - the reference's host/device seam, one entry point per
 * function of it, over plain records.
 *
 * SlimTracin's CUDA build mirrors the scene to the device through seven
 * functions of core/init.h (allocateDeviceScene, uploadScene,
 * uploadPrimitives, uploadMaterials, uploadLights, uploadSceneBVH,
 * uploadMeshBVHs) and renders with renderSceneOnGPU (render/raytracer.h:258).
 * Each st_seam_* function below replaces one of them with the same meaning:
 * the caller owns the scene on the host, hands over arrays whenever it
 * changes them, and asks for a frame with the viewport's projection plane and
 * camera. A shim in the reference's core/init.h converts its structs into the
 * records of st_blob.h and calls these. Unlike the reference, every function
 * returns a status (ST_OK or a negative ST_ERR_*; st_last_error() explains),
 * and the frame goes to caller memory or stays on the device.
 *
 * Differences kept on purpose (the reference CPU renderer is the contract):
 * texture mips hold (width+1)*(height+1) texel quads as the .texture format
 * stores them; primary directions are the CPU renderer's sequential sums;
 * traversal stacks are sized from the uploaded BVH heights.
 */
#ifndef SLIMTRACIN_SEAM_H
#define SLIMTRACIN_SEAM_H

#include <stddef.h>
#include <stdint.h>

#include "slimtracin.h"
#include "st_blob.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Capacities fixed at allocation, as allocateDeviceScene reads them from
   scene->settings (core/init.h:36-72). */
typedef struct StSeamSizes {
    uint32_t primitives, materials, lights, meshes, textures;
} StSeamSizes;

/* One mesh as uploadMeshBVHs hands it over (core/init.h:111-126): the Mesh
   fields the renderer reads (aabb, triangle/normal/uv counts, BVH height),
   its BVH nodes, its Triangle records in leaf order and its leaf ids. */
typedef struct StSeamMesh {
    float aabb[6];                 /* mesh->aabb min xyz, max xyz */
    uint32_t triangle_count, normals_count, uvs_count;
    uint32_t node_count, height;   /* mesh->bvh.node_count, mesh->bvh.height */
    const StBVHNodeRec *nodes;
    const StTriangleRec *triangles;
    const uint32_t *leaf_ids;      /* may be NULL: the renderer does not read them */
} StSeamMesh;

/* One texture mip chain as uploadScene hands it over (core/init.h:89-103). */
typedef struct StSeamMip {
    uint16_t width, height;
    const uint8_t *texel_quads;    /* (width+1)*(height+1) quads of 12 B (TexelQuad) */
} StSeamMip;

typedef struct StSeamTexture {
    uint16_t width, height;
    uint8_t mip_count, wrap, mipmap, pad;
    const StSeamMip *mips;
} StSeamTexture;

/* What renderSceneOnGPU reads from the Scene and the Viewport
   (render/raytracer.h:258-299). */
typedef struct StSeamFrame {
    float start[3], right[3], down[3];   /* viewport->projection_plane */
    float cone_angle;                    /* viewport->projection_plane.cone_angle */
    float camera_position[3];            /* camera->transform.position */
    float camera_rotation_inverted[4];   /* camera->transform.rotation_inverted: x, y, z, w */
    float ambient[3];                    /* scene->ambient_light.color */
    uint16_t width, height;              /* frame_buffer->dimensions: the traced grid (QCAA: window + 1) */
    uint32_t render_mode;                /* viewport->settings.render_mode (ST_MODE_*) */
    uint32_t depth;                      /* viewport->trace.depth */
} StSeamFrame;

/* allocateDeviceScene (core/init.h:36). The handle starts empty; it works
   with every other entry of slimtracin.h (st_scene_info, st_render, ...)
   once something has been uploaded and a frame rendered or
   st_seam_sync() called. */
int st_seam_allocate(const StSeamSizes *sizes, StSceneHandle **out);

/* uploadPrimitives (core/init.h:74): all `primitives` records, with each
   primitive's screen_bounds (min.x, min.y, max.x, max.y, from updateSceneSSB;
   render/SSB.h:137-173) and its flags' IS_VISIBLE bit gating primary rays. */
int st_seam_upload_primitives(StSceneHandle *scene, const StPrimRec *prims, const int32_t *screen_bounds);

/* uploadMaterials (core/init.h:77). */
int st_seam_upload_materials(StSceneHandle *scene, const StMaterialRec *materials);

/* uploadLights (core/init.h:80). Area lights are not read by the renderer
   (emissive quads are primitives). */
int st_seam_upload_lights(StSceneHandle *scene, const StLightRec *lights);

/* uploadScene (core/init.h:85): lights, primitives, materials and textures.
   textures may be NULL when the scene has none. */
int st_seam_upload_scene(StSceneHandle *scene, const StLightRec *lights, const StPrimRec *prims,
                         const int32_t *screen_bounds, const StMaterialRec *materials,
                         const StSeamTexture *textures);

/* uploadSceneBVH (core/init.h:106): scene->bvh.nodes[node_count] and the
   `primitives` leaf ids. */
int st_seam_upload_scene_bvh(StSceneHandle *scene, const StBVHNodeRec *nodes, uint32_t node_count,
                             uint32_t height, const uint32_t *leaf_ids);

/* uploadMeshBVHs (core/init.h:111): all `meshes` meshes. */
int st_seam_upload_mesh_bvhs(StSceneHandle *scene, const StSeamMesh *meshes);

/* Make the device copy current with everything uploaded so far (done by
   st_seam_render itself). */
int st_seam_sync(StSceneHandle *scene);

/* renderSceneOnGPU (render/raytracer.h:258-299): one frame of
   frame->width x frame->height FloatPixels into `pixels` (blocking, like the
   reference's download). hits and counts may be NULL. */
int st_seam_render(StSceneHandle *scene, const StSeamFrame *frame, StFloatPixel *pixels, StHitRec *hits,
                   StRayCounts *counts);

#ifdef __cplusplus
}
#endif
#endif
