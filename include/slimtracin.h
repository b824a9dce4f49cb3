/*
 * slimtracin.h - C ABI of the MI355X tracing library (libslimtracin.so).
 *
 * Two groups of entry points:
 *
 * 1. The reference's host/device seam (SlimTracin core/init.h:36-126 and
 *    render/raytracer.h:258-309): allocateDeviceScene, uploadScene,
 *    uploadPrimitives, uploadMaterials, uploadLights, uploadSceneBVH,
 *    uploadMeshBVHs and renderSceneOnGPU take the reference's own `Scene` and
 *    `Viewport` structs, so a reference application links this library where
 *    it used to compile its CUDA device mirror. They keep the reference's
 *    `void` signatures; st_last_status() reports what the last one did.
 *    They are declared in slimtracin_seam.h, which needs the reference's
 *    core/types.h.
 *
 * 2. The resident-scene API used by batching and serving: scenes are created
 *    once from a scene blob (st_blob.h) and stay in HBM; render requests
 *    (StView) name a scene. Every function returns an ST_* status;
 *    st_last_error() describes the last failure of the calling thread.
 *
 * All pointers are host pointers. Pixel buffers hold (width+1)*(height+1)
 * StFloatPixel records (the reference's QCAA grid, core/base.h:344-358).
 */
#ifndef SLIMTRACIN_H
#define SLIMTRACIN_H

#include <stddef.h>
#include <stdint.h>

#include "st_blob.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ST_OK = 0,
    ST_ERR_ARG = -1,      /* invalid argument or view */
    ST_ERR_SCENE = -2,    /* malformed scene blob or unsupported scene */
    ST_ERR_MEMORY = -3,   /* device or host allocation failed / over budget */
    ST_ERR_HIP = -4,      /* a HIP runtime call failed */
    ST_ERR_BUSY = -5,     /* serving queue full */
};

#define ST_BATCH_PERSISTENT 1u  /* one persistent launch over a global tile queue */

typedef struct StSceneHandle StSceneHandle;

typedef struct StSceneInfo {
    uint32_t prim_count, mesh_count, triangle_count, texture_count;
    uint32_t scene_node_count, scene_height, mesh_node_count, max_mesh_height;
    uint64_t device_bytes;
} StSceneInfo;

/* Select the device (one process per GPU). Called implicitly with device 0. */
int st_init(int device);
const char *st_last_error(void);

/* Parse a scene blob, build its BVHs (the reference's SAH builder unless the
   blob carries prebuilt ones) and copy it to HBM. */
int st_scene_create(const void *blob, size_t size, StSceneHandle **out);
/* Re-copy the host-side scene to the device (after st_scene_* edits). */
int st_scene_upload(StSceneHandle *scene);
void st_scene_destroy(StSceneHandle *scene);
int st_scene_info(const StSceneHandle *scene, StSceneInfo *info);
/* Scene BVH as built (reference node layout). nodes: 2*prim_count entries. */
int st_scene_bvh(const StSceneHandle *scene, StBVHNodeRec *nodes, uint32_t *leaf_ids, uint32_t *node_count,
                 uint32_t *height);

/* One request, one launch, blocking download: the plain per-request path,
   like renderSceneOnGPU (render/raytracer.h:258-299). hits and counts may be
   NULL. */
int st_render(StSceneHandle *scene, const StView *view, StFloatPixel *pixels, StHitRec *hits, StRayCounts *counts);

/* n requests fused into one launch. pixels[i] / hits[i] may be NULL. */
int st_render_batch(StSceneHandle *const *scenes, const StView *views, uint32_t n, StFloatPixel *const *pixels,
                    StHitRec *const *hits, StRayCounts *counts, uint32_t flags);

/* As st_render_batch without downloads; returns the render kernel's
   duration measured with HIP events on the launch stream. */
int st_render_timed(StSceneHandle *const *scenes, const StView *views, uint32_t n, uint32_t flags, float *kernel_ms,
                    StRayCounts *counts);

#ifdef __cplusplus
}
#endif
#endif
