/*
 * slimtracin.h - C ABI of the MI355X tracing library (libslimtracin.so).
 *
 * Two groups of entry points:
 *
 * 1. The reference's host/device seam (SlimTracin core/init.h:36-126 and
 *    render/raytracer.h:258-299), in slimtracin_seam.h: one st_seam_*
 *    function per function of the reference's device mirror
 *    (allocateDeviceScene, uploadScene, uploadPrimitives, uploadMaterials,
 *    uploadLights, uploadSceneBVH, uploadMeshBVHs) and renderSceneOnGPU,
 *    with the same meaning, over the plain records of st_blob.h; each
 *    returns an ST_* status. The reference-side binding that keeps the
 *    reference's `void` functions over its own Scene / Viewport structs and
 *    calls these is integration/init_amd.h (compiled with the reference's
 *    headers).
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
    ST_ERR_BUSY = -5,     /* over the serving HBM budget, or a wait timed out */
};

#define ST_BATCH_PERSISTENT 1u  /* one persistent launch over a global tile queue */
#define ST_BATCH_STAGE_TIMES 2u /* st_render_timed: also time every stage launch (st_last_stage_ms) */
#define ST_BATCH_GRAPH 4u       /* capture the launch's device work (upload, every stage kernel) in a HIP graph
                                   and replay it while the next launches have the same shape (scenes' stack
                                   sizes, request count, sizes, modes, depths); downloads stay outside */

typedef struct StSceneHandle StSceneHandle;

typedef struct StSceneInfo {
    uint32_t prim_count, mesh_count, triangle_count, texture_count;
    uint32_t scene_node_count, scene_height, mesh_node_count, max_mesh_height;
    uint64_t device_bytes;
} StSceneInfo;

/* Select the device (one process per GPU). Called implicitly with device 0. */
int st_init(int device);
const char *st_last_error(void);

/* The device the library runs on (initialising it with device 0 if no
   st_init came first): name and architecture as the runtime reports them,
   compute units, HBM bytes (name "AMD GPU <arch>" when the runtime reports
   none). No counterpart in the reference (its GPU build
   takes the first CUDA device); the tuning store keys launch shapes by it. */
typedef struct StDeviceInfo {
    char name[64];
    char arch[32];        /* e.g. "gfx950:sramecc+:xnack-" */
    int32_t device;
    uint32_t compute_units;
    uint64_t hbm_bytes;
} StDeviceInfo;
int st_device_info(StDeviceInfo *info);

/* Parse a scene blob, build its BVHs (the reference's SAH builder unless the
   blob carries prebuilt ones) and copy it to HBM. */
int st_scene_create(const void *blob, size_t size, StSceneHandle **out);
#define ST_SCENE_MESH_LBVH 1u   /* build meshes without a prebuilt BVH on the device (Morton LBVH, one triangle per leaf) */
#define ST_SCENE_MESH_BINNED 2u /* build them on the host with the binned SAH builder (one triangle per leaf) */

/* st_scene_create with options. Both flags replace the reference's host SAH
   mesh build (updateMeshBVH, builder_top_down.h:272-364), which takes
   minutes for meshes of millions of triangles: ST_SCENE_MESH_LBVH with a
   device build of milliseconds, ST_SCENE_MESH_BINNED with a host build of
   seconds whose trees take fewer node steps per ray. At most one of them;
   ST_ERR_ARG for both. */
int st_scene_create_ex(const void *blob, size_t size, uint32_t flags, StSceneHandle **out);

/* Mesh m's BVH node slots and Triangle records (leaf order) as resident on
   the device (reference layouts). node_cap: capacity of nodes. */
int st_scene_mesh_bvh(const StSceneHandle *scene, uint32_t m, StBVHNodeRec *nodes, uint32_t node_cap,
                      StTriangleRec *tris, uint32_t *node_count, uint32_t *height);

/* Re-copy the host-side scene to the device (after st_scene_* edits). */
int st_scene_upload(StSceneHandle *scene);
void st_scene_destroy(StSceneHandle *scene);
int st_scene_info(const StSceneHandle *scene, StSceneInfo *info);
/* Scene BVH as built (reference node layout). nodes: 2*prim_count entries. */
int st_scene_bvh(const StSceneHandle *scene, StBVHNodeRec *nodes, uint32_t *leaf_ids, uint32_t *node_count,
                 uint32_t *height);

#define ST_REBUILD_SAH 0u   /* the reference's SAH builder, on the host */
#define ST_REBUILD_LBVH 1u  /* Morton LBVH on the device */

/* Move primitives: replace records [first, first+count) (transforms, flags;
   updates both copies). The scene BVH is not touched: call st_scene_rebuild,
   or keep the stale BVH as the reference does when it only calls
   uploadPrimitives (SlimTracin examples/02_Geometry.c:70-75). */
int st_scene_set_prims(StSceneHandle *scene, const StPrimRec *prims, uint32_t first, uint32_t count);
/* Rebuild the scene BVH over the current primitives (the reference's
   updateSceneBVH, builder_top_down.h:366-381, when primitives move). `ms`
   (optional) receives the build time: device time for LBVH. */
int st_scene_rebuild(StSceneHandle *scene, uint32_t mode, float *ms);

/* One request, one launch, blocking download: the plain per-request path,
   like renderSceneOnGPU (render/raytracer.h:258-299). hits and counts may be
   NULL. */
int st_render(StSceneHandle *scene, const StView *view, StFloatPixel *pixels, StHitRec *hits, StRayCounts *counts);

/* st_render without any allocation: no hipMalloc, hipFree, page-locked or
   heap allocation in the call. No counterpart in the reference; for callers
   that cannot allocate (code under HIP graph capture, tight serving loops,
   harnesses that forbid allocation). Every per-launch device buffer comes
   from a file-scope __device__ scratch array of ST_NOALLOC_SCRATCH_BYTES
   (128 MiB) per GPU, held until the process ends, whose address st_init
   takes with hipGetSymbolAddress; a request whose buffers do not fit (about
   500k traced pixels in Beauty mode with 3 lights, more in the other modes)
   fails with ST_ERR_MEMORY. pixels ((width+1)*(height+1) FloatPixels) is
   required, hits and counts may be NULL. Every call on a GPU shares that
   GPU's array and its contents stay from the last call: calls are
   serialised by the library lock within a process, and calls that overlap on
   one GPU from several processes race. Results equal st_render's (tiles are
   claimed in tile order: the array keeps no tile history). */
int st_render_noalloc(StSceneHandle *scene, const StView *view, StFloatPixel *pixels, StHitRec *hits,
                      StRayCounts *counts);

/* n requests fused into one launch. pixels[i] / hits[i] may be NULL. */
int st_render_batch(StSceneHandle *const *scenes, const StView *views, uint32_t n, StFloatPixel *const *pixels,
                    StHitRec *const *hits, StRayCounts *counts, uint32_t flags);

/* n requests fused into one launch, each frame resolved on the device to its
   window image as drawViewportToWindowContent does on the host
   (viewport/viewport.h:12-66): bgra[i] receives width*height BGRA8 pixels
   (4 B each instead of the traced grid's 24-B FloatPixels). Alpha is 255. */
int st_render_batch_bgra8(StSceneHandle *const *scenes, const StView *views, uint32_t n, uint32_t *const *bgra,
                          StRayCounts *counts, uint32_t flags);

/* Launch shape of the primary stage (process-wide): tiles per wavefront and
   grid; the depth loop's stages run at their own fixed shapes. */
typedef struct StTuning {
    uint32_t tile_shift;     /* tile = (1 << tile_shift) x (64 >> tile_shift) pixels, one wavefront per tile */
    uint32_t blocks_per_cu;  /* persistent grid: blocks per compute unit (4 wavefronts each); one wavefront
                                per tile: at most this many blocks resident per compute unit (4 or more: no cap) */
    uint32_t persistent;     /* 1: persistent tile queue (st_render), tiles claimed from one counter; 2: the
                                same, each XCD's blocks first claiming its own eighth of the claim order (a
                                contiguous run of tiles, so its L2 serves a compact part of the frames), then
                                the other eighths as they run out; 0: one wavefront per tile */
    uint32_t cost_order;     /* 1: tiles claimed slowest first, from their times in the previous launch of the
                                same shape on the same stream; 0: in tile order (request by request, row by
                                row); 2: request by request in super-tiles of 8 x 8 tiles, so that a run of
                                claims covers a compact part of a frame */
    uint32_t ray_sort;       /* 1: the bounce stage takes its rays binned by request and direction octant
                                (a counting sort after the primary stage); 0: in the order they were queued */
    uint32_t stack_entries;  /* LDS slots per traversal stack and lane in the primary and bounce stages of a
                                persistent launch (a power of two, 4-64): older entries go to the launch's
                                spill buffer in device memory; 0: the whole stack the launch's scenes need */
    uint32_t primary_waves;  /* 4: the primary stage held to 4 waves per SIMD (when its stacks leave LDS for
                                them); 3 or 0: 3 waves per SIMD */
    uint32_t bounce_order;   /* which pixels' depth loops the bounce stage starts first: 0 the refracting
                                ones, 1 the others (the stage ends with its longest chains) */
} StTuning;

int st_set_tuning(const StTuning *tuning);
int st_get_tuning(StTuning *tuning);
/* Autotuner: times each candidate shape and claim order on these requests
   (median of reps render-kernel times), then the fastest with and without
   the bounce-ray sort, with either persistent tile claim (1 or 2), in
   super-tile claim order (cost_order 2), then with
   traversal-stack windows (whole, 8, 16, 32 slots) and primary-stage
   occupancies (3, 4 waves per SIMD), then with the other bounce_order, and
   keeps the fastest; returns it and its time. */
int st_autotune(StSceneHandle *const *scenes, const StView *views, uint32_t n, uint32_t reps, StTuning *best,
                float *best_ms);

/* The SSB pass of a view on the device (updateSceneSSB, render/SSB.h:137-173;
   render launches run it themselves): per primitive the rect as computeSSB
   leaves screen_bounds (min.x, min.y, max.x, max.y) and the flags with
   IS_VISIBLE recomputed, the same values as st_host_view_info. */
int st_ssb(StSceneHandle *scene, const StView *view, int32_t *rects, uint8_t *flags);

/* Picking at n window points (x, y pairs, window pixels from the top
   left): the closest primitive under each point, or a light's glow sphere
   in front of it, with the world hit point and its view-space distance, as
   manipulateSelection's first button press computes them
   (viewport/manipulation.h:55-124), on the device, from a fresh hit state
   per point. */
int st_pick(StSceneHandle *scene, const StView *view, uint32_t n, const uint32_t *xy, StPick *out);

/* Page-locked host memory for frame outputs: downloads into it run at the
   link's full rate (pageable destinations are staged by the runtime). */
int st_host_alloc(size_t bytes, void **out);
void st_host_free(void *p);

/* The same display resolve of a caller's traced grid (grid_w x grid_h
   FloatPixels, host memory) into (grid_w-1) x (grid_h-1) BGRA8 pixels. */
int st_resolve_bgra8(const StFloatPixel *grid, uint32_t grid_w, uint32_t grid_h, uint32_t *out);

/* As st_render_batch without downloads; returns the duration of the
   launch's device work (all stages) measured with HIP events on the launch
   stream. */
int st_render_timed(StSceneHandle *const *scenes, const StView *views, uint32_t n, uint32_t flags, float *kernel_ms,
                    StRayCounts *counts);

/* Per-stage durations (ms, HIP events on the launch stream) of the last
   st_render_timed call made with ST_BATCH_STAGE_TIMES, in launch order: the
   primary stage (with the first level of the Beauty depth loop: shading
   entry, shadow rays, shading), then, for a depth above 1, the bounce
   stage (every bounce ray and every further level's entry, shadow rays and
   shading). *count = number of stages; up to cap values are written. Not in the reference (bench
   instrumentation). */
int st_last_stage_ms(float *ms, uint32_t cap, uint32_t *count);

/* Test hook: every traversal of later launches stops after `cap` steps
   (0: only the bound no valid tree reaches, the node count). A traversal
   that stops this way makes its launch fail with ST_ERR_SCENE instead of
   looping; tests use a tiny cap to exercise that path. Not in the
   reference, whose traversal loops have no bound (trace.h:8-76,
   mesh.h:38-111). */
int st_debug_step_cap(uint32_t cap);

/* The redo stage: a mesh traversal of the main stages that takes more than
   abandon_steps node-pair steps (0: ST_ABANDON_STEPS, 2048) gives up and its
   pixel is traced again by a group of blocks, which splits traversals of
   meshes of at least coop_min_tris triangles (0: ST_COOP_MIN_TRIS, 4096)
   across its threads. Frames do not depend on either value; tests use tiny ones to send
   every pixel through the redo stage. */
int st_debug_redo(uint32_t abandon_steps, uint32_t coop_min_tris);
/* Pixels the redo stage traced again in the last launch on the library's
   main context (st_render, st_render_batch, st_render_timed). */
int st_last_redo_count(uint32_t *count);

/* Graphs captured and graph launches made so far (ST_BATCH_GRAPH). Not in
   the reference (the reference issues one blocking launch per frame). */
int st_graph_stats(uint64_t *captures, uint64_t *launches);

/* Host-side introspection (no device needed): what the library computes on
   the host for a blob and a view. */
/* SSB rects (min.x, min.y, max.x, max.y, as computeSSB leaves screen_bounds)
   and flags after updateSceneSSB (render/SSB.h:137-173); camera as 17 floats:
   projection plane start/right/down (core/init.h:545-559), rotation_inverted,
   rotation quaternions (scene/xform.h:8-20). */
int st_host_view_info(const void *blob, size_t size, const StView *view, int32_t *rects, uint8_t *flags,
                      float *camera);
/* Scene BVH built for a blob (updateSceneBVH, builder_top_down.h:366-381). */
int st_host_scene_bvh(const void *blob, size_t size, StBVHNodeRec *nodes, uint32_t *leaf_ids, uint32_t *node_count,
                      uint32_t *height);
/* Parse and validate a blob; `why` receives the reason of a rejection. */
int st_host_validate(const void *blob, size_t size, char *why, size_t cap);
/* Reference SAH build over n leaf boxes (6 floats: min xyz, max xyz each;
   buildBVH, builder_top_down.h:204-270). nodes: 2n entries. Returns the node
   count. */
uint32_t st_sah_build_boxes(const float *boxes, uint32_t n, uint32_t max_leaf, StBVHNodeRec *nodes, uint32_t *leaf_ids,
                            uint32_t *height);
/* Mesh BVH and triangle records (updateMeshBVH, builder_top_down.h:272-364). */
int st_build_mesh(uint32_t tri_count, const float *pos, const uint32_t *pidx, const float *nrm,
                  const uint32_t *nidx, uint32_t normals_count, const float *uvs, const uint32_t *uidx,
                  uint32_t uvs_count, StBVHNodeRec *nodes, uint32_t *node_count, uint32_t *height,
                  StTriangleRec *tris, uint32_t *leaf_ids);
/* Binned SAH build over n leaf boxes (layout as st_sah_build_boxes; 32
   centroid bins per axis). No counterpart in the reference: a fast builder
   of SAH-quality trees in the reference's BVH layout for meshes of millions
   of triangles. Returns the node count, 0 on failure. */
uint32_t st_binned_build_boxes(const float *boxes, uint32_t n, uint32_t max_leaf, StBVHNodeRec *nodes,
                               uint32_t *leaf_ids, uint32_t *height);
/* Mesh BVH (binned SAH, one triangle per leaf) and triangle records in leaf order,
   the outputs as st_build_mesh's (the triangle records are the same
   function of the triangle, builder_top_down.h:272-364). */
int st_build_mesh_binned(uint32_t tri_count, const float *pos, const uint32_t *pidx, const float *nrm,
                         const uint32_t *nidx, uint32_t normals_count, const float *uvs, const uint32_t *uidx,
                         uint32_t uvs_count, StBVHNodeRec *nodes, uint32_t *node_count, uint32_t *height,
                         StTriangleRec *tris, uint32_t *leaf_ids);

#ifdef __cplusplus
}
#endif
#endif
