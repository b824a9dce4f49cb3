/*
 * st_oracle.h - CPU oracle for the per-pixel tracing path.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle is a C restatement of the reference
 * CPU renderer (SlimTracin render/raytracer.h:81-171 and everything below it)
 * used by tests/ and by bench.py's CPU baseline leg to check the HIP library.
 * Product code never links it.
 */
#ifndef ST_ORACLE_H
#define ST_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "st_blob.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Algorithmic event counts of one frame (bytes per event use the reference's
   record sizes: node pair 64, root AABB 24, triangle hot part 60, primitive
   72, mesh-hit attributes 68, texel quad 12, material 84, light 44). */
typedef struct StOracleStats {
    uint64_t primary_rays, shadow_rays, bounce_rays;
    uint64_t node_pairs, root_tests, triangle_tests, primitive_visits;
    uint64_t mesh_attr_fetches, texel_fetches, material_reads, light_reads;
    /* the traversal events (node pairs, root tests, triangle tests,
       primitive visits) split by ray class: primary, shadow rays of the
       first depth-loop level, shadow rays of later levels, bounce rays */
    uint64_t class_events[4][4];
} StOracleStats;

enum { ST_ORC_PRIMARY = 0, ST_ORC_SHADOW0 = 1, ST_ORC_SHADOW_DEEP = 2, ST_ORC_BOUNCE = 3 };

/* Render one frame. pixels and hits have (width+1)*(height+1) entries; hits
   may be NULL. threads > 1 splits rows round-robin with one trace state per
   thread. Returns 0 on success. */
int oracle_render(const void *blob, size_t size, const StView *view, StFloatPixel *pixels,
                  StHitRec *hits, int threads, StOracleStats *stats);

/* oracle_render that also records each pixel's traversal events (node
   pairs + triangle tests + primitive visits over all of its rays) in steps
   ((width+1)*(height+1) entries): the work a pixel costs. */
int oracle_render_steps(const void *blob, size_t size, const StView *view, StFloatPixel *pixels, StHitRec *hits,
                        uint64_t *steps, int threads, StOracleStats *stats);

/* Load a scene blob once (NULL on a malformed blob), render any number of
   views from it (seconds: optional, the render call's own time), free it.
   flags & 1: build meshes that carry no BVH with the Morton LBVH (as
   st_scene_create_ex's ST_SCENE_MESH_LBVH) instead of the reference SAH. */
void *oracle_scene_open(const void *blob, size_t size, uint32_t flags);
int oracle_scene_render(void *scene, const StView *view, StFloatPixel *pixels, StHitRec *hits, int threads,
                        StOracleStats *stats, double *seconds);
void oracle_scene_close(void *scene);

/* Same, but loads once and renders `frames` times, recording per-frame
   seconds of the render call alone. */
int oracle_render_timed(const void *blob, size_t size, const StView *view, StFloatPixel *pixels,
                        int threads, int frames, double *seconds, StOracleStats *stats);

/* Picking at n window points (x, y pairs) on a fresh trace state per point
   (viewport/manipulation.h:55-124). */
int oracle_pick(const void *blob, size_t size, const StView *view, uint32_t n, const uint32_t *xy, StPick *out);

/* Scene BVH (SAH, leaves <= 2) as built for the blob. nodes: 2*prim_count. */
int oracle_scene_bvh(const void *blob, size_t size, StBVHNodeRec *nodes, uint32_t *leaf_ids,
                     uint32_t *node_count, uint32_t *height);

/* Mesh BVH (SAH, leaves <= 4) and triangle records of mesh m. nodes: 2*tris. */
int oracle_mesh_bvh(const void *blob, size_t size, uint32_t m, StBVHNodeRec *nodes,
                    StTriangleRec *tris, uint32_t *leaf_ids, uint32_t *node_count, uint32_t *height);

/* Mesh BVH of mesh m built as the Morton LBVH with one triangle per leaf
   (the device mesh builder's specification), and the triangle records in
   its leaf order. nodes: 2*tris slots (2*tris - 1 used). */
int oracle_mesh_lbvh(const void *blob, size_t size, uint32_t m, StBVHNodeRec *nodes, StTriangleRec *tris,
                     uint32_t *leaf_ids, uint32_t *node_count, uint32_t *height);

/* Mesh BVH of mesh m built by the binned SAH builder (one triangle per leaf;
   the HIP library's host builder for large meshes) and the triangle records in its
   leaf order. nodes: 2*tris slots. */
int oracle_mesh_binned(const void *blob, size_t size, uint32_t m, StBVHNodeRec *nodes, StTriangleRec *tris,
                       uint32_t *leaf_ids, uint32_t *node_count, uint32_t *height);

/* Screen-space bounds (4 ints per primitive) and flags for a view. */
int oracle_ssb(const void *blob, size_t size, const StView *view, int32_t *rects, uint8_t *flags);

/* Camera set-up for a view: 9 floats projection plane (start, right, down),
   4 floats inverted rotation quaternion, 4 floats rotation quaternion. */
int oracle_camera(const StView *view, float *out);

/* Plain SAH build over N caller-given leaf boxes (6 floats each), as
   buildBVH does for a scene or mesh. Used to check the HIP library's fast
   builder against the plain quicksort. Returns node count. */
uint32_t oracle_sah_build(const float *boxes, uint32_t n, uint32_t max_leaf, StBVHNodeRec *nodes,
                          uint32_t *leaf_ids, uint32_t *height);

/* Morton LBVH over primitive boxes (CPU restatement of the device builder).
   nodes: 2*n entries. Returns node count. */
uint32_t oracle_lbvh_build(const float *boxes, uint32_t n, StBVHNodeRec *nodes, uint32_t *leaf_ids,
                           uint32_t *height);

/* Binned SAH build over n caller-given leaf boxes (6 floats each).
   nodes: 2*n entries. Returns node count. */
uint32_t oracle_binned_build(const float *boxes, uint32_t n, uint32_t max_leaf, StBVHNodeRec *nodes,
                             uint32_t *leaf_ids, uint32_t *height);

/* World AABBs of the blob's primitives (getPrimitiveAABB + transformAABB). */
int oracle_prim_boxes(const void *blob, size_t size, float *boxes);

/* Display resolve of a traced QCAA grid (grid_w x grid_h FloatPixels) into
   (grid_w-1) x (grid_h-1) BGRA8 pixels (drawViewportToWindowContent). */
void oracle_resolve(const StFloatPixel *grid, uint32_t grid_w, uint32_t grid_h, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
