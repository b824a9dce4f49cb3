/*
 * st_host.h - host-side scene model of the HIP library (C).
 *
 * A StScene is a parsed scene blob whose BVHs are built and whose records
 * are already in the device layout of st_layout.h, ready to be copied to HBM
 * in one pass. Building follows the reference's init order (app.h:287-297):
 * mesh BVHs (updateMeshBVH), then the scene BVH over the primitives
 * (updateSceneBVH). Per-request state (camera basis, SSB rects) lives in a
 * StHostView.
 */
#ifndef ST_HOST_H
#define ST_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "st_blob.h"
#include "slimtracin_seam.h"
#include "../st_layout.h"
#include "../st_math.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct StScene {
    uint32_t prim_count, material_count, light_count, mesh_count, texture_count;
    float ambient[3];
    StDPrim *prims;
    StDMat *mats;
    StDLight *lights;
    /* scene BVH (reference node layout), node 0 = root */
    StBVHNodeRec *scene_nodes;
    uint32_t scene_node_count, scene_height;
    uint32_t *scene_leaf_ids;
    /* meshes: nodes of all meshes concatenated; mesh m's node k is
       mesh_nodes[meshes[m].node_offset + k]; offsets are odd so that child
       pairs (odd, even) start on 64-byte boundaries. */
    StDMesh *meshes;
    sbox *mesh_aabbs;
    StBVHNodeRec *mesh_nodes;
    uint32_t mesh_node_total;
    StDTriHot *tri_hot;
    StDTriCold *tri_cold;
    uint32_t tri_total;
    uint32_t max_mesh_height;
    /* textures */
    StDTex *textures;
    StDMip *mips;
    uint32_t mip_total;
    uint8_t *texels;
    uint64_t texel_bytes;
    uint32_t has_emissive_quads;
    uint32_t mesh_stack_size;
    uint32_t scene_real_height;    /* levels of the scene BVH as walked */
    /* meshes whose BVH and triangle records are built on the device: their
       host node/triangle ranges stay zero; geometry points into the caller's
       scene data and is valid only while the scene is being created */
    uint8_t *mesh_on_device;
    struct StMeshPart *deferred;
    uint32_t mesh_host_data;       /* 0: mesh_nodes / tri_hot / tri_cold are not held on the host */
} StScene;

/* A scene as separate arrays in the blob record layouts (st_blob.h); the
   pointers are borrowed. Meshes with rec.prebuilt set carry their BVH and
   Triangle records (nodes, tris) and may omit the geometry (pos/pidx NULL):
   such a mesh is device-ready, as uploadMeshBVHs hands it over. */
typedef struct StMeshPart StMeshPart;
struct StMeshPart {
    StMeshRec rec;
    const float *pos, *nrm, *uvs;
    const uint32_t *pidx, *nidx, *uidx;
    const StBVHNodeRec *nodes;
    const StTriangleRec *tris;
};

typedef struct StMipPart {
    uint16_t width, height;
    const uint8_t *quads;          /* (width+1)*(height+1) texel quads of 12 B */
} StMipPart;

typedef struct StTexturePart {
    StTextureRec rec;
    const StMipPart *mips;         /* rec.mip_count entries */
} StTexturePart;

typedef struct StSceneParts {
    uint32_t prim_count, material_count, light_count, mesh_count, texture_count;
    float ambient[3];
    const StPrimRec *prims;
    const StMaterialRec *mats;
    const StLightRec *lights;
    const StMeshPart *meshes;
    const StTexturePart *textures;
    /* optional scene BVH built elsewhere; NULL = SAH build over the primitives */
    const StBVHNodeRec *scene_nodes;
    uint32_t scene_node_count, scene_height;
    const uint32_t *scene_leaf_ids;
    /* builder of meshes without a prebuilt BVH: ST_MESH_BUILD_SAH (the
       reference's, host), ST_MESH_BUILD_DEVICE (left for the device LBVH),
       ST_MESH_BUILD_BINNED (binned SAH, host) */
    uint32_t mesh_builder;
} StSceneParts;

#define ST_MESH_BUILD_SAH 0u
#define ST_MESH_BUILD_DEVICE 1u
#define ST_MESH_BUILD_BINNED 2u

/* Validate the parts and build the device-ready scene. Returns NULL on
   malformed input; *err (if given) receives a short reason. */
StScene *st_scene_from_parts(const StSceneParts *parts, const char **err);

/* Parse a blob into parts and build its BVHs (st_scene_from_parts). */
StScene *st_scene_from_blob(const void *blob, size_t size, const char **err);
StScene *st_scene_from_blob_ex(const void *blob, size_t size, uint32_t mesh_builder, const char **err);

/* Device material record of a blob material record. */
StDMat st_mat_from_rec(const StMaterialRec *q);
void st_scene_free(StScene *sc);

/* Device bytes the resident scene occupies (for HBM budgeting). */
uint64_t st_scene_device_bytes(const StScene *sc);

/* Recompute the scene BVH over the current primitive transforms with the SAH
   builder (updateSceneBVH, builder_top_down.h:366-381). */
int st_scene_rebuild_sah(StScene *sc);

/* Tree check of a BVH in the reference layout; returns its height or -1. */
int st_bvh_validate(const StBVHNodeRec *nodes, uint32_t count, uint64_t leaf_limit);

/* Device record of a blob primitive record. */
StDPrim st_prim_from_rec(const StPrimRec *r);

/* World AABB of a primitive (getPrimitiveAABB + transformAABB). */
sbox st_prim_world_box(const StScene *sc, uint32_t i);

/* SAH BVH over leaf boxes (buildBVH, builder_top_down.h:204-270): returns the
   node count; nodes needs 2n entries, leaf_ids n. */
uint32_t st_sah_build(const sbox *boxes, uint32_t n, uint32_t max_leaf, StBVHNodeRec *nodes, uint32_t *leaf_ids,
                      uint32_t *height);

/* Binned SAH BVH over leaf boxes (32 centroid bins per axis; no counterpart
   in the reference): returns the node count (0 on failure); nodes needs
   2n entries, leaf_ids n. */
uint32_t st_binned_build(const sbox *boxes, uint32_t n, uint32_t max_leaf, StBVHNodeRec *nodes, uint32_t *leaf_ids,
                         uint32_t *height);

/* Mesh BVH + triangle records (updateMeshBVH, builder_top_down.h:272-364). */
int st_build_mesh(uint32_t tri_count, const float *pos, const uint32_t *pidx, const float *nrm,
                  const uint32_t *nidx, uint32_t normals_count, const float *uvs, const uint32_t *uidx,
                  uint32_t uvs_count, StBVHNodeRec *nodes, uint32_t *node_count, uint32_t *height,
                  StTriangleRec *tris, uint32_t *leaf_ids);

/* Camera + frame set-up for one request. */
typedef struct StHostView {
    StDView dv;            /* camera, grid, mode; pointers filled by the runtime */
    float start[3], right[3], down[3];
    float rot[4], rot_inv[4];
    sm3 rotation_matrix_inverted;
    float *row_starts;     /* 4 floats per row */
    int32_t *rects;        /* 4 ints per primitive, empty when not visible */
    uint8_t *flags;        /* primitive flags after the SSB pass */
    int32_t *raw_rects;    /* screen_bounds as computeSSB left them */
    float focal_length;
    int device_ssb;        /* rects not computed on the host: the device SSB pass fills them */
} StHostView;

int st_view_prepare(const StScene *sc, const StView *v, StHostView *hv);
/* with_ssb 0: no SSB pass (rects, raw_rects and flags are left unset; the
   device computes the gate rects) */
int st_view_prepare_ex(const StScene *sc, const StView *v, StHostView *hv, int with_ssb);
/* The same without allocating: row starts into the caller's buffer (4 floats
   per traced row, (height + 1) rows), no host SSB pass (the device computes
   the gate rects). Not to be passed to st_view_release. */
int st_view_prepare_rows(const StScene *sc, const StView *v, StHostView *hv, float *row_starts);

/* Set-up for a frame described the way renderSceneOnGPU receives it: the
   viewport's projection plane and camera, and per-primitive screen bounds
   already computed by the caller (gated by each primitive's IS_VISIBLE). */
int st_view_prepare_frame(const StScene *sc, const StSeamFrame *f, const int32_t *screen_bounds, StHostView *hv);
void st_view_release(StHostView *hv);

/* Record an error message for st_last_error; returns code. */
int st_host_error(int code, const char *msg);

#ifdef __cplusplus
}
#endif
#endif
