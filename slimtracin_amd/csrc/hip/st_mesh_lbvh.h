/*
 * st_mesh_lbvh.h - device mesh BVH builder (see st_mesh_lbvh.hip).
 */
#ifndef ST_MESH_LBVH_H
#define ST_MESH_LBVH_H

#include <hip/hip_runtime.h>

#include "../st_layout.h"
#include "../st_math.h"

/* A mesh's geometry in device memory (nrm/uvs NULL when absent). */
typedef struct StMeshGeomDev {
    const float *pos, *nrm, *uvs;
    const uint32_t *pidx, *nidx, *uidx;
    uint32_t T;
} StMeshGeomDev;

/* Radix-tree scratch: per internal node its sorted range, split, parent,
   arrival counter and box; per sorted leaf its parent. */
typedef struct StMeshLbvhWork {
    uint32_t *ifirst, *ilast, *isplit, *iparent, *leaf_parent, *flags;
    float *ibox;
} StMeshLbvhWork;

/* Steps 2-6 of a build over T boxes (centroids `cent`, 3 floats each, and
   nb per-256 partial centroid bounds, 6 floats each): Morton codes, the
   stable radix sort, the radix tree, bottom-up boxes and the node output
   with leaves of up to max_leaf boxes. `keys` and `order` (T words each)
   receive the sorted codes and the box index at each sorted position;
   `work` holds st_lbvh_tree_work_bytes(T). */
size_t st_lbvh_tree_work_bytes(uint32_t T);
int st_lbvh_tree(hipStream_t stream, const sbox *boxes, const float *cent, const float *partials, uint32_t nb,
                 uint32_t T, uint32_t max_leaf, char *work, uint32_t *keys, uint32_t *order, uint32_t *nodes,
                 uint32_t *height_dev);

/* Device scratch bytes for a T-triangle build (sort_bytes: 0, kept for the caller's layout). */
size_t st_mesh_lbvh_work_bytes(uint32_t T, size_t *sort_bytes);

/* Build into `nodes` (2T-1 slots of 8 words, reference layout) and the hot /
   cold triangle records (T each, in leaf order); *height_dev receives the
   tree height. All work is queued on `stream`. Returns 0 on success. */
int st_mesh_lbvh_build(hipStream_t stream, StMeshGeomDev in, char *work, size_t sort_bytes, uint32_t *nodes,
                       StDTriHot *hot, StDTriCold *cold, uint32_t *height_dev);

#endif
