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

/* Device scratch bytes for a T-triangle build (sort_bytes: the radix sort's share). */
size_t st_mesh_lbvh_work_bytes(uint32_t T, size_t *sort_bytes);

/* Build into `nodes` (2T-1 slots of 8 words, reference layout) and the hot /
   cold triangle records (T each, in leaf order); *height_dev receives the
   tree height. All work is queued on `stream`. Returns 0 on success. */
int st_mesh_lbvh_build(hipStream_t stream, StMeshGeomDev in, char *work, size_t sort_bytes, uint32_t *nodes,
                       StDTriHot *hot, StDTriCold *cold, uint32_t *height_dev);

#endif
