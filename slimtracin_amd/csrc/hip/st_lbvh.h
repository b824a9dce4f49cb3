/*
 * st_lbvh.h - device Morton LBVH builder (see st_lbvh.hip).
 */
#ifndef ST_LBVH_H
#define ST_LBVH_H

#include <hip/hip_runtime.h>

#include "../st_layout.h"
#include "../st_math.h"
#include "st_blob.h"

/* Device scratch bytes of a build over n primitives. */
size_t st_lbvh_work_bytes(uint32_t n);

/* Rebuilds the scene BVH of n primitives into nodes_out (2n-1 nodes of 8
   words) and leaf_ids_out (n), its height to *height_out; all work queued
   on `stream`. Returns 0 on success. */
extern "C" int st_lbvh_launch(hipStream_t stream, const StDPrim *prims, const float *mesh_aabbs, uint32_t mesh_count,
                              uint32_t n, char *work, uint32_t *nodes_out, uint32_t *leaf_ids_out,
                              uint32_t *height_out);

#endif
