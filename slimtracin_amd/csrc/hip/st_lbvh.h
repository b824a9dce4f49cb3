/*
 * st_lbvh.h - device Morton LBVH builder (see st_lbvh.hip).
 */
#ifndef ST_LBVH_H
#define ST_LBVH_H

#include <hip/hip_runtime.h>

#include "../st_layout.h"
#include "../st_math.h"
#include "st_blob.h"

#define ST_LBVH_MAX 4096u   /* primitives per scene BVH built in one workgroup */

/* Device scratch of one build (n primitives). */
typedef struct StLbvhWork {
    float *boxes;        /* 6 floats per primitive */
    float *cent;         /* 3 floats per primitive */
    uint32_t *codes;     /* n, sorted */
    uint32_t *parent;    /* 2n: radix-tree internal node owning each output slot */
    uint32_t *kind;      /* 2n: 1 = leaf */
    uint32_t *first;     /* 2n: first child slot or sorted position */
    uint32_t *depth;     /* 2n */
    uint32_t *slot;      /* n: output slot of each internal node */
    sbox *node_boxes;    /* 2n */
} StLbvhWork;

/* Bytes of StLbvhWork for n primitives (each array 256-byte aligned). */
static inline size_t st_lbvh_work_bytes(uint32_t n) {
    size_t a = 0;
    size_t sizes[9] = {24u * n, 12u * n, 4u * n, 8u * n, 8u * n, 8u * n, 8u * n, 4u * n, 48u * n};
    for (int i = 0; i < 9; i++) a += (sizes[i] + 255) & ~(size_t)255;
    return a;
}
static inline StLbvhWork st_lbvh_work_carve(char *base, uint32_t n) {
    StLbvhWork w;
    size_t sizes[9] = {24u * n, 12u * n, 4u * n, 8u * n, 8u * n, 8u * n, 8u * n, 4u * n, 48u * n};
    char *p[9];
    for (int i = 0; i < 9; i++) { p[i] = base; base += (sizes[i] + 255) & ~(size_t)255; }
    w.boxes = (float *)p[0]; w.cent = (float *)p[1]; w.codes = (uint32_t *)p[2];
    w.parent = (uint32_t *)p[3]; w.kind = (uint32_t *)p[4]; w.first = (uint32_t *)p[5];
    w.depth = (uint32_t *)p[6]; w.slot = (uint32_t *)p[7]; w.node_boxes = (sbox *)p[8];
    return w;
}

extern "C" int st_lbvh_launch(hipStream_t stream, const StDPrim *prims, const float *mesh_aabbs, uint32_t mesh_count,
                              uint32_t n, StLbvhWork work, uint32_t *nodes_out, uint32_t *leaf_ids_out,
                              uint32_t *height_out);

#endif
