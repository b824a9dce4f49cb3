/*
 * st_lbvh.hip - Morton LBVH rebuild of a scene BVH on the device.
 *
 * The reference rebuilds the scene BVH on the CPU with its SAH builder
 * whenever a primitive moves (updateSceneBVH, builder_top_down.h:366-381).
 * Here the rebuild runs on the GPU as a linear BVH: 30-bit Morton codes of
 * the primitives' world-box centroids, a stable radix sort of (code,
 * primitive index) pairs (st_sort.hip, so equal codes keep index order: the
 * order of the unique keys code << 32 | index), the binary radix tree of
 * Karras 2012 ("Maximizing parallelism in the construction of BVHs, octrees
 * and k-d trees") with one lane per internal node, boxes bottom-up, and
 * output in the reference's node layout so the traversal is unchanged:
 *   node 0 = root; the children of radix-tree internal node k are nodes
 *   2k+1 and 2k+2 (adjacent, as the traversal requires); a leaf holds one
 *   primitive (child_count 1, first_child_id = its sorted position, and
 *   leaf_ids[sorted position] = primitive index); internal boxes are
 *   merge(left, right); depth = tree depth.
 * The primitive boxes are the reference's leaf boxes (getPrimitiveAABB or
 * the mesh AABB, transformed by transformAABB, render/AABB.h:44-84). Steps
 * after the boxes are the mesh LBVH's (st_lbvh_tree, st_mesh_lbvh.hip) with
 * leaves of one primitive; every kernel is multi-block, so the primitive
 * count is not limited by a workgroup. The CPU oracle restates the same
 * specification with a top-down construction; tests compare the nodes bit
 * for bit.
 */
#include <hip/hip_runtime.h>

#include "st_lbvh.h"
#include "st_mesh_lbvh.h"

#define LBVH_THREADS 256u

/* transformAABB of the primitive's object-space box (render/AABB.h:44-84) */
__device__ static sbox prim_world_box(const StDPrim &p, const float *mesh_aabbs, uint32_t mesh_count) {
    uint32_t flags = p.flags_mat & 0xFFu;
    sbox bx;
    if (p.type == ST_PRIM_MESH && p.mesh_id < mesh_count) {
        const float *a = mesh_aabbs + 6 * p.mesh_id;
        bx.min = V3(a[0], a[1], a[2]);
        bx.max = V3(a[3], a[4], a[5]);
    } else {
        bx.max = V3s(p.type == ST_PRIM_TETRAHEDRON ? (1.73205080757f / 3) : 1);
        bx.min = v_neg(bx.max);
    }
    float x = bx.min.x, y = bx.min.y, z = bx.min.z, X = bx.max.x, Y = bx.max.y, Z = bx.max.z;
    sv3 c[8] = {V3(x, y, z), V3(x, y, Z), V3(x, Y, z), V3(x, Y, Z), V3(X, y, z), V3(X, y, Z), V3(X, Y, z), V3(X, Y, Z)};
    x = y = z = +INFINITY;
    X = Y = Z = -INFINITY;
    sv3 scale = V3(p.scale[0], p.scale[1], p.scale[2]);
    sv3 pos = V3(p.pos[0], p.pos[1], p.pos[2]);
    sq4 rot;
    rot.v = V3(p.rot[0], p.rot[1], p.rot[2]);
    rot.w = p.rot[3];
    for (int k = 0; k < 8; k++) {
        sv3 q = c[k];
        if (flags & ST_IS_SCALED) q = v_mul(q, scale);
        if (flags & ST_IS_ROTATED) q = q_rot(q, rot);
        if (flags & ST_IS_TRANSLATED) q = v_add(q, pos);
        if (q.x > X) X = q.x;
        if (q.y > Y) Y = q.y;
        if (q.z > Z) Z = q.z;
        if (q.x < x) x = q.x;
        if (q.y < y) y = q.y;
        if (q.z < z) z = q.z;
    }
    bx.min = V3(x, y, z);
    bx.max = V3(X, Y, Z);
    return bx;
}

/* world boxes, centroids (-0 folded to +0) and per-block centroid bounds */
__global__ void __launch_bounds__(LBVH_THREADS) k_sl_boxes(const StDPrim *prims, const float *mesh_aabbs,
                                                           uint32_t mesh_count, uint32_t n, sbox *boxes, float *cent,
                                                           float *partials) {
    __shared__ float red[6][LBVH_THREADS];
    const uint32_t tid = threadIdx.x, i = blockIdx.x * LBVH_THREADS + tid;
    float c[3] = {INFINITY, INFINITY, INFINITY}, C[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (i < n) {
        const sbox b = prim_world_box(prims[i], mesh_aabbs, mesh_count);
        boxes[i] = b;
        const float m[3] = {(b.min.x + b.max.x) * 0.5f + 0.0f, (b.min.y + b.max.y) * 0.5f + 0.0f,
                            (b.min.z + b.max.z) * 0.5f + 0.0f};
        for (int a = 0; a < 3; a++) {
            cent[3 * (size_t)i + a] = m[a];
            c[a] = m[a];
            C[a] = m[a];
        }
    }
    for (int a = 0; a < 3; a++) { red[a][tid] = c[a]; red[3 + a][tid] = C[a]; }
    __syncthreads();
    for (uint32_t s = LBVH_THREADS / 2; s > 0; s >>= 1) {
        if (tid < s)
            for (int a = 0; a < 3; a++) {
                float o = red[a][tid + s];
                red[a][tid] = o < red[a][tid] ? o : red[a][tid];
                o = red[3 + a][tid + s];
                red[3 + a][tid] = o > red[3 + a][tid] ? o : red[3 + a][tid];
            }
        __syncthreads();
    }
    if (tid < 6) partials[6 * (size_t)blockIdx.x + tid] = red[tid][0];
}

static size_t r256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t st_lbvh_work_bytes(uint32_t n) {
    const size_t nb = (n + LBVH_THREADS - 1) / LBVH_THREADS;
    return r256(24ull * n) + r256(12ull * n) + r256(24ull * nb) + r256(4ull * n) + st_lbvh_tree_work_bytes(n);
}

extern "C" int st_lbvh_launch(hipStream_t stream, const StDPrim *prims, const float *mesh_aabbs, uint32_t mesh_count,
                              uint32_t n, char *work, uint32_t *nodes_out, uint32_t *leaf_ids_out,
                              uint32_t *height_out) {
    if (n == 0) return -1;
    const uint32_t nb = (n + LBVH_THREADS - 1) / LBVH_THREADS;
    char *p = work;
    sbox *boxes = (sbox *)p; p += r256(24ull * n);
    float *cent = (float *)p; p += r256(12ull * n);
    float *partials = (float *)p; p += r256(24ull * nb);
    uint32_t *keys = (uint32_t *)p; p += r256(4ull * n);
    hipLaunchKernelGGL(k_sl_boxes, dim3(nb), dim3(LBVH_THREADS), 0, stream, prims, mesh_aabbs, mesh_count, n, boxes, cent,
                       partials);
    if (hipGetLastError() != hipSuccess) return -2;
    return st_lbvh_tree(stream, boxes, cent, partials, nb, n, 1u, p, keys, leaf_ids_out, nodes_out, height_out) ? -2 : 0;
}
