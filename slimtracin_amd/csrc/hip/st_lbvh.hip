/*
 * st_lbvh.hip - Morton LBVH rebuild of a scene BVH on the device.
 *
 * The reference rebuilds the scene BVH on the CPU with its SAH builder
 * whenever a primitive moves (updateSceneBVH, builder_top_down.h:366-381).
 * Here the rebuild runs on the GPU as a linear BVH: 30-bit Morton codes of
 * the primitives' world-box centroids, a sort of the unique 64-bit keys
 * (code << 32 | original index), the binary radix tree of Karras 2012
 * ("Maximizing parallelism in the construction of BVHs, octrees and k-d
 * trees") built with one lane per internal node, and output in the
 * reference's node layout so the traversal is unchanged:
 *   node 0 = root; the children of radix-tree internal node k are nodes
 *   2k+1 and 2k+2 (adjacent, as the traversal requires); a leaf holds one
 *   primitive (child_count 1, first_child_id = its sorted position, and
 *   leaf_ids[sorted position] = primitive index); internal boxes are
 *   merge(left, right), filled level by level from the deepest level up;
 *   depth = tree depth.
 * The primitive boxes are the reference's leaf boxes (getPrimitiveAABB or
 * the mesh AABB, transformed by transformAABB, render/AABB.h:44-84).
 *
 * One workgroup of 1024 lanes builds scene BVHs of up to ST_LBVH_MAX
 * primitives with all sorting in LDS. The CPU oracle restates the same
 * specification with a top-down construction; tests compare the nodes bit
 * for bit.
 */
#include <hip/hip_runtime.h>

#include "st_lbvh.h"

#define LBVH_THREADS 1024

__device__ static inline uint32_t expand10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

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

/* common-prefix length of the unique keys at sorted positions i and j;
   -1 outside [0, n) */
__device__ static inline int lcp(const uint32_t *codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = codes[i], b = codes[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

__global__ void __launch_bounds__(LBVH_THREADS) k_lbvh_small(const StDPrim *prims, const float *mesh_aabbs,
                                                               uint32_t mesh_count, uint32_t n, StLbvhWork work,
                                                               uint32_t *nodes_out, uint32_t *leaf_ids_out,
                                                               uint32_t *height_out) {
    __shared__ unsigned long long keys[ST_LBVH_MAX];
    __shared__ float red[6][LBVH_THREADS];
    __shared__ uint32_t s_max_depth;
    const uint32_t tid = threadIdx.x;
    sbox *boxes = reinterpret_cast<sbox *>(work.boxes);

    /* 1. world boxes, centroids (-0 folded to +0) and their bounds */
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = tid; i < n; i += LBVH_THREADS) {
        sbox b = prim_world_box(prims[i], mesh_aabbs, mesh_count);
        boxes[i] = b;
        float c[3] = {(b.min.x + b.max.x) * 0.5f + 0.0f, (b.min.y + b.max.y) * 0.5f + 0.0f,
                      (b.min.z + b.max.z) * 0.5f + 0.0f};
        for (int a = 0; a < 3; a++) {
            work.cent[3 * i + a] = c[a];
            mn[a] = c[a] < mn[a] ? c[a] : mn[a];
            mx[a] = c[a] > mx[a] ? c[a] : mx[a];
        }
    }
    for (int a = 0; a < 3; a++) { red[a][tid] = mn[a]; red[3 + a][tid] = mx[a]; }
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
    const float cmin[3] = {red[0][0], red[1][0], red[2][0]};
    float scale[3];
    for (int a = 0; a < 3; a++) {
        float e = red[3 + a][0] - red[a][0];
        scale[a] = e > 0 ? 1023.0f / e : 0.0f;
    }

    /* 2. keys = code << 32 | index, padded to a power of two */
    uint32_t m = 1;
    while (m < n) m <<= 1;
    for (uint32_t i = tid; i < m; i += LBVH_THREADS) {
        if (i < n) {
            uint32_t q[3];
            for (int a = 0; a < 3; a++) {
                float t = (work.cent[3 * i + a] - cmin[a]) * scale[a];
                t = t < 0 ? 0 : t;
                t = t > 1023.0f ? 1023.0f : t;
                q[a] = (uint32_t)t;
            }
            uint32_t code = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
            keys[i] = ((unsigned long long)code << 32) | i;
        } else
            keys[i] = ~0ull;
    }
    __syncthreads();

    /* 3. bitonic sort of the unique keys */
    for (uint32_t size = 2; size <= m; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = tid; i < m; i += LBVH_THREADS) {
                uint32_t j = i ^ stride;
                if (j > i) {
                    bool up = (i & size) == 0;
                    unsigned long long a = keys[i], b = keys[j];
                    if ((a > b) == up) { keys[i] = b; keys[j] = a; }
                }
            }
            __syncthreads();
        }
    for (uint32_t i = tid; i < n; i += LBVH_THREADS) {
        work.codes[i] = (uint32_t)(keys[i] >> 32);
        leaf_ids_out[i] = (uint32_t)keys[i];
    }
    if (tid == 0) s_max_depth = 0;
    __syncthreads();

    if (n == 1) {
        if (tid == 0) {
            sbox b = boxes[0];
            uint32_t *o = nodes_out;
            o[0] = __float_as_uint(b.min.x); o[1] = __float_as_uint(b.min.y); o[2] = __float_as_uint(b.min.z);
            o[3] = __float_as_uint(b.max.x); o[4] = __float_as_uint(b.max.y); o[5] = __float_as_uint(b.max.z);
            o[6] = 0; o[7] = 1;
            *height_out = 1;
        }
        return;
    }

    /* 4. radix tree: one lane per internal node (Karras 2012, section 4) */
    const int N = (int)n;
    for (int k = (int)tid; k < N - 1; k += LBVH_THREADS) {
        int d = (lcp(work.codes, N, k, k + 1) - lcp(work.codes, N, k, k - 1)) >= 0 ? 1 : -1;
        int dmin = lcp(work.codes, N, k, k - d);
        int lmax = 2;
        while (lcp(work.codes, N, k, k + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (lcp(work.codes, N, k, k + (l + t) * d) > dmin) l += t;
        int j = k + l * d;
        int dnode = lcp(work.codes, N, k, j);
        int s = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if (lcp(work.codes, N, k, k + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        int g = k + s * d + (d < 0 ? -1 : 0);
        int first = k < j ? k : j, last = k < j ? j : k;
        /* children -> output slots 2k+1 (left) and 2k+2 (right) */
        uint32_t lo = 2u * k + 1u, ro = 2u * k + 2u;
        work.parent[lo] = (uint32_t)k;
        work.parent[ro] = (uint32_t)k;
        if (first == g) { work.kind[lo] = 1u; work.first[lo] = (uint32_t)g; }
        else { work.kind[lo] = 0u; work.first[lo] = 2u * g + 1u; work.slot[g] = lo; }
        if (g + 1 == last) { work.kind[ro] = 1u; work.first[ro] = (uint32_t)(g + 1); }
        else { work.kind[ro] = 0u; work.first[ro] = 2u * (g + 1) + 1u; work.slot[g + 1] = ro; }
    }
    if (tid == 0) { work.slot[0] = 0; work.kind[0] = 0; work.first[0] = 1; work.parent[0] = 0xFFFFFFFFu; }
    __syncthreads();

    /* 5. depths (walk to the root) and the deepest level */
    const uint32_t count = 2u * n - 1u;
    for (uint32_t o = tid; o < count; o += LBVH_THREADS) {
        uint32_t dpt = 0, x = o;
        while (x != 0) { x = work.slot[work.parent[x]]; dpt++; }
        work.depth[o] = dpt;
        atomicMax(&s_max_depth, dpt);
    }
    __syncthreads();
    const uint32_t max_depth = s_max_depth;

    /* 6. boxes, deepest level first */
    for (int level = (int)max_depth; level >= 0; level--) {
        for (uint32_t o = tid; o < count; o += LBVH_THREADS) {
            if (work.depth[o] != (uint32_t)level) continue;
            sbox b;
            if (work.kind[o]) b = boxes[leaf_ids_out[work.first[o]]];
            else b = box_merge(work.node_boxes[work.first[o]], work.node_boxes[work.first[o] + 1]);
            work.node_boxes[o] = b;
            uint32_t *w = nodes_out + 8 * o;
            w[0] = __float_as_uint(b.min.x); w[1] = __float_as_uint(b.min.y); w[2] = __float_as_uint(b.min.z);
            w[3] = __float_as_uint(b.max.x); w[4] = __float_as_uint(b.max.y); w[5] = __float_as_uint(b.max.z);
            w[6] = work.first[o];
            w[7] = (work.kind[o] ? 1u : 0u) | (work.depth[o] << 16);
        }
        __syncthreads();
    }
    if (tid == 0) *height_out = max_depth + 1;
}

extern "C" int st_lbvh_launch(hipStream_t stream, const StDPrim *prims, const float *mesh_aabbs, uint32_t mesh_count,
                              uint32_t n, StLbvhWork work, uint32_t *nodes_out, uint32_t *leaf_ids_out,
                              uint32_t *height_out) {
    if (n == 0 || n > ST_LBVH_MAX) return -1;
    hipLaunchKernelGGL(k_lbvh_small, dim3(1), dim3(LBVH_THREADS), 0, stream, prims, mesh_aabbs, mesh_count, n, work,
                       nodes_out, leaf_ids_out, height_out);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
