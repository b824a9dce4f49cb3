/*
 * st_mesh_lbvh.hip - mesh BVH build on the device for large meshes.
 *
 * The reference builds each mesh BVH once on the CPU with its SAH builder
 * (updateMeshBVH, builder_top_down.h:272-364; in practice offline through
 * obj2mesh). For meshes of millions of triangles that takes minutes, so a
 * scene can instead have its meshes built here as a Morton LBVH in the
 * reference's node layout:
 *   1. triangle leaf boxes exactly as updateMeshBVH pads them, centroids
 *      (min+max)*0.5 (+0 folds -0), per-block centroid bounds;
 *   2. 30-bit Morton codes of the centroids in the centroid bounds
 *      (10 bits per axis, x highest);
 *   3. a stable device radix sort of (code, triangle index) pairs, so equal
 *      codes keep triangle order (st_sort.hip: LDS ranking, 8-bit digits);
 *   4. the binary radix tree over the sorted keys made unique by their sorted
 *      position (Karras 2012), one lane per internal node;
 *   5. boxes bottom-up: each leaf walks to the root and the second child to
 *      arrive at a node merges merge(left, right) (atomic arrival counters);
 *   6. output: node 0 = root, children of radix node k at slots 2k+1 and
 *      2k+2; a radix node covering at most max_leaf triangles becomes one
 *      leaf (child_count = its size, first_child_id = its first sorted
 *      position) and the slots below it stay unused; depth = tree depth;
 *      meshes get leaves of one triangle, so every radix node is a node;
 * Steps 2-6 also build the scene LBVH over primitive boxes with leaves of
 * one primitive (st_lbvh_tree, called from st_lbvh.hip).
 *   7. Triangle records in sorted order (st_mesh.h, the host builder's own
 *      per-triangle code), split into the device's hot and cold records.
 * The CPU oracle restates the same specification top-down; tests compare
 * nodes, leaf order and triangle records bit for bit.
 */
#include <hip/hip_runtime.h>

#include "../st_mesh.h"
#include "st_mesh_lbvh.h"
#include "st_sort.h"

#define MB_THREADS 256u
/* Triangles per mesh leaf. One: on 40 served scenes of 1-20M triangles the
   plain request rate rose from 780-829 to 965-999 requests/s against
   leaves of up to 4, the served p50 from 1.41-1.44 to 1.11-1.12 ms (a
   leaf's triangles are tested one after another, each on its own record
   load, while a node step loads both children's boxes at once). */
#define MB_MAX_LEAF 1u

__device__ static inline uint32_t mb_expand10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void __launch_bounds__(MB_THREADS) k_mb_boxes(StMeshGeomDev in, sbox *boxes, float *cent,
                                                         float *partials) {
    __shared__ float red[6][MB_THREADS];
    const uint32_t tid = threadIdx.x, i = blockIdx.x * MB_THREADS + tid;
    float c[3] = {INFINITY, INFINITY, INFINITY}, C[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (i < in.T) {
        const float *v1 = in.pos + 3 * (size_t)in.pidx[3 * (size_t)i];
        const float *v2 = in.pos + 3 * (size_t)in.pidx[3 * (size_t)i + 1];
        const float *v3 = in.pos + 3 * (size_t)in.pidx[3 * (size_t)i + 2];
        sbox b = st_tri_box(v1, v2, v3);
        boxes[i] = b;
        float m[3] = {(b.min.x + b.max.x) * 0.5f + 0.0f, (b.min.y + b.max.y) * 0.5f + 0.0f,
                      (b.min.z + b.max.z) * 0.5f + 0.0f};
        for (int a = 0; a < 3; a++) {
            cent[3 * (size_t)i + a] = m[a];
            c[a] = m[a];
            C[a] = m[a];
        }
    }
    for (int a = 0; a < 3; a++) { red[a][tid] = c[a]; red[3 + a][tid] = C[a]; }
    __syncthreads();
    for (uint32_t s = MB_THREADS / 2; s > 0; s >>= 1) {
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

/* bounds[0..2] = centroid min, bounds[3..5] = 1023 / extent (0 if flat) */
__global__ void __launch_bounds__(MB_THREADS) k_mb_bounds(const float *partials, uint32_t nblocks, float *bounds) {
    __shared__ float red[6][MB_THREADS];
    const uint32_t tid = threadIdx.x;
    float c[3] = {INFINITY, INFINITY, INFINITY}, C[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = tid; b < nblocks; b += MB_THREADS)
        for (int a = 0; a < 3; a++) {
            float lo = partials[6 * (size_t)b + a], hi = partials[6 * (size_t)b + 3 + a];
            c[a] = lo < c[a] ? lo : c[a];
            C[a] = hi > C[a] ? hi : C[a];
        }
    for (int a = 0; a < 3; a++) { red[a][tid] = c[a]; red[3 + a][tid] = C[a]; }
    __syncthreads();
    for (uint32_t s = MB_THREADS / 2; s > 0; s >>= 1) {
        if (tid < s)
            for (int a = 0; a < 3; a++) {
                float o = red[a][tid + s];
                red[a][tid] = o < red[a][tid] ? o : red[a][tid];
                o = red[3 + a][tid + s];
                red[3 + a][tid] = o > red[3 + a][tid] ? o : red[3 + a][tid];
            }
        __syncthreads();
    }
    if (tid < 3) {
        float e = red[3 + tid][0] - red[tid][0];
        bounds[tid] = red[tid][0];
        bounds[3 + tid] = e > 0 ? 1023.0f / e : 0.0f;
    }
}

__global__ void __launch_bounds__(MB_THREADS) k_mb_codes(const float *cent, const float *bounds, uint32_t T,
                                                         uint32_t *keys, uint32_t *vals) {
    const uint32_t i = blockIdx.x * MB_THREADS + threadIdx.x;
    if (i >= T) return;
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        float t = (cent[3 * (size_t)i + a] - bounds[a]) * bounds[3 + a];
        t = t < 0 ? 0 : t;
        t = t > 1023.0f ? 1023.0f : t;
        q[a] = (uint32_t)t;
    }
    keys[i] = (mb_expand10(q[0]) << 2) | (mb_expand10(q[1]) << 1) | mb_expand10(q[2]);
    vals[i] = i;
}

/* common-prefix length of the keys at sorted positions i and j made unique
   by their position; -1 outside [0, n) */
__device__ static inline int mb_lcp(const uint32_t *codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = codes[i], b = codes[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

__global__ void __launch_bounds__(MB_THREADS) k_mb_karras(const uint32_t *codes, uint32_t T, StMeshLbvhWork w) {
    const int N = (int)T;
    const int k = (int)(blockIdx.x * MB_THREADS + threadIdx.x);
    if (k >= N - 1) return;
    int d = (mb_lcp(codes, N, k, k + 1) - mb_lcp(codes, N, k, k - 1)) >= 0 ? 1 : -1;
    int dmin = mb_lcp(codes, N, k, k - d);
    int lmax = 2;
    while (mb_lcp(codes, N, k, k + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (mb_lcp(codes, N, k, k + (l + t) * d) > dmin) l += t;
    int j = k + l * d;
    int dnode = mb_lcp(codes, N, k, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (mb_lcp(codes, N, k, k + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int g = k + s * d + (d < 0 ? -1 : 0);
    int first = k < j ? k : j, last = k < j ? j : k;
    w.ifirst[k] = (uint32_t)first;
    w.ilast[k] = (uint32_t)last;
    w.isplit[k] = (uint32_t)g;
    w.flags[k] = 0;
    if (first == g) w.leaf_parent[g] = (uint32_t)k;
    else w.iparent[g] = (uint32_t)k;
    if (g + 1 == last) w.leaf_parent[g + 1] = (uint32_t)k;
    else w.iparent[g + 1] = (uint32_t)k;
    if (k == 0) w.iparent[0] = 0xFFFFFFFFu;
}

__device__ static inline sbox mb_box_load(const float *p) {
    /* written by another lane of another workgroup: agent-scope loads */
    float v[6];
    for (int a = 0; a < 6; a++)
        v[a] = __uint_as_float(
            __hip_atomic_load(reinterpret_cast<const uint32_t *>(p) + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    sbox b;
    b.min = V3(v[0], v[1], v[2]);
    b.max = V3(v[3], v[4], v[5]);
    return b;
}

__global__ void __launch_bounds__(MB_THREADS) k_mb_refit(const sbox *boxes, const uint32_t *order, uint32_t T,
                                                         StMeshLbvhWork w) {
    const uint32_t i = blockIdx.x * MB_THREADS + threadIdx.x;
    if (i >= T) return;
    uint32_t p = w.leaf_parent[i];
    /* each step climbs one level; a node is merged by exactly one lane (the
       second of its two children to arrive), so every lane stops within the
       tree height */
    for (uint32_t guard = 0; guard < T; guard++) {
        __threadfence();
        if (atomicAdd(&w.flags[p], 1u) == 0u) return;
        __threadfence();
        const uint32_t g = w.isplit[p];
        const sbox l = (w.ifirst[p] == g) ? boxes[order[g]] : mb_box_load(w.ibox + 6 * (size_t)g);
        const sbox r = (g + 1 == w.ilast[p]) ? boxes[order[g + 1]] : mb_box_load(w.ibox + 6 * (size_t)(g + 1));
        const sbox b = box_merge(l, r);
        float *o = w.ibox + 6 * (size_t)p;
        const float v[6] = {b.min.x, b.min.y, b.min.z, b.max.x, b.max.y, b.max.z};
        for (int a = 0; a < 6; a++)
            __hip_atomic_store(reinterpret_cast<uint32_t *>(o) + a, __float_as_uint(v[a]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        if (p == 0) return;
        p = w.iparent[p];
    }
}

__device__ static inline void mb_write_node(uint32_t *nodes, uint32_t slot, sbox b, uint32_t first, uint32_t count,
                                            uint32_t depth) {
    uint32_t *o = nodes + 8 * (size_t)slot;
    o[0] = __float_as_uint(b.min.x); o[1] = __float_as_uint(b.min.y); o[2] = __float_as_uint(b.min.z);
    o[3] = __float_as_uint(b.max.x); o[4] = __float_as_uint(b.max.y); o[5] = __float_as_uint(b.max.z);
    o[6] = first;
    o[7] = count | (depth << 16);
}

__global__ void __launch_bounds__(MB_THREADS) k_mb_emit(const sbox *boxes, const uint32_t *order, uint32_t T,
                                                        StMeshLbvhWork w, uint32_t *nodes, uint32_t *height,
                                                        uint32_t max_leaf) {
    const uint32_t k = blockIdx.x * MB_THREADS + threadIdx.x;
    if (k + 1 >= T) return;
    /* depth = ancestors; not emitted when an ancestor became one leaf */
    uint32_t dep = 0, x = k;
    for (uint32_t guard = 0; x != 0 && guard < T; guard++) {
        const uint32_t p = w.iparent[x];
        if (w.ilast[p] - w.ifirst[p] + 1u <= max_leaf) return;
        dep++;
        x = p;
    }
    uint32_t slot = 0;
    if (k != 0) {
        const uint32_t p = w.iparent[k];
        slot = (k == w.isplit[p]) ? 2u * p + 1u : 2u * p + 2u;
    }
    const sbox b = mb_box_load(w.ibox + 6 * (size_t)k);
    const uint32_t size = w.ilast[k] - w.ifirst[k] + 1u;
    if (size <= max_leaf) {
        mb_write_node(nodes, slot, b, w.ifirst[k], size, dep);
        atomicMax(height, dep + 1u);
        return;
    }
    mb_write_node(nodes, slot, b, 2u * k + 1u, 0u, dep);
    const uint32_t g = w.isplit[k];
    uint32_t deepest = dep + 1u;
    if (w.ifirst[k] == g) {
        mb_write_node(nodes, 2u * k + 1u, boxes[order[g]], g, 1u, dep + 1u);
        deepest = dep + 2u;
    }
    if (g + 1u == w.ilast[k]) {
        mb_write_node(nodes, 2u * k + 2u, boxes[order[g + 1u]], g + 1u, 1u, dep + 1u);
        deepest = dep + 2u;
    }
    atomicMax(height, deepest);
}

__global__ void __launch_bounds__(MB_THREADS) k_mb_single(const sbox *boxes, uint32_t *nodes, uint32_t *height) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        mb_write_node(nodes, 0, boxes[0], 0u, 1u, 0u);
        *height = 1u;
    }
}

__global__ void __launch_bounds__(MB_THREADS) k_mb_tris(StMeshGeomDev in, const uint32_t *order, StDTriHot *hot,
                                                        StDTriCold *cold) {
    const uint32_t j = blockIdx.x * MB_THREADS + threadIdx.x;
    if (j >= in.T) return;
    StTriangleRec t;
    st_tri_record(order[j], in.pos, in.pidx, in.nrm, in.nidx, in.uvs, in.uidx, &t);
    StDTriHot h;
    StDTriCold c;
    st_tri_split(&t, &h, &c);
    hot[j] = h;
    cold[j] = c;
}

__global__ void k_mb_zero(uint32_t *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

/* ---- host side ------------------------------------------------------------ */

static size_t r256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t st_lbvh_tree_work_bytes(uint32_t T) {
    return 2 * r256(4ull * T) + 6 * r256(4ull * T) + r256(24ull * T) + r256(st_radix_sort_scratch_bytes(T)) + 256;
}

int st_lbvh_tree(hipStream_t stream, const sbox *boxes, const float *cent, const float *partials, uint32_t nb,
                 uint32_t T, uint32_t max_leaf, char *work, uint32_t *keys, uint32_t *order, uint32_t *nodes,
                 uint32_t *height_dev) {
    const uint32_t nbt = (T + MB_THREADS - 1) / MB_THREADS;
    char *p = work;
    uint32_t *keys_tmp = (uint32_t *)p; p += r256(4ull * T);
    uint32_t *vals_tmp = (uint32_t *)p; p += r256(4ull * T);
    StMeshLbvhWork w;
    w.ifirst = (uint32_t *)p; p += r256(4ull * T);
    w.ilast = (uint32_t *)p; p += r256(4ull * T);
    w.isplit = (uint32_t *)p; p += r256(4ull * T);
    w.iparent = (uint32_t *)p; p += r256(4ull * T);
    w.leaf_parent = (uint32_t *)p; p += r256(4ull * T);
    w.flags = (uint32_t *)p; p += r256(4ull * T);
    w.ibox = (float *)p; p += r256(24ull * T);
    void *sort_tmp = p;
    float *bounds = (float *)((char *)sort_tmp + r256(st_radix_sort_scratch_bytes(T)));
    hipLaunchKernelGGL(k_mb_zero, dim3(1), dim3(64), 0, stream, height_dev, (size_t)1);
    if (T == 1) {
        hipLaunchKernelGGL(k_mb_single, dim3(1), dim3(64), 0, stream, boxes, nodes, height_dev);
        hipLaunchKernelGGL(k_mb_zero, dim3(1), dim3(64), 0, stream, order, (size_t)1);
        return hipGetLastError() == hipSuccess ? 0 : -2;
    }
    hipLaunchKernelGGL(k_mb_bounds, dim3(1), dim3(MB_THREADS), 0, stream, partials, nb, bounds);
    hipLaunchKernelGGL(k_mb_codes, dim3(nbt), dim3(MB_THREADS), 0, stream, cent, (const float *)bounds, T, keys,
                       order);
    if (st_radix_sort_pairs(stream, keys, order, keys_tmp, vals_tmp, T, 30, sort_tmp)) return -3;
    hipLaunchKernelGGL(k_mb_karras, dim3(nbt), dim3(MB_THREADS), 0, stream, (const uint32_t *)keys, T, w);
    hipLaunchKernelGGL(k_mb_refit, dim3(nbt), dim3(MB_THREADS), 0, stream, boxes, (const uint32_t *)order, T, w);
    hipLaunchKernelGGL(k_mb_emit, dim3(nbt), dim3(MB_THREADS), 0, stream, boxes, (const uint32_t *)order, T, w, nodes,
                       height_dev, max_leaf);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

size_t st_mesh_lbvh_work_bytes(uint32_t T, size_t *sort_bytes) {
    *sort_bytes = 0;
    const size_t nb = (T + MB_THREADS - 1) / MB_THREADS;
    return r256(24ull * T) + r256(12ull * T) + r256(24ull * nb) + 2 * r256(4ull * T) + st_lbvh_tree_work_bytes(T) +
           256;
}

int st_mesh_lbvh_build(hipStream_t stream, StMeshGeomDev in, char *work, size_t sort_bytes, uint32_t *nodes,
                       StDTriHot *hot, StDTriCold *cold, uint32_t *height_dev) {
    (void)sort_bytes;
    const uint32_t T = in.T;
    if (T == 0) return -1;
    const uint32_t nb = (T + MB_THREADS - 1) / MB_THREADS;
    char *p = work;
    sbox *boxes = (sbox *)p; p += r256(24ull * T);
    float *cent = (float *)p; p += r256(12ull * T);
    float *partials = (float *)p; p += r256(24ull * nb);
    uint32_t *keys = (uint32_t *)p; p += r256(4ull * T);
    uint32_t *order = (uint32_t *)p; p += r256(4ull * T);
    hipLaunchKernelGGL(k_mb_boxes, dim3(nb), dim3(MB_THREADS), 0, stream, in, boxes, cent, partials);
    if (hipGetLastError() != hipSuccess) return -2;
    const int rc = st_lbvh_tree(stream, boxes, cent, partials, nb, T, MB_MAX_LEAF, p, keys, order, nodes, height_dev);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mb_tris, dim3(nb), dim3(MB_THREADS), 0, stream, in, (const uint32_t *)order, hot, cold);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
