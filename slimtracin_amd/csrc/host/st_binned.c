/*
 * st_binned.c - binned SAH builder for large meshes (host).
 *
 * The reference builds every mesh BVH with its exact top-down SAH builder
 * (SlimTracin render/acceleration_structures/builder_top_down.h:204-364):
 * three full sorts per node, minutes per multi-million-triangle mesh. This
 * builder evaluates the same cost, SA_left * n_left + SA_right * n_right
 * (getSurfaceAreaOfAABB, half areas), at 31 candidate planes per axis over
 * 32 equal-width centroid bins, so a node costs a linear pass per axis and
 * a 5M-triangle mesh takes seconds. It has no counterpart in the reference:
 * its trees are what traversal speed depends on (fewer node steps per ray
 * than the Morton LBVH), and every field it writes follows the reference's
 * BVH layout (child pairs adjacent, leaves of up to max_leaf boxes as ranges
 * of leaf_ids), so the traversal kernels and the oracle take its trees like
 * any other.
 *
 * The tree is a function of the boxes alone; the oracle restates it and the
 * tests compare the two trees field for field:
 *   centroid c = (min + max) * 0.5 per axis;
 *   a range of n <= max_leaf boxes is a leaf (first_child_id = the range's
 *   start in leaf_ids, child_count = n);
 *   otherwise, per axis with centroid extent e = cmax - cmin > 0: bin of a
 *   box = (int)((c - cmin) * (32 / e)), clamped to 31; planes k = 1..31
 *   with boxes on both sides; cost from the bins' merged boxes; the
 *   smallest cost wins (strict <: first axis, then first plane on ties);
 *   the range is partitioned stably (bin < k first, each side in range
 *   order); with no axis of positive extent the range is halved at
 *   start + n / 2 in range order;
 *   the two children get the next free pair of node slots, the left child
 *   is built first (depth first), node boxes merge the range's boxes in
 *   range order (min/max by `<`/`>` ternaries), node depth counts from 0
 *   at the root, and the height is the largest depth + 1.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "st_host.h"
#include "../st_mesh.h"
#include "slimtracin.h"

#define BINS 32
/* Mesh trees get one triangle per leaf. Leaf caps 1-4 and 8 timed on the
   headline batch (256 requests over 8 scenes of 0.2-5M triangles, one
   launch): 26.3, 26.7, 28.2, 30.5 and 38.8 ms per step. A leaf's triangle
   records are loaded and tested one after another, each test waiting for
   its own 64-byte record, while a child pair's two boxes come in one
   load; so fewer triangle tests per ray (0.68 against 1.73 per ray on the
   200k-triangle scene) outweigh the extra node steps (8.61 against 7.69). */
#define MESH_LEAF 1

typedef struct {
    uint32_t start, end, node, depth;
} Task;

static void set_box(StBVHNodeRec *n, sbox b) {
    n->min[0] = b.min.x; n->min[1] = b.min.y; n->min[2] = b.min.z;
    n->max[0] = b.max.x; n->max[1] = b.max.y; n->max[2] = b.max.z;
}

static sbox range_box(const sbox *boxes, const uint32_t *ids, uint32_t s, uint32_t e) {
    sbox b = boxes[ids[s]];
    for (uint32_t i = s + 1; i < e; i++) b = box_merge(b, boxes[ids[i]]);
    return b;
}

static inline uint32_t bin_of(float c, float cmin, float scale) {
    int b = (int)((c - cmin) * scale);
    return b > BINS - 1 ? BINS - 1 : (uint32_t)b;
}

uint32_t st_binned_build(const sbox *boxes, uint32_t n, uint32_t max_leaf, StBVHNodeRec *nodes, uint32_t *leaf_ids,
                         uint32_t *height) {
    if (!n) { *height = 0; return 0; }
    if (max_leaf < 1) max_leaf = 1;
    float *cen = (float *)malloc(3 * sizeof(float) * (size_t)n);
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    /* a depth-first walk holds at most one pending right child per level;
       levels are bounded by n (each split leaves both sides non-empty) */
    Task *stack = (Task *)malloc(sizeof(Task) * ((size_t)n + 1));
    if (!cen || !tmp || !stack) { free(cen); free(tmp); free(stack); *height = 0; return 0; }
    for (uint32_t i = 0; i < n; i++) {
        cen[3 * (size_t)i + 0] = (boxes[i].min.x + boxes[i].max.x) * 0.5f;
        cen[3 * (size_t)i + 1] = (boxes[i].min.y + boxes[i].max.y) * 0.5f;
        cen[3 * (size_t)i + 2] = (boxes[i].min.z + boxes[i].max.z) * 0.5f;
        leaf_ids[i] = i;
    }
    memset(nodes, 0, sizeof *nodes);
    set_box(nodes, range_box(boxes, leaf_ids, 0, n));
    uint32_t node_count = 1, max_depth = 0, top = 0;
    stack[top++] = (Task){0, n, 0, 0};
    while (top) {
        Task t = stack[--top];
        StBVHNodeRec *node = nodes + t.node;
        uint32_t cnt = t.end - t.start;
        node->depth = (uint16_t)t.depth;
        if (t.depth > max_depth) max_depth = t.depth;
        if (cnt <= max_leaf) {
            node->child_count = (uint16_t)cnt;
            node->first_child_id = t.start;
            continue;
        }
        float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = t.start; i < t.end; i++) {
            const float *c = cen + 3 * (size_t)leaf_ids[i];
            for (int a = 0; a < 3; a++) {
                cmin[a] = c[a] < cmin[a] ? c[a] : cmin[a];
                cmax[a] = c[a] > cmax[a] ? c[a] : cmax[a];
            }
        }
        float best = INFINITY, best_scale = 0.0f;
        int best_axis = -1;
        uint32_t best_plane = 0;
        for (int a = 0; a < 3; a++) {
            float ext = cmax[a] - cmin[a];
            if (!(ext > 0.0f)) continue;
            float scale = (float)BINS / ext;
            sbox bb[BINS];
            uint32_t bc[BINS];
            memset(bc, 0, sizeof bc);
            for (uint32_t i = t.start; i < t.end; i++) {
                uint32_t id = leaf_ids[i];
                uint32_t k = bin_of(cen[3 * (size_t)id + a], cmin[a], scale);
                bb[k] = bc[k] ? box_merge(bb[k], boxes[id]) : boxes[id];
                bc[k]++;
            }
            /* right side areas and counts for planes 1..BINS-1, from the top */
            float r_area[BINS];
            uint32_t r_cnt[BINS];
            sbox R = boxes[0];   /* set from the first non-empty bin */
            uint32_t rn = 0;
            for (int k = BINS - 1; k >= 1; k--) {
                if (bc[k]) { R = rn ? box_merge(R, bb[k]) : bb[k]; rn += bc[k]; }
                r_cnt[k] = rn;
                r_area[k] = rn ? box_area(R) : 0.0f;
            }
            sbox L = boxes[0];
            uint32_t ln = 0;
            for (int k = 1; k < BINS; k++) {
                if (bc[k - 1]) { L = ln ? box_merge(L, bb[k - 1]) : bb[k - 1]; ln += bc[k - 1]; }
                if (!ln || !r_cnt[k]) continue;
                float cost = box_area(L) * (float)ln + r_area[k] * (float)r_cnt[k];
                if (cost < best) {
                    best = cost;
                    best_axis = a;
                    best_plane = (uint32_t)k;
                    best_scale = scale;
                }
            }
        }
        uint32_t mid;
        if (best_axis < 0) {
            mid = t.start + cnt / 2;
        } else {
            uint32_t l = t.start, r = 0;
            for (uint32_t i = t.start; i < t.end; i++) {
                uint32_t id = leaf_ids[i];
                if (bin_of(cen[3 * (size_t)id + best_axis], cmin[best_axis], best_scale) < best_plane)
                    leaf_ids[l++] = id;
                else
                    tmp[r++] = id;
            }
            memcpy(leaf_ids + l, tmp, sizeof(uint32_t) * r);
            mid = l;
        }
        uint32_t lc = node_count;
        node_count += 2;
        node->first_child_id = lc;
        node->child_count = 0;
        StBVHNodeRec *left = nodes + lc, *right = nodes + lc + 1;
        memset(left, 0, 2 * sizeof *left);
        set_box(left, range_box(boxes, leaf_ids, t.start, mid));
        set_box(right, range_box(boxes, leaf_ids, mid, t.end));
        stack[top++] = (Task){mid, t.end, lc + 1, t.depth + 1};
        stack[top++] = (Task){t.start, mid, lc, t.depth + 1};
    }
    free(cen);
    free(tmp);
    free(stack);
    *height = max_depth + 1;
    return node_count;
}

int st_build_mesh_binned(uint32_t T, const float *pos, const uint32_t *pidx, const float *nrm, const uint32_t *nidx,
                         uint32_t normals_count, const float *uvs, const uint32_t *uidx, uint32_t uvs_count,
                         StBVHNodeRec *nodes, uint32_t *node_count, uint32_t *height, StTriangleRec *tris,
                         uint32_t *leaf_ids) {
    sbox *boxes = (sbox *)calloc((size_t)T + 1, sizeof(sbox));
    uint32_t *order = (uint32_t *)malloc(((size_t)T + 1) * sizeof(uint32_t));
    if (!boxes || !order) { free(boxes); free(order); return -1; }
    for (uint32_t i = 0; i < T; i++)
        boxes[i] = st_tri_box(pos + 3 * (size_t)pidx[3 * (size_t)i], pos + 3 * (size_t)pidx[3 * (size_t)i + 1],
                              pos + 3 * (size_t)pidx[3 * (size_t)i + 2]);
    *node_count = st_binned_build(boxes, T, MESH_LEAF, nodes, order, height);
    int ok = *node_count != 0;
    for (uint32_t i = 0; ok && i < T; i++) {
        uint32_t tid = order[i];
        leaf_ids[tid] = i;
        st_tri_record(tid, pos, pidx, normals_count ? nrm : NULL, nidx, uvs_count ? uvs : NULL, uidx, tris + i);
    }
    free(boxes);
    free(order);
    return ok ? 0 : -1;
}

uint32_t st_binned_build_boxes(const float *boxes, uint32_t n, uint32_t max_leaf, StBVHNodeRec *nodes,
                               uint32_t *leaf_ids, uint32_t *height) {
    sbox *b = (sbox *)malloc(((size_t)n + 1) * sizeof(sbox));
    if (!b) return 0;
    for (uint32_t i = 0; i < n; i++) {
        b[i].min = V3(boxes[6 * (size_t)i], boxes[6 * (size_t)i + 1], boxes[6 * (size_t)i + 2]);
        b[i].max = V3(boxes[6 * (size_t)i + 3], boxes[6 * (size_t)i + 4], boxes[6 * (size_t)i + 5]);
    }
    uint32_t c = st_binned_build(b, n, max_leaf, nodes, leaf_ids, height);
    free(b);
    return c;
}
