/*
 * st_sah.c - the reference's top-down SAH BVH builder with a fast sort.
 *
 * Node topology, leaf order and every float of the output equal the
 * reference builder's (SlimTracin render/acceleration_structures/
 * builder_top_down.h:66-270): per node and per axis the primitive ids are
 * ordered by the reference's quicksort (Lomuto partition, pivot = last
 * element, key = aabb.max[axis], strict `<`), prefix and suffix boxes are
 * merged in the same order, the split with the smallest
 * SA_left * n_left + SA_right * n_right wins (first minimum, first axis on
 * ties), and nodes are numbered in the reference's depth-first order.
 *
 * The reference sort is quadratic on the inputs it sees most: every child
 * range arrives sorted on its parent's split axis. The result of that
 * quicksort on a non-decreasing range is known in closed form: the pivot is
 * the range's maximum, so each partition step leaves the prefix in place and
 * swaps the ends of the trailing run of equal keys, and the run itself then
 * recurses as an all-equal range. Net effect: every maximal run of equal keys
 * is rotated right by one element and nothing else moves. The sort below
 * checks each popped range for that case (a linear scan that stops at the
 * first descent) and applies the rotation instead of partitioning, which
 * yields the identical permutation in linear time. Tests compare this
 * builder with a plain restatement of the reference quicksort on random,
 * sorted and duplicate-heavy inputs.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "st_host.h"
#include "../st_mesh.h"
#include "slimtracin.h"

typedef struct {
    const sbox *boxes;
    float *key[3];
    uint32_t *ids;
    uint32_t *sorted[3];
    float *lsa, *rsa;
    int32_t *stack;
    uint32_t *it_start, *it_end, *it_node;
    uint32_t left_count[3];
    float smallest[3];
} Builder;

static void rotate_runs(const float *key, uint32_t *ids, int32_t s, int32_t e) {
    int32_t i = s;
    while (i <= e) {
        int32_t j = i;
        while (j < e && key[ids[j + 1]] == key[ids[i]]) j++;
        if (j > i) {
            uint32_t last = ids[j];
            memmove(ids + i + 1, ids + i, (size_t)(j - i) * sizeof(uint32_t));
            ids[i] = last;
        }
        i = j + 1;
    }
}

/* builder_top_down.h:66-123 with the sorted-range shortcut described above. */
static void sort_ids(const float *key, uint32_t *ids, int32_t start, int32_t end, int32_t *stack) {
    int32_t top = -1;
    stack[++top] = start;
    stack[++top] = end;
    while (top >= 0) {
        end = stack[top--];
        start = stack[top--];
        int32_t i = start;
        while (i < end && key[ids[i]] <= key[ids[i + 1]]) i++;
        if (i == end) {
            rotate_runs(key, ids, start, end);
            continue;
        }
        int32_t p = start;
        uint32_t pivot = ids[end];
        float pk = key[pivot];
        for (i = start; i < end; i++) {
            if (key[ids[i]] < pk) {
                uint32_t t = ids[i];
                ids[i] = ids[p];
                ids[p] = t;
                p++;
            }
        }
        uint32_t t = ids[end];
        ids[end] = ids[p];
        ids[p] = t;
        if (p - 1 > start) { stack[++top] = start; stack[++top] = p - 1; }
        if (p + 1 < end) { stack[++top] = p + 1; stack[++top] = end; }
    }
}

/* builder_top_down.h:125-166 */
static void partition_axis(Builder *b, int axis, uint32_t N) {
    uint32_t *sorted = b->sorted[axis];
    sort_ids(b->key[axis], sorted, 0, (int32_t)N - 1, b->stack);
    const sbox *bx = b->boxes;
    sbox L = bx[sorted[0]], R = bx[sorted[N - 1]];
    b->lsa[0] = box_area(L);
    b->rsa[N - 1] = box_area(R);
    for (uint32_t li = 1, ri = N - 2; li < N; li++, ri--) {
        L = box_merge(L, bx[sorted[li]]);
        R = box_merge(R, bx[sorted[ri]]);
        b->lsa[li] = box_area(L);
        b->rsa[ri] = box_area(R);
    }
    uint32_t last = N - 1, next = 1, lcount = 1, rcount = N - 1;
    b->left_count[axis] = 1;
    b->smallest[axis] = INFINITY;
    for (uint32_t cur = 0; cur < last; cur++, next++, lcount++, rcount--) {
        float sa = b->lsa[cur] * (float)lcount + b->rsa[next] * (float)rcount;
        if (b->smallest[axis] > sa) {
            b->smallest[axis] = sa;
            b->left_count[axis] = next;
        }
    }
}

static void node_clear(StBVHNodeRec *n) { memset(n, 0, sizeof *n); }

static void node_set_box(StBVHNodeRec *n, sbox b) {
    n->min[0] = b.min.x; n->min[1] = b.min.y; n->min[2] = b.min.z;
    n->max[0] = b.max.x; n->max[1] = b.max.y; n->max[2] = b.max.z;
}
static sbox node_box(const StBVHNodeRec *n) {
    sbox b;
    b.min = V3(n->min[0], n->min[1], n->min[2]);
    b.max = V3(n->max[0], n->max[1], n->max[2]);
    return b;
}

/* builder_top_down.h:168-202 */
static uint32_t split_node(Builder *b, StBVHNodeRec *nodes, uint32_t *node_count, StBVHNodeRec *node, uint32_t start,
                           uint32_t end) {
    uint32_t N = end - start;
    uint32_t *ids = b->ids + start;
    node->first_child_id = *node_count;
    StBVHNodeRec *left = nodes + (*node_count)++;
    StBVHNodeRec *right = nodes + (*node_count)++;
    node_clear(left);
    node_clear(right);
    float smallest = INFINITY;
    int chosen = 0;
    for (int axis = 0; axis < 3; axis++) {
        memcpy(b->sorted[axis], ids, N * sizeof(uint32_t));
        partition_axis(b, axis, N);
        if (b->smallest[axis] < smallest) { smallest = b->smallest[axis]; chosen = axis; }
    }
    uint32_t lc = b->left_count[chosen];
    const uint32_t *s = b->sorted[chosen];
    /* the chosen axis' left.aabbs[lc-1] and right.aabbs[lc], merged in the
       reference's order */
    sbox L = b->boxes[s[0]];
    for (uint32_t i = 1; i < lc; i++) L = box_merge(L, b->boxes[s[i]]);
    sbox R = b->boxes[s[N - 1]];
    for (uint32_t i = N - 1; i-- > lc;) R = box_merge(R, b->boxes[s[i]]);
    node_set_box(left, L);
    node_set_box(right, R);
    memcpy(ids, s, N * sizeof(uint32_t));
    return start + lc;
}

uint32_t st_sah_build(const sbox *boxes, uint32_t n, uint32_t max_leaf, StBVHNodeRec *nodes, uint32_t *leaf_ids,
                      uint32_t *height) {
    node_clear(nodes);
    if (n <= max_leaf) { /* builder_top_down.h:208-222 */
        sbox root;
        root.min = V3s(INFINITY);
        root.max = V3s(-INFINITY);
        for (uint32_t i = 0; i < n; i++) {
            leaf_ids[i] = i;
            root = box_merge(root, boxes[i]);
        }
        node_set_box(nodes, root);
        nodes->child_count = (uint16_t)n;
        nodes->first_child_id = 0;
        nodes->depth = 0;
        *height = 1;
        return 1;
    }
    Builder b;
    memset(&b, 0, sizeof b);
    size_t m = (size_t)n + 2;
    b.boxes = boxes;
    for (int a = 0; a < 3; a++) {
        b.key[a] = (float *)malloc(m * sizeof(float));
        b.sorted[a] = (uint32_t *)malloc(m * sizeof(uint32_t));
    }
    for (uint32_t i = 0; i < n; i++) {
        b.key[0][i] = boxes[i].max.x;
        b.key[1][i] = boxes[i].max.y;
        b.key[2][i] = boxes[i].max.z;
    }
    b.ids = (uint32_t *)malloc(m * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; i++) b.ids[i] = i;
    b.lsa = (float *)malloc(m * sizeof(float));
    b.rsa = (float *)malloc(m * sizeof(float));
    b.stack = (int32_t *)malloc((2 * m + 4) * sizeof(int32_t));
    b.it_start = (uint32_t *)malloc(m * sizeof(uint32_t));
    b.it_end = (uint32_t *)malloc(m * sizeof(uint32_t));
    b.it_node = (uint32_t *)malloc(m * sizeof(uint32_t));

    uint32_t node_count = 1;
    uint32_t middle = split_node(&b, nodes, &node_count, nodes, 0, n);
    b.it_start[0] = 0; b.it_end[0] = middle; b.it_node[0] = 1;
    b.it_start[1] = middle; b.it_end[1] = n; b.it_node[1] = 2;
    int32_t depth = 1;
    uint32_t leaf_count = 0;
    *height = 1;
    while (depth >= 0) {
        uint32_t ls = b.it_start[depth], le = b.it_end[depth];
        StBVHNodeRec *node = nodes + b.it_node[depth];
        uint32_t cnt = le - ls;
        if (cnt <= max_leaf) {
            node->depth = (uint16_t)depth;
            node->child_count = (uint16_t)cnt;
            node->first_child_id = leaf_count;
            for (uint32_t i = 0; i < cnt; i++) leaf_ids[leaf_count + i] = b.ids[ls + i];
            leaf_count += cnt;
            depth--;
        } else {
            middle = split_node(&b, nodes, &node_count, node, ls, le);
            nodes[node_count - 1].depth = nodes[node_count - 2].depth = (uint16_t)depth;
            b.it_start[depth] = ls; b.it_end[depth] = middle; b.it_node[depth] = node->first_child_id;
            depth++;
            b.it_start[depth] = middle; b.it_end[depth] = le; b.it_node[depth] = node->first_child_id + 1;
            if (depth > (int32_t)*height) *height = (uint32_t)depth;
        }
    }
    sbox root = box_merge(node_box(nodes + 1), node_box(nodes + 2));
    node_set_box(nodes, root);

    for (int a = 0; a < 3; a++) { free(b.key[a]); free(b.sorted[a]); }
    free(b.ids); free(b.lsa); free(b.rsa); free(b.stack);
    free(b.it_start); free(b.it_end); free(b.it_node);
    return node_count;
}

/* builder_top_down.h:272-364: padded triangle boxes, leaf size 4, triangle
   records in leaf order, leaf_ids as the inverse permutation. */
int st_build_mesh(uint32_t T, const float *pos, const uint32_t *pidx, const float *nrm, const uint32_t *nidx,
                  uint32_t normals_count, const float *uvs, const uint32_t *uidx, uint32_t uvs_count,
                  StBVHNodeRec *nodes, uint32_t *node_count, uint32_t *height, StTriangleRec *tris,
                  uint32_t *leaf_ids) {
    sbox *boxes = (sbox *)calloc((size_t)T + 1, sizeof(sbox));
    uint32_t *order = (uint32_t *)malloc(((size_t)T + 1) * sizeof(uint32_t));
    if (!boxes || !order) { free(boxes); free(order); return -1; }
    for (uint32_t i = 0; i < T; i++)
        boxes[i] = st_tri_box(pos + 3 * (size_t)pidx[3 * (size_t)i], pos + 3 * (size_t)pidx[3 * (size_t)i + 1],
                              pos + 3 * (size_t)pidx[3 * (size_t)i + 2]);
    *node_count = st_sah_build(boxes, T, 4, nodes, order, height);
    for (uint32_t i = 0; i < T; i++) {
        uint32_t tid = order[i];
        leaf_ids[tid] = i;
        st_tri_record(tid, pos, pidx, normals_count ? nrm : NULL, nidx, uvs_count ? uvs : NULL, uidx, tris + i);
    }
    free(boxes);
    free(order);
    return 0;
}

uint32_t st_sah_build_boxes(const float *boxes, uint32_t n, uint32_t max_leaf, StBVHNodeRec *nodes, uint32_t *leaf_ids,
                            uint32_t *height) {
    sbox *b = (sbox *)malloc(((size_t)n + 1) * sizeof(sbox));
    if (!b) return 0;
    for (uint32_t i = 0; i < n; i++) {
        b[i].min = V3(boxes[6 * (size_t)i], boxes[6 * (size_t)i + 1], boxes[6 * (size_t)i + 2]);
        b[i].max = V3(boxes[6 * (size_t)i + 3], boxes[6 * (size_t)i + 4], boxes[6 * (size_t)i + 5]);
    }
    uint32_t c = st_sah_build(b, n, max_leaf, nodes, leaf_ids, height);
    free(b);
    return c;
}
