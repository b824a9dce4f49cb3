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
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

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

typedef struct {
    const sbox *boxes;
    const float *cen;   /* centroids, 3 per box */
    uint32_t *ids;      /* leaf_ids: permuted in place, range by range */
    uint32_t *tmp;      /* partition scratch, used at the range's own offsets */
    uint32_t max_leaf;
} Ctx;

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

/* Split the range [s, e) of more than max_leaf boxes: best binned plane (or
   the halving), stable partition in place; returns the split point and the
   two sides' boxes. Touches only ids[s, e) and tmp[s, e). */
static uint32_t split_range(const Ctx *c, uint32_t s, uint32_t e, sbox *lbox, sbox *rbox) {
    const sbox *boxes = c->boxes;
    const float *cen = c->cen;
    uint32_t *ids = c->ids;
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = s; i < e; i++) {
        const float *p = cen + 3 * (size_t)ids[i];
        for (int a = 0; a < 3; a++) {
            cmin[a] = p[a] < cmin[a] ? p[a] : cmin[a];
            cmax[a] = p[a] > cmax[a] ? p[a] : cmax[a];
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
        for (uint32_t i = s; i < e; i++) {
            uint32_t id = ids[i];
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
        mid = s + (e - s) / 2;
    } else {
        uint32_t *tmp = c->tmp + s;
        uint32_t l = s, r = 0;
        for (uint32_t i = s; i < e; i++) {
            uint32_t id = ids[i];
            if (bin_of(cen[3 * (size_t)id + best_axis], cmin[best_axis], best_scale) < best_plane)
                ids[l++] = id;
            else
                tmp[r++] = id;
        }
        memcpy(ids + l, tmp, sizeof(uint32_t) * r);
        mid = l;
    }
    *lbox = range_box(boxes, ids, s, mid);
    *rbox = range_box(boxes, ids, mid, e);
    return mid;
}

/* Depth-first build of the subtree over [s, e) whose root is nodes[0] (its
   box set by the caller): child pairs take slots 1, 3, 5, ... in the order
   a single-threaded build of the whole tree allocates them within this
   subtree, the left subtree first. Returns the slots used. */
static uint32_t build_subtree(const Ctx *c, uint32_t s, uint32_t e, uint32_t depth0, StBVHNodeRec *nodes,
                              Task *stack, uint32_t *max_depth) {
    uint32_t count = 1, top = 0;
    stack[top++] = (Task){s, e, 0, depth0};
    while (top) {
        Task t = stack[--top];
        StBVHNodeRec *node = nodes + t.node;
        uint32_t cnt = t.end - t.start;
        node->depth = (uint16_t)t.depth;
        if (t.depth > *max_depth) *max_depth = t.depth;
        if (cnt <= c->max_leaf) {
            node->child_count = (uint16_t)cnt;
            node->first_child_id = t.start;
            continue;
        }
        sbox lb, rb;
        uint32_t mid = split_range(c, t.start, t.end, &lb, &rb);
        uint32_t lc = count;
        count += 2;
        node->first_child_id = lc;
        node->child_count = 0;
        StBVHNodeRec *left = nodes + lc, *right = nodes + lc + 1;
        memset(left, 0, 2 * sizeof *left);
        set_box(left, lb);
        set_box(right, rb);
        stack[top++] = (Task){mid, t.end, lc + 1, t.depth + 1};
        stack[top++] = (Task){t.start, mid, lc, t.depth + 1};
    }
    return count;
}

/* ---- parallel build ------------------------------------------------------
 * Large inputs: the top of the tree is split on the calling thread down to
 * ranges of at most `frontier` boxes; those subtrees are built by worker
 * threads, each into its own slot block (ranges are disjoint, so their
 * in-place partitions of ids and tmp never meet); a replay of the
 * depth-first slot allocation over the top then gives every pair the slot a
 * single-threaded build gives it, so the output is the same tree, slot for
 * slot. */
#define PAR_MIN_BOXES (1u << 17)
#define PAR_MAX_THREADS 16
enum { TOP_LEAF, TOP_INNER, TOP_FRONTIER };

typedef struct {
    uint32_t s, e, depth, kind, left, work;
    sbox box;
} TopNode;

typedef struct {
    uint32_t s, e, depth, count, max_depth;
    StBVHNodeRec *nodes;
} Work;

typedef struct {
    const Ctx *c;
    Work *work;
    uint32_t nwork;
    uint32_t next;   /* shared claim counter */
    int failed;
} Pool;

static void *pool_run(void *arg) {
    Pool *p = (Pool *)arg;
    for (;;) {
        uint32_t w = __atomic_fetch_add(&p->next, 1u, __ATOMIC_RELAXED);
        if (w >= p->nwork) break;
        Work *k = p->work + w;
        uint32_t n = k->e - k->s;
        k->nodes = (StBVHNodeRec *)calloc(2 * (size_t)n, sizeof(StBVHNodeRec));
        Task *stack = (Task *)malloc(sizeof(Task) * ((size_t)n + 1));
        if (!k->nodes || !stack) {
            free(stack);
            __atomic_store_n(&p->failed, 1, __ATOMIC_RELAXED);
            continue;
        }
        k->max_depth = 0;
        k->count = build_subtree(p->c, k->s, k->e, k->depth, k->nodes, stack, &k->max_depth);
        free(stack);
    }
    return NULL;
}

/* Writes top node u at slot g and everything below it; *next = next free
   slot of the depth-first allocation. */
static void place(const TopNode *top, const Work *work, uint32_t u, uint32_t g, StBVHNodeRec *nodes,
                  uint32_t *next) {
    const TopNode *t = top + u;
    StBVHNodeRec *rec = nodes + g;
    if (t->kind == TOP_FRONTIER) {
        const Work *k = work + t->work;
        const uint32_t base = *next;   /* local slot i >= 1 goes to base + i - 1 */
        *rec = k->nodes[0];
        if (!rec->child_count) rec->first_child_id = base + rec->first_child_id - 1;
        set_box(rec, t->box);
        for (uint32_t i = 1; i < k->count; i++) {
            StBVHNodeRec r = k->nodes[i];
            if (!r.child_count) r.first_child_id = base + r.first_child_id - 1;
            nodes[base + i - 1] = r;
        }
        *next = base + k->count - 1;
        return;
    }
    memset(rec, 0, sizeof *rec);
    set_box(rec, t->box);
    rec->depth = (uint16_t)t->depth;
    if (t->kind == TOP_LEAF) {
        rec->child_count = (uint16_t)(t->e - t->s);
        rec->first_child_id = t->s;
        return;
    }
    const uint32_t gl = *next;
    *next += 2;
    rec->first_child_id = gl;
    place(top, work, t->left, gl, nodes, next);
    place(top, work, t->left + 1, gl + 1, nodes, next);
}

static uint32_t build_parallel(const Ctx *c, uint32_t n, sbox root, StBVHNodeRec *nodes, uint32_t *height) {
    /* subtrees of at most n / 64 boxes: the top's levels are single-threaded
       passes over every box, and 64 or more subtrees keep the threads busy */
    const uint32_t frontier = n / 64u > 4096u ? n / 64u : 4096u;
    /* top slots: 4n / frontier covers balanced splits; chains of uneven
       splits grow the arrays */
    uint32_t cap = 4u * (n / frontier) + 8u;
    TopNode *top = (TopNode *)calloc(cap, sizeof(TopNode));
    uint32_t *stack = (uint32_t *)malloc(sizeof(uint32_t) * cap);
    Work *work = (Work *)calloc(cap, sizeof(Work));
    uint32_t ntop = 1, nwork = 0, sp = 0, max_depth = 0, result = 0;
    if (!top || !stack || !work) goto out;
    top[0] = (TopNode){0, n, 0, TOP_LEAF, 0, 0, root};
    stack[sp++] = 0;
    while (sp) {   /* the same depth-first order as build_subtree */
        const uint32_t u = stack[--sp];
        TopNode *t = top + u;
        const uint32_t cnt = t->e - t->s;
        if (t->depth > max_depth) max_depth = t->depth;
        if (cnt <= c->max_leaf) {
            t->kind = TOP_LEAF;
        } else if (cnt <= frontier) {
            t->kind = TOP_FRONTIER;
            t->work = nwork;
            work[nwork++] = (Work){t->s, t->e, t->depth, 0, 0, NULL};
        } else {
            if (ntop + 2 > cap) {   /* the stack and the work list never hold more than ntop entries */
                const uint32_t ncap = 2u * cap;
                TopNode *t2 = (TopNode *)realloc(top, sizeof(TopNode) * ncap);
                if (t2) top = t2;
                uint32_t *s2 = (uint32_t *)realloc(stack, sizeof(uint32_t) * ncap);
                if (s2) stack = s2;
                Work *w2 = (Work *)realloc(work, sizeof(Work) * ncap);
                if (w2) work = w2;
                if (!t2 || !s2 || !w2) goto out;
                cap = ncap;
                t = top + u;
            }
            sbox lb, rb;
            const uint32_t mid = split_range(c, t->s, t->e, &lb, &rb);
            t->kind = TOP_INNER;
            t->left = ntop;
            top[ntop] = (TopNode){t->s, mid, t->depth + 1, TOP_LEAF, 0, 0, lb};
            top[ntop + 1] = (TopNode){mid, t->e, t->depth + 1, TOP_LEAF, 0, 0, rb};
            stack[sp++] = ntop + 1;
            stack[sp++] = ntop;
            ntop += 2;
        }
    }
    {
        long online = sysconf(_SC_NPROCESSORS_ONLN);
        uint32_t nthreads = online > 1 ? (uint32_t)online : 1u;
        if (nthreads > PAR_MAX_THREADS) nthreads = PAR_MAX_THREADS;
        if (nthreads > nwork) nthreads = nwork;
        Pool pool = {c, work, nwork, 0, 0};
        pthread_t th[PAR_MAX_THREADS];
        uint32_t started = 0;
        for (uint32_t i = 1; i < nthreads; i++)
            if (pthread_create(&th[started], NULL, pool_run, &pool) == 0) started++;
        pool_run(&pool);   /* the calling thread works too */
        for (uint32_t i = 0; i < started; i++) pthread_join(th[i], NULL);
        if (pool.failed) goto out;
    }
    for (uint32_t w = 0; w < nwork; w++)
        if (work[w].max_depth > max_depth) max_depth = work[w].max_depth;
    {
        uint32_t next = 1;
        place(top, work, 0, 0, nodes, &next);
        result = next;
    }
    *height = max_depth + 1;
out:
    for (uint32_t w = 0; work && w < nwork; w++) free(work[w].nodes);
    free(work);
    free(stack);
    free(top);
    return result;
}

uint32_t st_binned_build(const sbox *boxes, uint32_t n, uint32_t max_leaf, StBVHNodeRec *nodes, uint32_t *leaf_ids,
                         uint32_t *height) {
    if (!n) { *height = 0; return 0; }
    if (max_leaf < 1) max_leaf = 1;
    float *cen = (float *)malloc(3 * sizeof(float) * (size_t)n);
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    if (!cen || !tmp) { free(cen); free(tmp); *height = 0; return 0; }
    for (uint32_t i = 0; i < n; i++) {
        cen[3 * (size_t)i + 0] = (boxes[i].min.x + boxes[i].max.x) * 0.5f;
        cen[3 * (size_t)i + 1] = (boxes[i].min.y + boxes[i].max.y) * 0.5f;
        cen[3 * (size_t)i + 2] = (boxes[i].min.z + boxes[i].max.z) * 0.5f;
        leaf_ids[i] = i;
    }
    const Ctx c = {boxes, cen, leaf_ids, tmp, max_leaf};
    const sbox root = range_box(boxes, leaf_ids, 0, n);
    uint32_t count = 0;
    if (n >= PAR_MIN_BOXES) count = build_parallel(&c, n, root, nodes, height);
    if (!count) {
        /* single-threaded (also when the parallel path ran out of memory: it
           starts again from the initial order, which the tree depends on) */
        for (uint32_t i = 0; i < n; i++) leaf_ids[i] = i;
        /* a depth-first walk holds at most one pending right child per
           level; levels are bounded by n (both sides of a split are
           non-empty) */
        Task *stack = (Task *)malloc(sizeof(Task) * ((size_t)n + 1));
        if (stack) {
            uint32_t max_depth = 0;
            memset(nodes, 0, sizeof *nodes);
            set_box(nodes, root);
            count = build_subtree(&c, 0, n, 0, nodes, stack, &max_depth);
            *height = max_depth + 1;
        }
        free(stack);
    }
    free(cen);
    free(tmp);
    if (!count) *height = 0;
    return count;
}

/* Per-index work over [0, n) in chunks on up to PAR_MAX_THREADS threads
   (the calling thread too); single-threaded below PAR_MIN_BOXES. Each index
   writes only its own outputs, so the result does not depend on the split. */
typedef struct {
    void (*fn)(void *, uint32_t, uint32_t);
    void *arg;
    uint32_t n, chunk, next;
} PFor;

static void *pfor_run(void *a) {
    PFor *p = (PFor *)a;
    for (;;) {
        uint32_t b = __atomic_fetch_add(&p->next, p->chunk, __ATOMIC_RELAXED);
        if (b >= p->n) break;
        p->fn(p->arg, b, p->n - b < p->chunk ? p->n : b + p->chunk);
    }
    return NULL;
}

static void parallel_for(uint32_t n, void (*fn)(void *, uint32_t, uint32_t), void *arg) {
    if (n < PAR_MIN_BOXES) { fn(arg, 0, n); return; }
    long online = sysconf(_SC_NPROCESSORS_ONLN);
    uint32_t nthreads = online > 1 ? (uint32_t)online : 1u;
    if (nthreads > PAR_MAX_THREADS) nthreads = PAR_MAX_THREADS;
    PFor p = {fn, arg, n, 1u << 15, 0};
    pthread_t th[PAR_MAX_THREADS];
    uint32_t started = 0;
    for (uint32_t i = 1; i < nthreads; i++)
        if (pthread_create(&th[started], NULL, pfor_run, &p) == 0) started++;
    pfor_run(&p);
    for (uint32_t i = 0; i < started; i++) pthread_join(th[i], NULL);
}

typedef struct {
    const float *pos, *nrm, *uvs;
    const uint32_t *pidx, *nidx, *uidx, *order;
    sbox *boxes;
    StTriangleRec *tris;
    uint32_t *leaf_ids;
} MeshJob;

static void mesh_boxes(void *a, uint32_t b, uint32_t e) {
    const MeshJob *j = (const MeshJob *)a;
    for (uint32_t i = b; i < e; i++)
        j->boxes[i] = st_tri_box(j->pos + 3 * (size_t)j->pidx[3 * (size_t)i],
                                 j->pos + 3 * (size_t)j->pidx[3 * (size_t)i + 1],
                                 j->pos + 3 * (size_t)j->pidx[3 * (size_t)i + 2]);
}

static void mesh_records(void *a, uint32_t b, uint32_t e) {
    const MeshJob *j = (const MeshJob *)a;
    for (uint32_t i = b; i < e; i++) {
        const uint32_t tid = j->order[i];
        j->leaf_ids[tid] = i;
        st_tri_record(tid, j->pos, j->pidx, j->nrm, j->nidx, j->uvs, j->uidx, j->tris + i);
    }
}

int st_build_mesh_binned(uint32_t T, const float *pos, const uint32_t *pidx, const float *nrm, const uint32_t *nidx,
                         uint32_t normals_count, const float *uvs, const uint32_t *uidx, uint32_t uvs_count,
                         StBVHNodeRec *nodes, uint32_t *node_count, uint32_t *height, StTriangleRec *tris,
                         uint32_t *leaf_ids) {
    sbox *boxes = (sbox *)calloc((size_t)T + 1, sizeof(sbox));
    uint32_t *order = (uint32_t *)malloc(((size_t)T + 1) * sizeof(uint32_t));
    if (!boxes || !order) { free(boxes); free(order); return -1; }
    MeshJob j = {pos, normals_count ? nrm : NULL, uvs_count ? uvs : NULL, pidx, nidx, uidx, order, boxes, tris,
                 leaf_ids};
    parallel_for(T, mesh_boxes, &j);
    *node_count = st_binned_build(boxes, T, MESH_LEAF, nodes, order, height);
    int ok = *node_count != 0;
    if (ok) parallel_for(T, mesh_records, &j);
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
