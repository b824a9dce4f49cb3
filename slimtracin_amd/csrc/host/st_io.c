/*
 * st_io.c - .mesh / .texture readers and writers, and the obj2mesh /
 * bmp2texture conversions (include/slimtracin_io.h).
 *
 * The formats are the reference's (scene/io.h:6-175) in the Windows layout
 * its tools write (u32 = 4 bytes). The conversions follow the tools' steps:
 * obj2mesh (src/obj2mesh.c:25-284) and bmp2texture's loadTexture
 * (src/bmp2texture.c:170-376), with the mesh BVH from the library's
 * reference SAH builder (st_build_mesh = updateMeshBVH).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "st_host.h"
#include "slimtracin.h"
#include "slimtracin_io.h"

/* ---- .mesh ----------------------------------------------------------------- */

#define MESH_HEADER_BYTES (24 + 7 * 4)

int st_mesh_file_parse(const void *data, size_t size, StMeshFile *m) {
    if (!data || !m) return st_host_error(ST_ERR_ARG, "null argument");
    const uint8_t *p = (const uint8_t *)data;
    if (size < MESH_HEADER_BYTES) return st_host_error(ST_ERR_ARG, ".mesh: shorter than its header");
    memset(m, 0, sizeof *m);
    memcpy(m->aabb, p, 24);
    uint32_t c[7];
    memcpy(c, p + 24, sizeof c);
    m->vertex_count = c[0];
    m->triangle_count = c[1];
    m->edge_count = c[2];
    m->uvs_count = c[3];
    m->normals_count = c[4];
    m->node_count = c[5];
    m->height = c[6];
    const uint64_t T = m->triangle_count;
    uint64_t need = MESH_HEADER_BYTES + 12ull * m->vertex_count + 12 * T + 8ull * m->edge_count +
                    (m->uvs_count ? 8ull * m->uvs_count + 12 * T : 0) +
                    (m->normals_count ? 12ull * m->normals_count + 12 * T : 0) + sizeof(StTriangleRec) * T +
                    sizeof(StBVHNodeRec) * (uint64_t)m->node_count + 4 * T;
    if (need != size) return st_host_error(ST_ERR_ARG, ".mesh: the counts in the header do not match the file size");
    if (T && (m->node_count == 0 || m->node_count > 2 * T))
        return st_host_error(ST_ERR_ARG, ".mesh: node count out of range");
    size_t off = MESH_HEADER_BYTES;
#define TAKE(field, type, bytes) do { m->field = (type)(p + off); off += (size_t)(bytes); } while (0)
    TAKE(positions, const float *, 12ull * m->vertex_count);
    TAKE(position_indices, const uint32_t *, 12 * T);
    TAKE(edges, const uint32_t *, 8ull * m->edge_count);
    if (m->uvs_count) {
        TAKE(uvs, const float *, 8ull * m->uvs_count);
        TAKE(uv_indices, const uint32_t *, 12 * T);
    }
    if (m->normals_count) {
        TAKE(normals, const float *, 12ull * m->normals_count);
        TAKE(normal_indices, const uint32_t *, 12 * T);
    }
    TAKE(triangles, const StTriangleRec *, sizeof(StTriangleRec) * T);
    TAKE(nodes, const StBVHNodeRec *, sizeof(StBVHNodeRec) * (uint64_t)m->node_count);
    TAKE(leaf_ids, const uint32_t *, 4 * T);
#undef TAKE
    /* indices must stay inside their arrays */
    for (uint64_t i = 0; i < 3 * T; i++) {
        if (m->position_indices[i] >= m->vertex_count) return st_host_error(ST_ERR_ARG, ".mesh: vertex index out of range");
        if (m->uv_indices && m->uv_indices[i] >= m->uvs_count) return st_host_error(ST_ERR_ARG, ".mesh: uv index out of range");
        if (m->normal_indices && m->normal_indices[i] >= m->normals_count)
            return st_host_error(ST_ERR_ARG, ".mesh: normal index out of range");
    }
    if (T && st_bvh_validate(m->nodes, m->node_count, T) < 0) return st_host_error(ST_ERR_ARG, ".mesh: malformed BVH");
    return ST_OK;
}

int st_mesh_file_write(const StMeshFile *m, void **out, size_t *size) {
    if (!m || !out || !size) return st_host_error(ST_ERR_ARG, "null argument");
    const uint64_t T = m->triangle_count;
    const int has_uv = m->uvs_count != 0, has_n = m->normals_count != 0;
    if (!m->positions || !m->position_indices || (m->edge_count && !m->edges) || !m->triangles || !m->nodes ||
        !m->leaf_ids || (has_uv && (!m->uvs || !m->uv_indices)) || (has_n && (!m->normals || !m->normal_indices)))
        return st_host_error(ST_ERR_ARG, "mesh arrays missing");
    uint64_t bytes = MESH_HEADER_BYTES + 12ull * m->vertex_count + 12 * T + 8ull * m->edge_count +
                     (has_uv ? 8ull * m->uvs_count + 12 * T : 0) + (has_n ? 12ull * m->normals_count + 12 * T : 0) +
                     sizeof(StTriangleRec) * T + sizeof(StBVHNodeRec) * (uint64_t)m->node_count + 4 * T;
    uint8_t *p = (uint8_t *)malloc((size_t)bytes);
    if (!p) return st_host_error(ST_ERR_MEMORY, "out of host memory");
    size_t off = 0;
#define PUT(src, n) do { memcpy(p + off, (src), (size_t)(n)); off += (size_t)(n); } while (0)
    PUT(m->aabb, 24);
    const uint32_t c[7] = {m->vertex_count, m->triangle_count, m->edge_count, m->uvs_count, m->normals_count,
                           m->node_count, m->height};
    PUT(c, sizeof c);
    PUT(m->positions, 12ull * m->vertex_count);
    PUT(m->position_indices, 12 * T);
    if (m->edge_count) PUT(m->edges, 8ull * m->edge_count);
    if (has_uv) {
        PUT(m->uvs, 8ull * m->uvs_count);
        PUT(m->uv_indices, 12 * T);
    }
    if (has_n) {
        PUT(m->normals, 12ull * m->normals_count);
        PUT(m->normal_indices, 12 * T);
    }
    PUT(m->triangles, sizeof(StTriangleRec) * T);
    PUT(m->nodes, sizeof(StBVHNodeRec) * (uint64_t)m->node_count);
    PUT(m->leaf_ids, 4 * T);
#undef PUT
    *out = p;
    *size = off;
    return ST_OK;
}

/* ---- obj2mesh ------------------------------------------------------------------ */

enum { OBJ_NONE, OBJ_POS, OBJ_POS_UV, OBJ_POS_UV_N };

/* next line of at most 1023 characters (fgets into a 1024-byte buffer) */
static const char *obj_line(const char *s, const char *end, char *buf) {
    if (s >= end) return NULL;
    size_t n = 0;
    while (s < end && n < 1023) {
        char ch = *s++;
        buf[n++] = ch;
        if (ch == '\n') break;
    }
    buf[n] = 0;
    return s;
}

/* sscanf-style fields: floats after a 1- or 2-character tag */
static int scan_floats(const char *s, float *out, int n) {
    for (int i = 0; i < n; i++) {
        char *e;
        out[i] = strtof(s, &e);
        if (e == s) return i;
        s = e;
    }
    return n;
}

/* "%d", "%d/%d" or "%d/%d/%d" per vertex after "f" */
static int scan_face(const char *s, int per_vertex, int v[3][3]) {
    s += 1;
    for (int k = 0; k < 3; k++) {
        for (int j = 0; j < per_vertex; j++) {
            char *e;
            long x = strtol(s, &e, 10);
            if (e == s) return 0;
            v[k][j] = (int)x;
            s = e;
            if (j + 1 < per_vertex) {
                if (*s != '/') return 0;
                s++;
            }
        }
    }
    return 1;
}

typedef struct { uint32_t from, to; } ObjEdge;

/* unique undirected edges in first-seen order (obj2mesh.c:218-242) through
   an open-addressing set */
static int obj_edges(const uint32_t *pidx, uint32_t T, ObjEdge **out, uint32_t *count) {
    size_t cap = 16;
    while (cap < (size_t)T * 6 + 16) cap <<= 1;
    uint64_t *set = (uint64_t *)malloc(cap * sizeof(uint64_t));
    ObjEdge *e = (ObjEdge *)malloc(((size_t)T * 3 + 1) * sizeof(ObjEdge));
    if (!set || !e) { free(set); free(e); return -1; }
    memset(set, 0xFF, cap * sizeof(uint64_t));
    uint32_t n = 0;
    for (uint32_t i = 0; i < T; i++)
        for (int from = 0, to = 1; from < 3; from++, to = (to + 1) % 3) {
            uint32_t a = pidx[3 * (size_t)i + from], b = pidx[3 * (size_t)i + to];
            if (a > b) { uint32_t t = a; a = b; b = t; }
            const uint64_t key = ((uint64_t)a << 32) | b;
            size_t h = (size_t)((key * 0x9E3779B97F4A7C15ull) >> 20) & (cap - 1);
            int found = 0;
            while (set[h] != ~0ull) {
                if (set[h] == key) { found = 1; break; }
                h = (h + 1) & (cap - 1);
            }
            if (!found) {
                set[h] = key;
                e[n].from = a;
                e[n].to = b;
                n++;
            }
        }
    free(set);
    *out = e;
    *count = n;
    return 0;
}

int st_obj_to_mesh(const char *text, size_t len, int invert_winding, void **out, size_t *size) {
    if (!text || !out || !size) return st_host_error(ST_ERR_ARG, "null argument");
    const char *end = text + len;
    char line[1024];
    uint32_t nv = 0, nn = 0, nt = 0, T = 0;
    int mode = OBJ_NONE;
    /* counting pass (obj2mesh.c:44-69) */
    for (const char *s = text; (s = obj_line(s, end, line));) {
        if (strncmp(line, "vn", 2) == 0) nn++;
        if (strncmp(line, "vt", 2) == 0) nt++;
        if (strncmp(line, "v ", 2) == 0) nv++;
        if (strncmp(line, "f ", 2) == 0) {
            T++;
            if (mode == OBJ_NONE) {
                int slashes = 0;
                for (const char *ch = line; *ch; ch++) slashes += *ch == '/';
                mode = slashes == 0 ? OBJ_POS : slashes == 3 ? OBJ_POS_UV : slashes == 6 ? OBJ_POS_UV_N : OBJ_NONE;
            }
        }
    }
    if (!T || mode == OBJ_NONE) return st_host_error(ST_ERR_ARG, "OBJ: no faces of a supported form");
    if ((nt && mode == OBJ_POS) || (nn && mode != OBJ_POS_UV_N))
        return st_host_error(ST_ERR_ARG, "OBJ: uv or normal lines without the matching face form");
    float *pos = (float *)calloc((size_t)nv * 3 + 3, sizeof(float));
    float *nrm = nn ? (float *)calloc((size_t)nn * 3, sizeof(float)) : NULL;
    float *uvs = nt ? (float *)calloc((size_t)nt * 2, sizeof(float)) : NULL;
    uint32_t *pidx = (uint32_t *)calloc((size_t)T * 3, sizeof(uint32_t));
    uint32_t *uidx = mode != OBJ_POS ? (uint32_t *)calloc((size_t)T * 3, sizeof(uint32_t)) : NULL;
    uint32_t *nidx = mode == OBJ_POS_UV_N ? (uint32_t *)calloc((size_t)T * 3, sizeof(uint32_t)) : NULL;
    StBVHNodeRec *nodes = (StBVHNodeRec *)calloc((size_t)T * 2, sizeof(StBVHNodeRec));
    StTriangleRec *tris = (StTriangleRec *)calloc((size_t)T, sizeof(StTriangleRec));
    uint32_t *leaf = (uint32_t *)calloc((size_t)T, sizeof(uint32_t));
    ObjEdge *edges = NULL;
    int rc = ST_OK;
    if (!pos || (nn && !nrm) || (nt && !uvs) || !pidx || (mode != OBJ_POS && !uidx) || (mode == OBJ_POS_UV_N && !nidx) ||
        !nodes || !tris || !leaf) {
        rc = st_host_error(ST_ERR_MEMORY, "out of host memory");
        goto done;
    }
    {
        /* parsing pass (obj2mesh.c:110-187); face vertex order 1,2,3 or 1,3,2 */
        const int v1 = 0, v2 = invert_winding ? 2 : 1, v3 = invert_winding ? 1 : 2;
        const int per_vertex = mode == OBJ_POS ? 1 : mode == OBJ_POS_UV ? 2 : 3;
        uint32_t iv = 0, in = 0, it = 0, f = 0;
        for (const char *s = text; (s = obj_line(s, end, line));) {
            if (strncmp(line, "v ", 2) == 0) {
                scan_floats(line + 1, pos + 3 * (size_t)iv++, 3);
            } else if (strncmp(line, "vn", 2) == 0) {
                scan_floats(line + 2, nrm + 3 * (size_t)in++, 3);
            } else if (strncmp(line, "vt", 2) == 0) {
                scan_floats(line + 2, uvs + 2 * (size_t)it++, 2);
            } else if (strncmp(line, "f ", 2) == 0) {
                int v[3][3];
                if (!scan_face(line, per_vertex, v)) {
                    rc = st_host_error(ST_ERR_ARG, "OBJ: a face does not have the first face's form");
                    goto done;
                }
                const int slot[3] = {v1, v2, v3};
                for (int k = 0; k < 3; k++) {
                    pidx[3 * (size_t)f + slot[k]] = (uint32_t)(v[k][0] - 1);
                    if (uidx) uidx[3 * (size_t)f + slot[k]] = (uint32_t)(v[k][1] - 1);
                    if (nidx) nidx[3 * (size_t)f + slot[k]] = (uint32_t)(v[k][2] - 1);
                }
                f++;
            }
        }
        for (size_t i = 0; i < 3 * (size_t)T; i++)
            if (pidx[i] >= nv || (uidx && uidx[i] >= nt) || (nidx && nidx[i] >= nn)) {
                rc = st_host_error(ST_ERR_ARG, "OBJ: face index out of range");
                goto done;
            }
    }
    {
        /* bound from the origin and the vertices, then recentre
           (obj2mesh.c:191-216) */
        float mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
        for (uint32_t i = 0; i < nv; i++)
            for (int a = 0; a < 3; a++) {
                const float x = pos[3 * (size_t)i + a];
                mn[a] = mn[a] < x ? mn[a] : x;
                mx[a] = mx[a] > x ? mx[a] : x;
            }
        float c[3];
        for (int a = 0; a < 3; a++) c[a] = (mn[a] + mx[a]) * 0.5f;
        if (c[0] != 0 || c[1] != 0 || c[2] != 0) {
            for (int a = 0; a < 3; a++) {
                mn[a] = mn[a] - c[a];
                mx[a] = mx[a] - c[a];
            }
            for (uint32_t i = 0; i < nv; i++)
                for (int a = 0; a < 3; a++) pos[3 * (size_t)i + a] = pos[3 * (size_t)i + a] - c[a];
        }
        uint32_t edge_count = 0, node_count = 0, height = 0;
        if (obj_edges(pidx, T, &edges, &edge_count)) {
            rc = st_host_error(ST_ERR_MEMORY, "out of host memory");
            goto done;
        }
        if (st_build_mesh(T, pos, pidx, nrm, nidx, nn, uvs, uidx, nt, nodes, &node_count, &height, tris, leaf)) {
            rc = st_host_error(ST_ERR_MEMORY, "out of host memory");
            goto done;
        }
        StMeshFile m;
        memset(&m, 0, sizeof m);
        for (int a = 0; a < 3; a++) {
            m.aabb[a] = mn[a];
            m.aabb[3 + a] = mx[a];
        }
        m.vertex_count = nv;
        m.triangle_count = T;
        m.edge_count = edge_count;
        m.uvs_count = mode != OBJ_POS ? nt : 0;
        m.normals_count = mode == OBJ_POS_UV_N ? nn : 0;
        m.node_count = node_count;
        m.height = height;
        m.positions = pos;
        m.position_indices = pidx;
        m.edges = (const uint32_t *)edges;
        m.uvs = m.uvs_count ? uvs : NULL;
        m.uv_indices = m.uvs_count ? uidx : NULL;
        m.normals = m.normals_count ? nrm : NULL;
        m.normal_indices = m.normals_count ? nidx : NULL;
        m.triangles = tris;
        m.nodes = nodes;
        m.leaf_ids = leaf;
        rc = st_mesh_file_write(&m, out, size);
    }
done:
    free(pos);
    free(nrm);
    free(uvs);
    free(pidx);
    free(uidx);
    free(nidx);
    free(nodes);
    free(tris);
    free(leaf);
    free(edges);
    return rc;
}

/* ---- .texture ----------------------------------------------------------------- */

int st_texture_file_parse(const void *data, size_t size, StTextureFile *t) {
    if (!data || !t) return st_host_error(ST_ERR_ARG, "null argument");
    const uint8_t *p = (const uint8_t *)data;
    if (size < 7) return st_host_error(ST_ERR_ARG, ".texture: shorter than its header");
    memset(t, 0, sizeof *t);
    memcpy(&t->width, p, 2);
    memcpy(&t->height, p + 2, 2);
    t->mipmap = p[4];
    t->wrap = p[5];
    t->mip_count = p[6];
    if (t->mip_count == 0 || t->mip_count > 16) return st_host_error(ST_ERR_ARG, ".texture: mip count out of range");
    size_t off = 7;
    for (uint32_t i = 0; i < t->mip_count; i++) {
        if (off + 4 > size) return st_host_error(ST_ERR_ARG, ".texture: truncated");
        memcpy(&t->mips[i].width, p + off, 2);
        memcpy(&t->mips[i].height, p + off + 2, 2);
        off += 4;
        const uint64_t q = (uint64_t)(t->mips[i].width + 1u) * (t->mips[i].height + 1u) * 12u;
        if (off + q > size) return st_host_error(ST_ERR_ARG, ".texture: truncated");
        t->mips[i].quads = p + off;
        off += (size_t)q;
    }
    if (off != size) return st_host_error(ST_ERR_ARG, ".texture: trailing bytes");
    return ST_OK;
}

/* One mip level being built: float texels and their texel quads
   (TexelQuadLoader: TL, TR, BL, BR, each r, g, b). */
typedef struct { float r, g, b; } TexF;
typedef struct { TexF q[4]; } QuadF;   /* TL, TR, BL, BR */
typedef struct {
    uint16_t width, height;
    TexF *texels;
    QuadF *quads;                        /* (height+1) x (width+1) */
} MipF;

enum { QTL = 0, QTR = 1, QBL = 2, QBR = 3 };

/* loadTextureMipLoader (bmp2texture.c:182-225): every texel goes to the four
   quads around it; border quads take the edge texels (clamp) or the
   opposite edge's (wrap). Quads are written in the tool's order, so where
   a quad corner is set twice the later write wins, as in the tool. */
static void mip_quads(MipF *m, int wrap) {
    const uint32_t W = m->width, H = m->height, stride = W + 1;
    const uint32_t last_x = W - 1, last_y = H - 1, l = 0, r = W;
    QuadF *top = m->quads, *bottom = m->quads + (size_t)H * stride;
    const TexF *tex = m->texels;
    for (uint32_t y = 0; y < H; y++) {
        QuadF *cur = m->quads + (size_t)y * stride, *next = cur + stride;
        const int T = y == 0, B = y == last_y;
        for (uint32_t x = 0; x < W; x++, tex++) {
            QuadF *TL = cur + x, *TR = cur + x + 1, *BL = next + x, *BR = next + x + 1;
            const int L = x == 0, R = x == last_x;
            TL->q[QBR] = TR->q[QBL] = BL->q[QTR] = BR->q[QTL] = *tex;
            if (L) TL->q[QBL] = BL->q[QTL] = *tex;
            else if (R) TR->q[QBR] = BR->q[QTR] = *tex;
            if (wrap) {
                if (L) cur[r].q[QBR] = next[r].q[QTR] = *tex;
                else if (R) cur[l].q[QBL] = next[l].q[QTL] = *tex;
                if (T) {
                    bottom[x].q[QBR] = bottom[x + 1].q[QBL] = *tex;
                    if (L) bottom[r].q[QBR] = *tex;
                    else if (R) bottom[l].q[QBL] = *tex;
                } else if (B) {
                    top[x].q[QTR] = top[x + 1].q[QTL] = *tex;
                    if (L) top[r].q[QTR] = *tex;
                    else if (R) top[l].q[QTL] = *tex;
                }
            } else {
                if (T) {
                    TL->q[QTR] = TR->q[QTL] = *tex;
                    if (L) TL->q[QTL] = *tex;
                    else if (R) TR->q[QTR] = *tex;
                } else if (B) {
                    BL->q[QBR] = BR->q[QBL] = *tex;
                    if (L) BL->q[QBL] = *tex;
                    else if (R) BR->q[QBR] = *tex;
                }
            }
        }
    }
}

static int mip_alloc(MipF *m, uint16_t w, uint16_t h) {
    m->width = w;
    m->height = h;
    m->texels = (TexF *)calloc((size_t)w * h + 1, sizeof(TexF));
    m->quads = (QuadF *)calloc((size_t)(w + 1) * (h + 1), sizeof(QuadF));
    return m->texels && m->quads ? 0 : -1;
}

/* (u8) of a float in [0, 255] (bmp2texture.c:344-355) */
static uint8_t to_u8(float x) { return (uint8_t)x; }

int st_texture_from_texels(const uint8_t *texels, uint16_t width, uint16_t height, uint32_t bpt, int wrap,
                           int mipmap, void **out, size_t *size) {
    if (!texels || !out || !size) return st_host_error(ST_ERR_ARG, "null argument");
    if (width == 0 || height == 0 || (bpt != 3 && bpt != 4)) return st_host_error(ST_ERR_ARG, "bad texture size");
    /* mip count (bmp2texture.c:234-241) */
    uint32_t count = 1;
    {
        uint16_t w = width, h = height;
        if (mipmap)
            while (w > 4 && h > 4) {
                w /= 2;
                h /= 2;
                count++;
            }
    }
    if (count > 16) return st_host_error(ST_ERR_ARG, "too many mip levels");
    MipF mips[16];
    memset(mips, 0, sizeof mips);
    int rc = ST_OK;
    if (mip_alloc(&mips[0], width, height)) {
        rc = st_host_error(ST_ERR_MEMORY, "out of host memory");
        goto done;
    }
    for (size_t i = 0; i < (size_t)width * height; i++) {
        mips[0].texels[i].r = (float)texels[bpt * i];
        mips[0].texels[i].g = (float)texels[bpt * i + 1];
        mips[0].texels[i].b = (float)texels[bpt * i + 2];
    }
    mip_quads(&mips[0], wrap);
    /* each coarser texel: the mean of the finer level's quad at its 2x2
       block's shared corner, summed TL, TR, BL, BR then times 0.25
       (bmp2texture.c:268-300) */
    for (uint32_t k = 1; k < count; k++) {
        const MipF *src = &mips[k - 1];
        const uint16_t w = src->width / 2, h = src->height / 2;
        if (mip_alloc(&mips[k], w, h)) {
            rc = st_host_error(ST_ERR_MEMORY, "out of host memory");
            goto done;
        }
        const uint32_t sstride = src->width + 1u;
        for (uint32_t y = 0; y < h; y++)
            for (uint32_t x = 0; x < w; x++) {
                const QuadF *cq = src->quads + (size_t)(2 * y + 1) * sstride + 2 * x + 1;
                float R = cq->q[0].r, G = cq->q[0].g, B = cq->q[0].b;
                for (int i = 1; i < 4; i++) {
                    R += cq->q[i].r;
                    G += cq->q[i].g;
                    B += cq->q[i].b;
                }
                TexF *t = mips[k].texels + (size_t)w * y + x;
                t->r = R * 0.25f;
                t->g = G * 0.25f;
                t->b = B * 0.25f;
            }
        mip_quads(&mips[k], wrap);
    }
    {
        size_t bytes = 7;
        for (uint32_t k = 0; k < count; k++) bytes += 4 + (size_t)(mips[k].width + 1) * (mips[k].height + 1) * 12;
        uint8_t *p = (uint8_t *)malloc(bytes);
        if (!p) {
            rc = st_host_error(ST_ERR_MEMORY, "out of host memory");
            goto done;
        }
        memcpy(p, &width, 2);
        memcpy(p + 2, &height, 2);
        p[4] = (uint8_t)(mipmap != 0);
        p[5] = (uint8_t)(wrap != 0);
        p[6] = (uint8_t)count;
        size_t off = 7;
        for (uint32_t k = 0; k < count; k++) {
            memcpy(p + off, &mips[k].width, 2);
            memcpy(p + off + 2, &mips[k].height, 2);
            off += 4;
            const size_t nq = (size_t)(mips[k].width + 1) * (mips[k].height + 1);
            for (size_t i = 0; i < nq; i++, off += 12) {
                const QuadF *q = mips[k].quads + i;
                /* TexelQuad: R{TL,TR,BL,BR}, G{...}, B{...} */
                for (int c = 0; c < 4; c++) {
                    p[off + c] = to_u8(q->q[c].r);
                    p[off + 4 + c] = to_u8(q->q[c].g);
                    p[off + 8 + c] = to_u8(q->q[c].b);
                }
            }
        }
        *out = p;
        *size = off;
    }
done:
    for (uint32_t k = 0; k < 16; k++) {
        free(mips[k].texels);
        free(mips[k].quads);
    }
    return rc;
}

int st_bmp_to_texture(const void *bmp, size_t size, int wrap, int mipmap, void **out, size_t *out_size) {
    if (!bmp || !out || !out_size) return st_host_error(ST_ERR_ARG, "null argument");
    const uint8_t *p = (const uint8_t *)bmp;
    if (size < 54 || p[0] != 'B' || p[1] != 'M') return st_host_error(ST_ERR_ARG, "not a BMP file");
    uint32_t data_off, image_size, compression;
    int32_t w, h;
    uint16_t bits;
    memcpy(&data_off, p + 10, 4);
    memcpy(&w, p + 18, 4);
    memcpy(&h, p + 22, 4);
    memcpy(&bits, p + 28, 2);
    memcpy(&compression, p + 30, 4);
    memcpy(&image_size, p + 34, 4);
    if ((bits != 24 && bits != 32) || compression != 0 || w <= 0 || h <= 0 || w > 65535 || h > 65535)
        return st_host_error(ST_ERR_ARG, "BMP: only uncompressed 24/32-bit images");
    const uint32_t step = bits == 32 ? 4 : 3;
    /* the tool reads width*height texels from the pixel data as stored
       (bottom-up rows, no row padding skipped) and swaps B and R */
    const size_t need = (size_t)w * (size_t)h * step;
    if (data_off > size || size - data_off < need) return st_host_error(ST_ERR_ARG, "BMP: truncated pixel data");
    uint8_t *tex = (uint8_t *)malloc(need);
    if (!tex) return st_host_error(ST_ERR_MEMORY, "out of host memory");
    memcpy(tex, p + data_off, need);
    for (size_t i = 0; i + 2 < need; i += step) {
        uint8_t t = tex[i];
        tex[i] = tex[i + 2];
        tex[i + 2] = t;
    }
    int rc = st_texture_from_texels(tex, (uint16_t)w, (uint16_t)h, step, wrap, mipmap, out, out_size);
    free(tex);
    (void)image_size;
    return rc;
}
