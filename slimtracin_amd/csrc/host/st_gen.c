/*
 * st_gen.c - seeded synthetic scenes, written as scene blobs.
 *
 * Random numbers come from PCG32 (O'Neill 2014, "PCG: a family of simple
 * fast space-efficient statistically good algorithms for random number
 * generation": 64-bit LCG state, XSH-RR output). Rotations are composed
 * with the reference's own rotation formulas (getRotationAroundAxis,
 * mulQuat, normQuat: math/quat.h:15-42,132-139) so that a reference host fed
 * the same amounts builds identical quaternions. Defaults of every record
 * are those of initPrimitive / initMaterial (core/init.h:461-493).
 *
 * Meshes are noise-displaced spheres: a geodesic icosphere of frequency k
 * (every icosahedron face split into k*k triangles, 20*k*k in total) or a
 * latitude/longitude sphere with an exact triangle count. Vertices move
 * radially by 3-octave value noise; vertex normals are the normalised sums
 * of the adjacent face normals; uvs are spherical coordinates.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "st_host.h"
#include "slimtracin.h"
#include "slimtracin_gen.h"

/* ---- PCG32 --------------------------------------------------------------- */
typedef struct { uint64_t state, inc; } Pcg;
static uint32_t pcg_next(Pcg *r) {
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}
static void pcg_seed(Pcg *r, uint64_t seed, uint64_t seq) {
    r->state = 0;
    r->inc = (seq << 1u) | 1u;
    pcg_next(r);
    r->state += seed;
    pcg_next(r);
}
static float rnd01(Pcg *r) { return (float)(pcg_next(r) >> 8) * (1.0f / 16777216.0f); }
static float rnd(Pcg *r, float lo, float hi) { return lo + (hi - lo) * rnd01(r); }

/* ---- blob writer ---------------------------------------------------------- */
typedef struct { uint8_t *p; size_t n, cap; int bad; } Buf;
static void put(Buf *b, const void *src, size_t n) {
    size_t pad = (4 - (n & 3)) & 3;
    if (b->n + n + pad > b->cap) {
        size_t cap = b->cap ? b->cap * 2 : 1 << 16;
        while (cap < b->n + n + pad) cap *= 2;
        uint8_t *q = (uint8_t *)realloc(b->p, cap);
        if (!q) { b->bad = 1; return; }
        b->p = q;
        b->cap = cap;
    }
    memcpy(b->p + b->n, src, n);
    memset(b->p + b->n + n, 0, pad);
    b->n += n + pad;
}

typedef struct {
    float *pos, *nrm, *uv;
    uint32_t *idx;
    uint32_t vcount, tcount;
    float aabb[6];
} GMesh;

static void gmesh_free(GMesh *m) { free(m->pos); free(m->nrm); free(m->uv); free(m->idx); memset(m, 0, sizeof *m); }

static StPrimRec prim_default(uint32_t type, uint32_t id) {
    StPrimRec p;
    memset(&p, 0, sizeof p);
    p.scale[0] = p.scale[1] = p.scale[2] = 1;
    p.rotation[3] = 1;
    p.type = type;
    p.mesh_id = id;
    p.flags = ST_ALL_FLAGS;
    return p;
}
static StMaterialRec mat_default(void) {
    StMaterialRec m;
    memset(&m, 0, sizeof m);
    m.albedo[0] = m.albedo[1] = m.albedo[2] = 1;
    m.reflectivity[0] = m.reflectivity[1] = m.reflectivity[2] = 1;
    m.uv_repeat[0] = m.uv_repeat[1] = 1;
    m.roughness = 1;
    m.n1_over_n2 = m.n2_over_n1 = 1;
    m.brdf = ST_BRDF_LAMBERT;
    m.use = ST_NORMAL_MAP | ST_ALBEDO_MAP;
    return m;
}
static void set3(float *d, float x, float y, float z) { d[0] = x; d[1] = y; d[2] = z; }
static void quat_out(float *d, sq4 q) { d[0] = q.v.x; d[1] = q.v.y; d[2] = q.v.z; d[3] = q.w; }

/* yaw about Y, then pitch about X, then roll about Z, each by the
   reference's rational amount, composed as normQuat(mulQuat(a, b)). */
static sq4 ypr_quat(float yaw, float pitch, float roll) {
    sq4 qy = q_axis(V3(0, 1, 0), yaw), qp = q_axis(V3(1, 0, 0), pitch), qr = q_axis(V3(0, 0, 1), roll);
    return q_norm(q_mul(q_norm(q_mul(qy, qp)), qr));
}

static void write_header(Buf *b, uint32_t prims, uint32_t mats, uint32_t lights, uint32_t meshes, uint32_t textures,
                         const float *ambient) {
    StBlobHeader h;
    memset(&h, 0, sizeof h);
    h.magic = ST_BLOB_MAGIC;
    h.version = ST_BLOB_VERSION;
    h.prim_count = prims;
    h.material_count = mats;
    h.light_count = lights;
    h.mesh_count = meshes;
    h.texture_count = textures;
    memcpy(h.ambient, ambient, 12);
    put(b, &h, sizeof h);
}

static void write_mesh(Buf *b, const GMesh *m) {
    StMeshRec r;
    memset(&r, 0, sizeof r);
    r.vertex_count = m->vcount;
    r.triangle_count = m->tcount;
    r.normals_count = m->vcount;
    r.uvs_count = m->vcount;
    memcpy(r.aabb, m->aabb, sizeof r.aabb);
    put(b, &r, sizeof r);
    put(b, m->pos, 12u * (size_t)m->vcount);
    put(b, m->idx, 12u * (size_t)m->tcount);
    put(b, m->nrm, 12u * (size_t)m->vcount);
    put(b, m->idx, 12u * (size_t)m->tcount);
    put(b, m->uv, 8u * (size_t)m->vcount);
    put(b, m->idx, 12u * (size_t)m->tcount);
}

/* ---- value noise (3 octaves) -------------------------------------------- */
static uint32_t hash3(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    uint32_t h = seed * 0x9E3779B1u;
    h ^= (uint32_t)x * 0x85EBCA77u; h = (h << 13) | (h >> 19); h *= 5u; h += 0xE6546B64u;
    h ^= (uint32_t)y * 0xC2B2AE3Du; h = (h << 13) | (h >> 19); h *= 5u; h += 0xE6546B64u;
    h ^= (uint32_t)z * 0x27D4EB2Fu; h = (h << 13) | (h >> 19); h *= 5u; h += 0xE6546B64u;
    h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16;
    return h;
}
static double lattice(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    return (double)(hash3(x, y, z, seed) >> 8) * (2.0 / 16777216.0) - 1.0;
}
static double smooth(double t) { return t * t * (3 - 2 * t); }
static double value_noise(double x, double y, double z, uint32_t seed) {
    double fx = floor(x), fy = floor(y), fz = floor(z);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    double tx = smooth(x - fx), ty = smooth(y - fy), tz = smooth(z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; a++)
        for (int b2 = 0; b2 < 2; b2++)
            for (int d = 0; d < 2; d++) c[a][b2][d] = lattice(ix + a, iy + b2, iz + d, seed);
    double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx, x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx, x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    double y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}
static double fbm3(double x, double y, double z, uint32_t seed) {
    return (value_noise(x, y, z, seed) + 0.5 * value_noise(2 * x, 2 * y, 2 * z, seed + 1) +
            0.25 * value_noise(4 * x, 4 * y, 4 * z, seed + 2)) / 1.75;
}

/* Radial displacement, smooth normals, spherical uvs and the AABB. */
static void finish_sphere(GMesh *m, double *dpos, uint32_t seed, double amplitude, double frequency) {
    for (uint32_t v = 0; v < m->vcount; v++) {
        double *p = dpos + 3 * (size_t)v;
        double n = fbm3(p[0] * frequency, p[1] * frequency, p[2] * frequency, seed);
        double r = 1.0 + amplitude * n;
        double u = atan2(p[2], p[0]) / (2 * M_PI) + 0.5;
        double w = acos(p[1] < -1 ? -1 : p[1] > 1 ? 1 : p[1]) / M_PI;
        m->uv[2 * (size_t)v] = (float)u;
        m->uv[2 * (size_t)v + 1] = (float)w;
        for (int k = 0; k < 3; k++) m->pos[3 * (size_t)v + k] = (float)(p[k] * r);
    }
    double *acc = (double *)calloc(3 * (size_t)m->vcount, sizeof(double));
    for (uint32_t t = 0; t < m->tcount; t++) {
        const uint32_t *f = m->idx + 3 * (size_t)t;
        const float *a = m->pos + 3 * (size_t)f[0], *b = m->pos + 3 * (size_t)f[1], *c = m->pos + 3 * (size_t)f[2];
        double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++) acc[3 * (size_t)f[k] + j] += n[j];
    }
    for (uint32_t v = 0; v < m->vcount; v++) {
        double *n = acc + 3 * (size_t)v;
        double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (l == 0) l = 1;
        for (int k = 0; k < 3; k++) m->nrm[3 * (size_t)v + k] = (float)(n[k] / l);
    }
    free(acc);
    for (int k = 0; k < 3; k++) { m->aabb[k] = INFINITY; m->aabb[3 + k] = -INFINITY; }
    for (uint32_t v = 0; v < m->vcount; v++)
        for (int k = 0; k < 3; k++) {
            float c = m->pos[3 * (size_t)v + k];
            m->aabb[k] = m->aabb[k] < c ? m->aabb[k] : c;
            m->aabb[3 + k] = m->aabb[3 + k] > c ? m->aabb[3 + k] : c;
        }
}

static int gmesh_alloc(GMesh *m, uint32_t vcount, uint32_t tcount) {
    memset(m, 0, sizeof *m);
    m->vcount = vcount;
    m->tcount = tcount;
    m->pos = (float *)malloc(12u * (size_t)vcount);
    m->nrm = (float *)malloc(12u * (size_t)vcount);
    m->uv = (float *)malloc(8u * (size_t)vcount);
    m->idx = (uint32_t *)malloc(12u * (size_t)tcount);
    return (m->pos && m->nrm && m->uv && m->idx) ? 0 : -1;
}

/* Geodesic icosphere of frequency k: 10k^2+2 vertices, 20k^2 triangles.
   Vertices on icosahedron edges are shared by computing them from the edge's
   two corners in a fixed (low id first) order. */
static int gen_icosphere(GMesh *m, uint32_t k, uint32_t seed, double amplitude) {
    static const int F[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                                 {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                                 {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    const double t = (1.0 + sqrt(5.0)) / 2.0;
    double V[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                       {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    uint32_t vcount = 10 * k * k + 2, tcount = 20 * k * k;
    if (gmesh_alloc(m, vcount, tcount)) return -1;
    double *dpos = (double *)malloc(24u * (size_t)vcount);
    /* vertex numbering: 12 corners, then (k-1) per edge (30 edges), then
       (k-1)(k-2)/2 interior points per face */
    int edge_id[12][12];
    memset(edge_id, -1, sizeof edge_id);
    int ne = 0;
    for (int f = 0; f < 20; f++)
        for (int e = 0; e < 3; e++) {
            int a = F[f][e], b = F[f][(e + 1) % 3];
            int lo = a < b ? a : b, hi = a < b ? b : a;
            if (edge_id[lo][hi] < 0) edge_id[lo][hi] = ne++;
        }
    uint32_t edge_base = 12, face_base = 12 + 30 * (k - 1);
    for (int c = 0; c < 12; c++) memcpy(dpos + 3 * c, V[c], sizeof V[c]);
    for (int lo = 0; lo < 12; lo++)
        for (int hi = lo + 1; hi < 12; hi++) {
            int e = edge_id[lo][hi];
            if (e < 0) continue;
            for (uint32_t s = 1; s < k; s++) {
                double *p = dpos + 3 * (size_t)(edge_base + e * (k - 1) + (s - 1));
                for (int q = 0; q < 3; q++) p[q] = V[lo][q] + (V[hi][q] - V[lo][q]) * (double)s / k;
            }
        }
    uint32_t *grid = (uint32_t *)malloc(sizeof(uint32_t) * (k + 1) * (k + 1));
    uint32_t tri = 0;
    for (int f = 0; f < 20; f++) {
        int A = F[f][0], B = F[f][1], C = F[f][2];
        uint32_t interior = 0;
        for (uint32_t i = 0; i <= k; i++)
            for (uint32_t j = 0; i + j <= k; j++) {
                uint32_t id;
                uint32_t l = k - i - j; /* barycentric weights: A*l + B*i + C*j */
                if (i == 0 && j == 0) id = (uint32_t)A;
                else if (i == k) id = (uint32_t)B;
                else if (j == k) id = (uint32_t)C;
                else if (j == 0 || i == 0 || l == 0) {
                    int a, b;
                    uint32_t s;
                    if (j == 0) { a = A; b = B; s = i; }
                    else if (i == 0) { a = A; b = C; s = j; }
                    else { a = B; b = C; s = j; }
                    int lo = a < b ? a : b, hi = a < b ? b : a;
                    uint32_t step = a == lo ? s : k - s;
                    id = edge_base + (uint32_t)edge_id[lo][hi] * (k - 1) + (step - 1);
                } else {
                    id = face_base + (uint32_t)f * ((k - 1) * (k - 2) / 2) + interior++;
                    double *p = dpos + 3 * (size_t)id;
                    for (int q = 0; q < 3; q++) p[q] = (V[A][q] * l + V[B][q] * i + V[C][q] * j) / k;
                }
                grid[i * (k + 1) + j] = id;
            }
        for (uint32_t i = 0; i < k; i++)
            for (uint32_t j = 0; i + j < k; j++) {
                uint32_t *o = m->idx + 3 * (size_t)tri++;
                o[0] = grid[i * (k + 1) + j];
                o[1] = grid[(i + 1) * (k + 1) + j];
                o[2] = grid[i * (k + 1) + j + 1];
                if (i + j + 1 < k) {
                    o = m->idx + 3 * (size_t)tri++;
                    o[0] = grid[(i + 1) * (k + 1) + j];
                    o[1] = grid[(i + 1) * (k + 1) + j + 1];
                    o[2] = grid[i * (k + 1) + j + 1];
                }
            }
    }
    free(grid);
    for (uint32_t v = 0; v < vcount; v++) {
        double *p = dpos + 3 * (size_t)v;
        double l = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (int q = 0; q < 3; q++) p[q] /= l;
    }
    finish_sphere(m, dpos, seed, amplitude, 4.0);
    free(dpos);
    return tri == tcount ? 0 : -1;
}

/* Latitude/longitude sphere with exactly 2*rings*segments triangles
   (poles are rings of degenerate-free thin triangles around a tiny cap). */
static int gen_latlong(GMesh *m, uint32_t rings, uint32_t segments, uint32_t seed, double amplitude) {
    uint32_t vcount = (rings + 1) * (segments + 1), tcount = 2 * rings * segments;
    if (gmesh_alloc(m, vcount, tcount)) return -1;
    double *dpos = (double *)malloc(24u * (size_t)vcount);
    for (uint32_t r = 0; r <= rings; r++) {
        double th = 0.02 + (M_PI - 0.04) * r / rings;
        for (uint32_t s = 0; s <= segments; s++) {
            double ph = 2 * M_PI * s / segments;
            double *p = dpos + 3 * (size_t)(r * (segments + 1) + s);
            p[0] = sin(th) * cos(ph);
            p[1] = cos(th);
            p[2] = sin(th) * sin(ph);
        }
    }
    uint32_t t = 0;
    for (uint32_t r = 0; r < rings; r++)
        for (uint32_t s = 0; s < segments; s++) {
            uint32_t a = r * (segments + 1) + s, b = a + 1, c = a + segments + 1, d = c + 1;
            uint32_t *o = m->idx + 3 * (size_t)t;
            o[0] = a; o[1] = c; o[2] = b;
            o[3] = b; o[4] = c; o[5] = d;
            t += 2;
        }
    finish_sphere(m, dpos, seed, amplitude, 4.0);
    free(dpos);
    return 0;
}

static int finish_blob(Buf *b, void **out, size_t *size) {
    if (b->bad) { free(b->p); return ST_ERR_MEMORY; }
    *out = b->p;
    *size = b->n;
    return ST_OK;
}

/* Config "implicit": 4 quads + 4 boxes + 4 spheres + 4 tetrahedra at
   U[-5,5]^3, random yaw/pitch/roll amounts in [-1,1], uniform scale
   U[0.5,2]; 3 white point lights with intensity U[1,10] at U[-8,8] x
   U[6,12] x U[-8,8]; one Blinn material per primitive with albedo U[0,1]^3,
   reflectivity 1 and roughness 0.5; ambient 0. */
int st_gen_implicit(uint64_t seed, void **out, size_t *size) {
    Pcg r;
    pcg_seed(&r, seed, 1);
    Buf b = {0};
    float ambient[3] = {0, 0, 0};
    const uint32_t P = 16, L = 3;
    write_header(&b, P, P, L, 0, 0, ambient);
    static const uint32_t types[4] = {ST_PRIM_QUAD, ST_PRIM_BOX, ST_PRIM_SPHERE, ST_PRIM_TETRAHEDRON};
    StMaterialRec mats[16];
    for (uint32_t i = 0; i < P; i++) {
        StPrimRec p = prim_default(types[i / 4], i);
        set3(p.position, rnd(&r, -5, 5), rnd(&r, -5, 5), rnd(&r, -5, 5));
        float yaw = rnd(&r, -1, 1), pitch = rnd(&r, -1, 1), roll = rnd(&r, -1, 1);
        quat_out(p.rotation, ypr_quat(yaw, pitch, roll));
        float s = rnd(&r, 0.5f, 2.0f);
        set3(p.scale, s, s, s);
        p.material_id = i;
        put(&b, &p, sizeof p);
        mats[i] = mat_default();
        mats[i].brdf = ST_BRDF_BLINN;
        mats[i].roughness = 0.5f;
        set3(mats[i].albedo, rnd01(&r), rnd01(&r), rnd01(&r));
    }
    put(&b, mats, sizeof mats);
    for (uint32_t i = 0; i < L; i++) {
        StLightRec l;
        set3(l.position, rnd(&r, -8, 8), rnd(&r, 6, 12), rnd(&r, -8, 8));
        set3(l.color, 1, 1, 1);
        l.intensity = rnd(&r, 1, 10);
        put(&b, &l, sizeof l);
    }
    return finish_blob(&b, out, size);
}

/* Config "pbr": one noise-displaced geodesic icosphere (frequency k, 20k^2
   triangles, radius 1 +- 0.1, scaled by 3) with a Cook-Torrance material
   (roughness U[0.05,1], metallic 0 or 1), a Cook-Torrance floor quad, a glass
   sphere (n1/n2 = 1/1.5), a mirror quad, 4 point lights and 2 emissive quads. */
int st_gen_pbr(uint64_t seed, uint32_t k, void **out, size_t *size) {
    Pcg r;
    pcg_seed(&r, seed, 2);
    GMesh m;
    if (gen_icosphere(&m, k, (uint32_t)seed, 0.1)) { gmesh_free(&m); return ST_ERR_MEMORY; }
    Buf b = {0};
    float ambient[3] = {0.004f, 0.004f, 0.007f};
    enum { MESH, FLOOR, GLASS, MIRROR, AREA1, AREA2, NPRIM };
    write_header(&b, NPRIM, NPRIM, 4, 1, 0, ambient);
    StPrimRec p[NPRIM];
    p[MESH] = prim_default(ST_PRIM_MESH, 0);
    set3(p[MESH].position, 0, 3, 2);
    set3(p[MESH].scale, 3, 3, 3);
    quat_out(p[MESH].rotation, ypr_quat(rnd(&r, -1, 1), rnd(&r, -0.3f, 0.3f), 0));
    p[FLOOR] = prim_default(ST_PRIM_QUAD, FLOOR);
    set3(p[FLOOR].scale, 40, 1, 40);
    p[GLASS] = prim_default(ST_PRIM_SPHERE, GLASS);
    set3(p[GLASS].position, -5, 2, -1);
    set3(p[GLASS].scale, 2, 2, 2);
    p[MIRROR] = prim_default(ST_PRIM_QUAD, MIRROR);
    set3(p[MIRROR].position, 7, 5, 6);
    set3(p[MIRROR].scale, 4, 1, 6);
    quat_out(p[MIRROR].rotation, q_norm(q_mul(q_axis(V3(0, 0, 1), 1.0f), q_axis(V3(1, 0, 0), 0.2f))));
    p[AREA1] = prim_default(ST_PRIM_QUAD, AREA1);
    set3(p[AREA1].position, -15, 5, 5);
    set3(p[AREA1].scale, 4, 1, 8);
    { sq4 q; q.v = V3(0, 0, -0.70710678118f); q.w = 0.70710678118f; quat_out(p[AREA1].rotation, q); }
    p[AREA2] = prim_default(ST_PRIM_QUAD, AREA2);
    set3(p[AREA2].position, 15, 5, 5);
    set3(p[AREA2].scale, 4, 1, 8);
    { sq4 q; q.v = V3(0, 0, -0.70710678118f); q.w = -0.70710678118f; quat_out(p[AREA2].rotation, q); }
    for (int i = 0; i < NPRIM; i++) p[i].material_id = (uint32_t)i;
    put(&b, p, sizeof p);
    StMaterialRec mt[NPRIM];
    for (int i = 0; i < NPRIM; i++) mt[i] = mat_default();
    mt[MESH].brdf = ST_BRDF_COOK_TORRANCE;
    mt[MESH].roughness = rnd(&r, 0.05f, 1.0f);
    mt[MESH].metallic = (pcg_next(&r) & 1u) ? 1.0f : 0.0f;
    set3(mt[MESH].albedo, rnd(&r, 0.2f, 1), rnd(&r, 0.2f, 1), rnd(&r, 0.2f, 1));
    for (int c = 0; c < 3; c++) /* lerpVec3(plastic 0.04, albedo, metallic) */
        mt[MESH].reflectivity[c] = (mt[MESH].albedo[c] - 0.04f) * mt[MESH].metallic + 0.04f;
    mt[FLOOR].brdf = ST_BRDF_COOK_TORRANCE;
    mt[FLOOR].roughness = rnd(&r, 0.05f, 1.0f);
    set3(mt[FLOOR].albedo, 0.8f, 0.8f, 0.8f);
    set3(mt[FLOOR].reflectivity, 0.04f, 0.04f, 0.04f);
    mt[GLASS].brdf = ST_BRDF_BLINN;
    mt[GLASS].is = ST_REFRACTIVE;
    mt[GLASS].roughness = 0;
    mt[GLASS].n1_over_n2 = 1.0f / 1.5f;
    mt[GLASS].n2_over_n1 = 1.5f;
    set3(mt[GLASS].reflectivity, 0.7f, 0.9f, 0.7f);
    mt[MIRROR].brdf = ST_BRDF_BLINN;
    mt[MIRROR].is = ST_REFLECTIVE;
    mt[MIRROR].roughness = 0;
    set3(mt[MIRROR].reflectivity, 0.9f, 0.9f, 0.9f);
    mt[AREA1].is = ST_EMISSIVE;
    set3(mt[AREA1].emission, 0.9f, 0.7f, 0.7f);
    mt[AREA2].is = ST_EMISSIVE;
    set3(mt[AREA2].emission, 0.7f, 0.7f, 0.9f);
    put(&b, mt, sizeof mt);
    for (int i = 0; i < 4; i++) {
        StLightRec l;
        set3(l.position, i % 2 ? -10.0f : 10.0f, rnd(&r, 8, 14), i < 2 ? -8.0f : 12.0f);
        set3(l.color, rnd(&r, 0.6f, 1), rnd(&r, 0.6f, 1), rnd(&r, 0.6f, 1));
        l.intensity = rnd(&r, 40, 120);
        put(&b, &l, sizeof l);
    }
    write_mesh(&b, &m);
    gmesh_free(&m);
    return finish_blob(&b, out, size);
}

/* A noise-displaced latitude/longitude sphere with `triangles` triangles
   (rounded to an even rings*segments factorisation), as a standalone mesh
   scene: the mesh, a floor and 3 point lights. Used for the batched and
   served workloads' resident scenes. */
int st_gen_mesh_scene(uint64_t seed, uint32_t triangles, void **out, size_t *size) {
    Pcg r;
    pcg_seed(&r, seed, 3);
    uint32_t half = triangles / 2 ? triangles / 2 : 1;
    uint32_t rings = (uint32_t)sqrt((double)half / 2.0);
    if (rings < 2) rings = 2;
    uint32_t segments = half / rings;
    if (segments < 3) segments = 3;
    GMesh m;
    if (gen_latlong(&m, rings, segments, (uint32_t)seed, 0.1)) { gmesh_free(&m); return ST_ERR_MEMORY; }
    Buf b = {0};
    float ambient[3] = {0.004f, 0.004f, 0.007f};
    write_header(&b, 3, 3, 3, 1, 0, ambient);
    StPrimRec p[3];
    p[0] = prim_default(ST_PRIM_MESH, 0);
    set3(p[0].position, 0, 3, 0);
    set3(p[0].scale, 3, 3, 3);
    quat_out(p[0].rotation, ypr_quat(rnd(&r, -1, 1), rnd(&r, -0.5f, 0.5f), rnd(&r, -0.2f, 0.2f)));
    p[1] = prim_default(ST_PRIM_QUAD, 1);
    set3(p[1].scale, 40, 1, 40);
    p[2] = prim_default(ST_PRIM_SPHERE, 2);
    set3(p[2].position, rnd(&r, -6, 6), 1.5f, rnd(&r, -6, 6));
    set3(p[2].scale, 1.5f, 1.5f, 1.5f);
    for (int i = 0; i < 3; i++) p[i].material_id = (uint32_t)i;
    put(&b, p, sizeof p);
    StMaterialRec mt[3];
    for (int i = 0; i < 3; i++) mt[i] = mat_default();
    mt[0].brdf = ST_BRDF_COOK_TORRANCE;
    mt[0].roughness = rnd(&r, 0.05f, 1.0f);
    mt[0].metallic = (pcg_next(&r) & 1u) ? 1.0f : 0.0f;
    set3(mt[0].albedo, rnd(&r, 0.2f, 1), rnd(&r, 0.2f, 1), rnd(&r, 0.2f, 1));
    for (int c = 0; c < 3; c++) mt[0].reflectivity[c] = (mt[0].albedo[c] - 0.04f) * mt[0].metallic + 0.04f;
    mt[1].brdf = ST_BRDF_COOK_TORRANCE;
    mt[1].roughness = 0.6f;
    set3(mt[1].reflectivity, 0.04f, 0.04f, 0.04f);
    mt[2].brdf = ST_BRDF_BLINN;
    mt[2].is = (pcg_next(&r) & 1u) ? ST_REFLECTIVE : ST_REFRACTIVE;
    mt[2].roughness = 0;
    mt[2].n1_over_n2 = 1.0f / 1.5f;
    mt[2].n2_over_n1 = 1.5f;
    set3(mt[2].reflectivity, 0.9f, 0.9f, 0.9f);
    put(&b, mt, sizeof mt);
    for (int i = 0; i < 3; i++) {
        StLightRec l;
        set3(l.position, rnd(&r, -12, 12), rnd(&r, 8, 14), rnd(&r, -12, 12));
        set3(l.color, 1, 1, 1);
        l.intensity = rnd(&r, 40, 120);
        put(&b, &l, sizeof l);
    }
    write_mesh(&b, &m);
    gmesh_free(&m);
    return finish_blob(&b, out, size);
}

/* ---- Perlin textures ------------------------------------------------------ */

/* Tileable 2D gradient noise (Perlin 1985/2002: fade 6t^5-15t^4+10t^3,
   pseudo-random unit gradients on the lattice) with period `per` cells. */
static double g_grad[256][2];
static int g_grad_ready = 0;
static double grad_dot(int32_t ix, int32_t iy, int32_t per, uint32_t seed, double dx, double dy) {
    if (!g_grad_ready) {
        for (int k = 0; k < 256; k++) {
            g_grad[k][0] = cos(k * (2 * M_PI / 256.0));
            g_grad[k][1] = sin(k * (2 * M_PI / 256.0));
        }
        g_grad_ready = 1;
    }
    uint32_t h = hash3(((ix % per) + per) % per, ((iy % per) + per) % per, 0, seed) & 255u;
    return g_grad[h][0] * dx + g_grad[h][1] * dy;
}
static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
static double perlin2(double x, double y, int32_t per, uint32_t seed) {
    int32_t ix = (int32_t)floor(x), iy = (int32_t)floor(y);
    double fx = x - ix, fy = y - iy;
    double n00 = grad_dot(ix, iy, per, seed, fx, fy), n10 = grad_dot(ix + 1, iy, per, seed, fx - 1, fy);
    double n01 = grad_dot(ix, iy + 1, per, seed, fx, fy - 1), n11 = grad_dot(ix + 1, iy + 1, per, seed, fx - 1, fy - 1);
    double u = fade(fx), v = fade(fy);
    double a = n00 + u * (n10 - n00), b = n01 + u * (n11 - n01);
    return a + v * (b - a);
}
/* 5 octaves, base period 8 cells over the texture, result about [-1, 1] */
static double perlin_fbm(double u, double v, uint32_t seed) {
    double sum = 0, amp = 1, norm = 0;
    int32_t per = 8;
    for (int o = 0; o < 5; o++) {
        sum += amp * perlin2(u * per, v * per, per, seed + (uint32_t)o);
        norm += amp;
        amp *= 0.5;
        per *= 2;
    }
    return sum / norm * 1.6;
}

/* One RGB8 texture of size x size and its mip chain down to 4x4 (each level
   the rounded mean of 2x2 texels of the one above), written as TexelQuad
   levels: quad (x, y) of a w x h level holds the four texels around grid
   corner (x, y), indices clamped, per channel TL, TR, BL, BR. */
static void write_texture(Buf *b, const uint8_t *rgb, uint32_t size) {
    uint32_t levels = 0;
    for (uint32_t w = size; w >= 4; w >>= 1) { levels++; if (w == 4) break; }
    StTextureRec tr;
    memset(&tr, 0, sizeof tr);
    tr.width = (uint16_t)size;
    tr.height = (uint16_t)size;
    tr.mip_count = (uint8_t)levels;
    tr.wrap = 1;
    tr.mipmap = 1;
    put(b, &tr, sizeof tr);
    uint8_t *cur = (uint8_t *)malloc(3u * size * size);
    memcpy(cur, rgb, 3u * size * size);
    uint32_t w = size;
    uint8_t *quads = (uint8_t *)malloc(12u * (size + 1) * (size + 1));
    for (uint32_t l = 0; l < levels; l++) {
        uint16_t dims[2] = {(uint16_t)w, (uint16_t)w};
        put(b, dims, 4);
        for (uint32_t y = 0; y <= w; y++)
            for (uint32_t x = 0; x <= w; x++) {
                uint32_t x0 = x ? x - 1 : 0, y0 = y ? y - 1 : 0, x1 = x < w ? x : w - 1, y1 = y < w ? y : w - 1;
                const uint8_t *tl = cur + 3 * ((size_t)y0 * w + x0), *trp = cur + 3 * ((size_t)y0 * w + x1);
                const uint8_t *bl = cur + 3 * ((size_t)y1 * w + x0), *br = cur + 3 * ((size_t)y1 * w + x1);
                uint8_t *q = quads + 12 * ((size_t)y * (w + 1) + x);
                for (int c = 0; c < 3; c++) {
                    q[4 * c] = tl[c]; q[4 * c + 1] = trp[c]; q[4 * c + 2] = bl[c]; q[4 * c + 3] = br[c];
                }
            }
        put(b, quads, 12u * (size_t)(w + 1) * (w + 1));
        if (l + 1 < levels) {
            uint32_t h = w / 2;
            uint8_t *nxt = (uint8_t *)malloc(3u * h * h);
            for (uint32_t y = 0; y < h; y++)
                for (uint32_t x = 0; x < h; x++)
                    for (int c = 0; c < 3; c++) {
                        uint32_t s4 = cur[3 * ((2 * y) * (size_t)w + 2 * x) + c] + cur[3 * ((2 * y) * (size_t)w + 2 * x + 1) + c] +
                                      cur[3 * ((2 * y + 1) * (size_t)w + 2 * x) + c] +
                                      cur[3 * ((2 * y + 1) * (size_t)w + 2 * x + 1) + c];
                        nxt[3 * ((size_t)y * h + x) + c] = (uint8_t)((s4 + 2) / 4);
                    }
            free(cur);
            cur = nxt;
            w = h;
        }
    }
    free(quads);
    free(cur);
}

/* Albedo (three noise channels around a seeded base colour) and a normal
   map from the gradient of a noise height field, both size x size RGB8. */
static void perlin_textures(uint32_t size, uint32_t seed, uint8_t *albedo, uint8_t *normal) {
    double *hgt = (double *)malloc(sizeof(double) * size * size);
    for (uint32_t y = 0; y < size; y++)
        for (uint32_t x = 0; x < size; x++) {
            double u = (x + 0.5) / size, v = (y + 0.5) / size;
            double r = 0.55 + 0.35 * perlin_fbm(u, v, seed), g = 0.45 + 0.35 * perlin_fbm(u, v, seed + 17),
                   bb = 0.35 + 0.35 * perlin_fbm(u, v, seed + 31);
            double c[3] = {r, g, bb};
            for (int k = 0; k < 3; k++) {
                double t = c[k] < 0 ? 0 : c[k] > 1 ? 1 : c[k];
                albedo[3 * ((size_t)y * size + x) + k] = (uint8_t)(t * 255.0 + 0.5);
            }
            hgt[(size_t)y * size + x] = perlin_fbm(u, v, seed + 101);
        }
    for (uint32_t y = 0; y < size; y++)
        for (uint32_t x = 0; x < size; x++) {
            double hx = hgt[(size_t)y * size + (x + 1) % size] - hgt[(size_t)y * size + (x + size - 1) % size];
            double hy = hgt[(size_t)((y + 1) % size) * size + x] - hgt[(size_t)((y + size - 1) % size) * size + x];
            double n[3] = {-hx * size * 0.02, -hy * size * 0.02, 1.0};
            double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int k = 0; k < 3; k++)
                normal[3 * ((size_t)y * size + x) + k] = (uint8_t)((n[k] / l * 0.5 + 0.5) * 255.0 + 0.5);
        }
    free(hgt);
}

/* Config "instanced": `instances` mesh primitives drawing from `mesh_count`
   noise-displaced sphere meshes of mesh_triangles[i] triangles, random
   positions over a 60 x 60 field, random yaw/pitch/roll amounts, uniform
   scale U[0.8, 2.5]; a floor; 3 point lights; materials textured with a
   Perlin albedo map and a Perlin normal map (tex_size^2, full mip chains)
   - Cook-Torrance and Blinn, one untextured mirror. */
int st_gen_instanced(uint64_t seed, uint32_t instances, const uint32_t *mesh_triangles, uint32_t mesh_count,
                     uint32_t tex_size, void **out, size_t *size) {
    if (!mesh_count || !instances || tex_size < 4 || (tex_size & (tex_size - 1))) return ST_ERR_ARG;
    Pcg r;
    pcg_seed(&r, seed, 5);
    Buf b = {0};
    float ambient[3] = {0.01f, 0.01f, 0.012f};
    const uint32_t nmat = 4;
    write_header(&b, instances + 1, nmat, 3, mesh_count, 2, ambient);
    for (uint32_t i = 0; i < instances; i++) {
        StPrimRec p = prim_default(ST_PRIM_MESH, pcg_next(&r) % mesh_count);
        set3(p.position, rnd(&r, -30, 30), rnd(&r, 1.5f, 6), rnd(&r, -20, 40));
        float sc = rnd(&r, 0.8f, 2.5f);
        set3(p.scale, sc, sc, sc);
        quat_out(p.rotation, ypr_quat(rnd(&r, -1, 1), rnd(&r, -1, 1), rnd(&r, -1, 1)));
        p.material_id = pcg_next(&r) % nmat;
        put(&b, &p, sizeof p);
    }
    StPrimRec floor_p = prim_default(ST_PRIM_QUAD, 0);
    set3(floor_p.scale, 80, 1, 80);
    floor_p.material_id = 0;
    put(&b, &floor_p, sizeof floor_p);
    StMaterialRec m[4];
    for (int i = 0; i < 4; i++) m[i] = mat_default();
    m[0].brdf = ST_BRDF_COOK_TORRANCE;
    m[0].roughness = 0.6f;
    set3(m[0].reflectivity, 0.04f, 0.04f, 0.04f);
    m[0].texture_count = 2;
    m[0].texture_ids[0] = 0;
    m[0].texture_ids[1] = 1;
    m[0].uv_repeat[0] = m[0].uv_repeat[1] = 4;
    m[1] = m[0];
    m[1].roughness = 0.25f;
    m[1].metallic = 1;
    set3(m[1].reflectivity, 0.9f, 0.8f, 0.6f);
    m[1].uv_repeat[0] = m[1].uv_repeat[1] = 2;
    m[2] = m[0];
    m[2].brdf = ST_BRDF_BLINN;
    m[2].roughness = 0.5f;
    set3(m[2].reflectivity, 0.5f, 0.5f, 0.5f);
    m[3].brdf = ST_BRDF_BLINN;
    m[3].is = ST_REFLECTIVE;
    m[3].roughness = 0;
    m[3].use = 0;
    set3(m[3].reflectivity, 0.9f, 0.9f, 0.9f);
    put(&b, m, sizeof m);
    for (int i = 0; i < 3; i++) {
        StLightRec l;
        set3(l.position, rnd(&r, -25, 25), rnd(&r, 12, 20), rnd(&r, -15, 35));
        set3(l.color, 1, 1, 1);
        l.intensity = rnd(&r, 150, 300);
        put(&b, &l, sizeof l);
    }
    for (uint32_t k = 0; k < mesh_count; k++) {
        uint32_t half = mesh_triangles[k] / 2 ? mesh_triangles[k] / 2 : 1;
        uint32_t rings = (uint32_t)sqrt((double)half / 2.0);
        if (rings < 2) rings = 2;
        uint32_t segments = half / rings;
        if (segments < 3) segments = 3;
        GMesh gm;
        if (gen_latlong(&gm, rings, segments, (uint32_t)seed + 77 * k, 0.15)) { gmesh_free(&gm); free(b.p); return ST_ERR_MEMORY; }
        write_mesh(&b, &gm);
        gmesh_free(&gm);
    }
    uint8_t *albedo = (uint8_t *)malloc(3u * tex_size * tex_size), *normal = (uint8_t *)malloc(3u * tex_size * tex_size);
    if (!albedo || !normal) { free(albedo); free(normal); free(b.p); return ST_ERR_MEMORY; }
    perlin_textures(tex_size, (uint32_t)seed * 7u + 3u, albedo, normal);
    write_texture(&b, albedo, tex_size);
    write_texture(&b, normal, tex_size);
    free(albedo);
    free(normal);
    return finish_blob(&b, out, size);
}

void st_free(void *p) { free(p); }
