/*
 * st_oracle.c - CPU restatement of the reference per-pixel tracing path.
 *
 * TEST INFRASTRUCTURE ONLY (see st_oracle.h). Every function below follows a
 * reference function operation for operation: same fp32 operation order, no
 * contraction (built with -ffp-contract=off on baseline x86-64, where the
 * reference's fast_mul_add(a,b,c) is a*b+c, SlimTracin core/base.h:71-76),
 * the same ternary min/max forms and the same integer casts. Each function
 * cites the reference lines it restates. The shading state that the
 * reference aliases through `Trace` (closest/current/mesh hits overwritten by
 * shadow and area-light passes) is kept in the same shape so the data flow,
 * not just the formulas, matches.
 *
 * Two places where the reference reads memory it never wrote get a defined
 * value here, and the HIP library uses the same value:
 *   - hitTetrahedron's local RayHit (render/shaders/intersection/
 *     tetrahedra.h:11) copies uninitialised cone_angle/dUV/ids: zero here.
 *   - the `dUV.v` of shadeSurface's bounce branch and of the Normals mode
 *     block (closest_hit/surface.h:152,230; render/raytracer.h:103-104) is
 *     never assigned when no earlier branch set it: zero here.
 * Both only reach the mip level of a texture fetch.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "st_oracle.h"

/* ---- constants (core/base.h:107-173) ---------------------------------- */
#define O_EPS 0.0001f
#define O_SQRT2 1.41421356237f
#define O_SQRT3 1.73205080757f
#define O_PI 3.14159265f
#define O_ONE_OVER_PI (1.0f / O_PI)
#define O_UNIT_SPHERE_AREA_OVER_SIX ((4.0f * O_PI) / 6.0f)
#define O_COLOR_TO_FLOAT 0.00392156862f
#define O_TRACE_OFFSET 0.0001f
#define O_TET_A 0.577350259f
#define O_TET_B 0.288675159f

/* ---- value types (core/base.h:193-208, core/types.h) ------------------- */
typedef struct { float x, y, z; } V3;
typedef struct { float x, y; } V2;
typedef struct { V3 v; float w; } Q;
typedef struct { V3 X, Y, Z; } M3;
typedef struct { V3 min, max; } Box3;

static inline V3 v3(float x, float y, float z) { V3 o; o.x = x; o.y = y; o.z = z; return o; }
static inline V3 v3s(float s) { return v3(s, s, s); }
static inline V2 v2(float x, float y) { V2 o; o.x = x; o.y = y; return o; }

/* math/vec3.h:143-266 */
static inline V3 add3(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub3(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 mul3(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V3 scl3(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 neg3(V3 a) { return v3(-a.x, -a.y, -a.z); }
static inline V3 mad3(V3 v, V3 f, V3 a) { return v3(v.x * f.x + a.x, v.y * f.y + a.y, v.z * f.z + a.z); }
static inline V3 smad3(V3 v, float s, V3 a) { return v3(v.x * s + a.x, v.y * s + a.y, v.z * s + a.z); }
static inline float dot3(V3 a, V3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
static inline V3 cross3(V3 a, V3 b) {
    return v3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
static inline float sqlen3(V3 v) { return (v.x * v.x) + (v.y * v.y) + (v.z * v.z); }
static inline float len3(V3 v) { return sqrtf(sqlen3(v)); }
static inline V3 norm3(V3 v) { return scl3(v, 1.0f / len3(v)); }
static inline V3 rcp3(V3 v) { return v3(1.0f / v.x, 1.0f / v.y, 1.0f / v.z); }
static inline V3 one_minus3(V3 v) { return v3(1.0f - v.x, 1.0f - v.y, 1.0f - v.z); }
/* core/base.h:390-393 */
static inline float clamp01(float v) { float mn = v < 1.0f ? v : 1.0f; return mn > 0.0f ? mn : 0.0f; }
static inline float dotc3(V3 a, V3 b) { return clamp01(dot3(a, b)); }
static inline float maxc3(V3 v) { float o = v.x; if (v.y > o) o = v.y; if (v.z > o) o = v.z; return o; }
/* math/vec3.h:203-211 */
static inline V3 vmat3(V3 in, M3 m) {
    return v3(in.x * m.X.x + in.y * m.Y.x + in.z * m.Z.x,
              in.x * m.X.y + in.y * m.Y.y + in.z * m.Z.y,
              in.x * m.X.z + in.y * m.Y.z + in.z * m.Z.z);
}

/* math/quat.h:15-51 */
static inline Q qnorm(Q q) {
    float f = 1.0f / sqrtf(q.v.x * q.v.x + q.v.y * q.v.y + q.v.z * q.v.z + q.w * q.w);
    Q o; o.v = scl3(q.v, f); o.w = q.w * f; return o;
}
static inline V3 qrot(V3 v, Q q) {
    V3 out = cross3(q.v, v);
    V3 qqv = cross3(q.v, out);
    out = smad3(out, q.w, qqv);
    return smad3(out, 2.0f, v);
}
static inline Q qconj(Q q) { Q o; o.w = q.w; o.v = neg3(q.v); return o; }

/* ---- scene records ------------------------------------------------------ */
typedef struct {
    V3 position, scale;
    Q rotation;
    uint32_t type, mesh_id, material_id;
    uint8_t flags;
    int32_t rect[4]; /* screen bounds min.x, min.y, max.x, max.y */
} OPrim;

typedef struct {
    V3 albedo, reflectivity, emission;
    V2 uv_repeat;
    float metallic, roughness, n1_over_n2, n2_over_n1;
    uint32_t brdf;
    uint8_t is, use, texture_count, texture_ids[16];
} OMat;

typedef struct { V3 position, color; float intensity; } OLight;

typedef struct { Box3 aabb; uint32_t first_child_id; uint16_t child_count, depth; } ONode;

typedef struct {
    M3 world_to_tangent;
    V3 position, normal, vertex_normals[3];
    V2 uvs[3];
    float area_of_parallelogram, area_of_uv;
} OTri;

typedef struct {
    Box3 aabb;
    uint32_t vertex_count, triangle_count, normals_count, uvs_count;
    uint32_t node_count, height;
    ONode *nodes;
    OTri *triangles;
    uint32_t *leaf_ids;
} OMesh;

typedef struct { uint16_t width, height; uint8_t *quads; } OMip; /* quads: 12 B each */
typedef struct { uint16_t width, height; uint8_t mip_count, wrap, mipmap; OMip *mips; } OTex;

typedef struct {
    uint32_t prim_count, material_count, light_count, mesh_count, texture_count;
    V3 ambient;
    OPrim *prims;
    OMat *mats;
    OLight *lights;
    OMesh *meshes;
    OTex *textures;
    ONode *bvh_nodes;
    uint32_t *bvh_leaf_ids;
    uint32_t bvh_node_count, bvh_height;
} OScene;

/* ---- per-ray state (core/types.h:176-200) ------------------------------ */
typedef struct {
    V3 position, normal;
    V2 uv, dUV;
    float distance, distance_squared, cone_angle, cone_width, NdotV, area, uv_area;
    uint32_t material_id, object_id, object_type;
    uint8_t from_behind;
} OHit;

typedef struct { float b, c, t_near, t_far, furthest, closest_hit_distance, closest_hit_density; } OSphereHit;

typedef struct {
    V3 origin, scaled_origin, direction, direction_reciprocal;
    uint8_t octant[3];
} ORay;

typedef struct {
    OSphereHit sphere_hit;
    OHit closest_hit, closest_mesh_hit, current_hit;
    ORay local;
    uint32_t scene_stack[1024], mesh_stack[256];
    uint8_t depth, mesh_stack_size;
    int32_t primary_triangle; /* triangle index consumed by the last hitPrimitives */
    uint32_t ray_class, level;  /* class of the ray being traced (ST_ORC_*), depth-loop level */
    StOracleStats st;
} OTrace;

typedef struct {
    V3 albedo, position, normal, viewing_direction, viewing_origin, reflected_direction, light_direction;
    V3 quad_vertices[4];
    V2 uv;
    const OPrim *primitive;
    const OMat *material;
} OShaded;

/* ---- traversal helpers ------------------------------------------------- */

/* core/init.h:536-543 */
static inline void prep_ray(ORay *r) {
    r->octant[0] = signbit(r->direction.x) ? 3 : 0;
    r->octant[1] = signbit(r->direction.y) ? 3 : 0;
    r->octant[2] = signbit(r->direction.z) ? 3 : 0;
    r->scaled_origin.x = -r->origin.x * r->direction_reciprocal.x;
    r->scaled_origin.y = -r->origin.y * r->direction_reciprocal.y;
    r->scaled_origin.z = -r->origin.z * r->direction_reciprocal.z;
}

/* render/AABB.h:6-23: the near plane is max when the octant bit is set. */
static inline int hit_aabb(const Box3 *b, const ORay *r, float closest, float *dist) {
    V3 lo = v3(r->octant[0] ? b->max.x : b->min.x, r->octant[1] ? b->max.y : b->min.y, r->octant[2] ? b->max.z : b->min.z);
    V3 hi = v3(r->octant[0] ? b->min.x : b->max.x, r->octant[1] ? b->min.y : b->max.y, r->octant[2] ? b->min.z : b->max.z);
    lo = mad3(lo, r->direction_reciprocal, r->scaled_origin);
    hi = mad3(hi, r->direction_reciprocal, r->scaled_origin);
    hi.x = hi.x < hi.y ? hi.x : hi.y;
    hi.x = hi.x < hi.z ? hi.x : hi.z;
    hi.x = hi.x < closest ? hi.x : closest;
    lo.x = lo.x > lo.y ? lo.x : lo.y;
    lo.x = lo.x > lo.z ? lo.x : lo.z;
    lo.x = lo.x > 0 ? lo.x : 0;
    *dist = lo.x;
    return lo.x <= hi.x;
}

static inline Box3 merge_boxes(Box3 a, Box3 b) {
    Box3 m;
    m.min.x = a.min.x < b.min.x ? a.min.x : b.min.x;
    m.min.y = a.min.y < b.min.y ? a.min.y : b.min.y;
    m.min.z = a.min.z < b.min.z ? a.min.z : b.min.z;
    m.max.x = a.max.x > b.max.x ? a.max.x : b.max.x;
    m.max.y = a.max.y > b.max.y ? a.max.y : b.max.y;
    m.max.z = a.max.z > b.max.z ? a.max.z : b.max.z;
    return m;
}

static inline float box_area(Box3 b) {
    V3 e = sub3(b.max, b.min);
    return e.x * e.y + e.y * e.z + e.z * e.x;
}

/* ---- instance transforms (scene/primitive.h:8-59) ---------------------- */
static inline void to_object(V3 pos, V3 dir, const OPrim *p, V3 *opos, V3 *odir) {
    *opos = (p->flags & ST_IS_TRANSLATED) ? sub3(pos, p->position) : pos;
    if (p->flags & ST_IS_ROTATED) {
        Q inv = qconj(p->rotation);
        *opos = qrot(*opos, inv);
        *odir = qrot(dir, inv);
    } else
        *odir = dir;
    if (p->flags & ST_IS_SCALED) {
        V3 inv_scale = rcp3(p->scale);
        *opos = mul3(*opos, inv_scale);
        if (p->flags & ST_IS_SCALED_NON_UNIFORMLY) *odir = norm3(mul3(*odir, inv_scale));
    }
}
static inline V3 pos_to_world(V3 pos, const OPrim *p) {
    if (p->flags & ST_IS_SCALED) pos = mul3(pos, p->scale);
    if (p->flags & ST_IS_ROTATED) pos = qrot(pos, p->rotation);
    if (p->flags & ST_IS_TRANSLATED) pos = add3(pos, p->position);
    return pos;
}
static inline V3 pos_to_object(V3 pos, const OPrim *p) {
    if (p->flags & ST_IS_TRANSLATED) pos = sub3(pos, p->position);
    if (p->flags & ST_IS_ROTATED) pos = qrot(pos, qconj(p->rotation));
    if (p->flags & ST_IS_SCALED) pos = mul3(pos, rcp3(p->scale));
    return pos;
}
static inline V3 dir_to_world(V3 dir, const OPrim *p) {
    if (p->flags & ST_IS_SCALED_NON_UNIFORMLY) dir = mul3(dir, rcp3(p->scale));
    if (p->flags & ST_IS_ROTATED) dir = qrot(dir, p->rotation);
    return dir;
}

/* ---- intersection shaders --------------------------------------------- */

/* render/shaders/common.h:115-119; the float->u8 cast truncates through a
   32-bit integer as the reference's x86 build does. */
static inline int transparent_uv(V2 uv) {
    uint8_t v = (uint8_t)(int32_t)(uv.y * 4);
    uint8_t u = (uint8_t)(int32_t)(uv.x * 4);
    return (v % 2 != 0) == (u % 2);
}

/* common.h:177-195 */
static inline int hit_plane(V3 po, V3 pn, V3 ro, V3 rd, OHit *hit) {
    hit->NdotV = -dot3(rd, pn);
    if (hit->NdotV == 0) return 0;
    int facing = hit->NdotV > 0;
    V3 RP = sub3(po, ro);
    float RPdotN = dot3(RP, pn);
    hit->from_behind = RPdotN > 0;
    if (hit->from_behind == facing) return 0;
    hit->distance = fabsf(RPdotN / hit->NdotV);
    hit->position = scl3(rd, hit->distance);
    hit->position = add3(ro, hit->position);
    hit->normal = pn;
    return 1;
}

/* intersection/quad.h:7-42 */
static inline int hit_quad(OHit *hit, V3 Ro, V3 Rd, uint8_t flags) {
    if (Rd.y == 0) return 0;
    if (Ro.y == 0) return 0;
    hit->from_behind = Ro.y < 0;
    if (hit->from_behind == (Rd.y < 0)) return 0;
    hit->distance = fabsf(Ro.y / Rd.y);
    hit->position = smad3(Rd, hit->distance, Ro);
    if (hit->position.x < -1 || hit->position.x > +1 || hit->position.z < -1 || hit->position.z > +1) return 0;
    hit->uv.x = (hit->position.x + 1.0f) * 0.5f;
    hit->uv.y = (hit->position.z + 1.0f) * 0.5f;
    if ((flags & ST_IS_TRANSPARENT) && transparent_uv(hit->uv)) return 0;
    hit->normal = v3(0, 1, 0);
    hit->NdotV = -Rd.y;
    hit->area = 4;
    hit->uv_area = 1;
    return 1;
}

enum { SIDE_NONE = 0, SIDE_TOP = 1, SIDE_BOTTOM = 2, SIDE_LEFT = 4, SIDE_RIGHT = 8, SIDE_FRONT = 16, SIDE_BACK = 32 };

/* common.h:121-130 */
static inline int box_side(V3 oct, int axis) {
    switch (axis) {
        case 0: return oct.x > 0 ? SIDE_RIGHT : SIDE_LEFT;
        case 3: return oct.x > 0 ? SIDE_LEFT : SIDE_RIGHT;
        case 1: return oct.y > 0 ? SIDE_TOP : SIDE_BOTTOM;
        case 4: return oct.y > 0 ? SIDE_BOTTOM : SIDE_TOP;
        case 2: return oct.z > 0 ? SIDE_FRONT : SIDE_BACK;
        default: return oct.z > 0 ? SIDE_BACK : SIDE_FRONT;
    }
}
/* common.h:132-168 */
static inline V2 cube_uv(V3 p, int side) {
    V2 uv;
    switch (side) {
        case SIDE_TOP: uv.x = p.x; uv.y = p.z; break;
        case SIDE_BOTTOM: uv.x = -p.x; uv.y = -p.z; break;
        case SIDE_LEFT: uv.x = -p.z; uv.y = p.y; break;
        case SIDE_RIGHT: uv.x = p.z; uv.y = p.y; break;
        case SIDE_FRONT: uv.x = p.x; uv.y = p.y; break;
        default: uv.x = -p.x; uv.y = p.y; break;
    }
    uv.x += 1;
    uv.y += 1;
    uv.x *= 0.5f;
    uv.y *= 0.5f;
    return uv;
}

/* intersection/box.h:7-72 */
static inline int hit_box(OHit *hit, V3 Ro, V3 Rd, uint8_t flags) {
    V3 rcp = rcp3(Rd), oct;
    oct.x = signbit(Rd.x) ? 1.0f : -1.0f;
    oct.y = signbit(Rd.y) ? 1.0f : -1.0f;
    oct.z = signbit(Rd.z) ? 1.0f : -1.0f;
    float t[6];
    t[0] = (oct.x - Ro.x) * rcp.x;
    t[1] = (oct.y - Ro.y) * rcp.y;
    t[2] = (oct.z - Ro.z) * rcp.z;
    t[3] = (-oct.x - Ro.x) * rcp.x;
    t[4] = (-oct.y - Ro.y) * rcp.y;
    t[5] = (-oct.z - Ro.z) * rcp.z;
    int max_axis = t[3] < t[4] ? 3 : 4;
    if (t[5] < t[max_axis]) max_axis = 5;
    float maxt = t[max_axis];
    if (maxt < 0) return SIDE_NONE;
    int min_axis = t[0] > t[1] ? 0 : 1;
    if (t[2] > t[min_axis]) min_axis = 2;
    float mint = t[min_axis];
    if (maxt < (mint > 0 ? mint : 0)) return SIDE_NONE;
    hit->from_behind = mint < 0;
    if (hit->from_behind) { mint = maxt; min_axis = max_axis; }
    int side = box_side(oct, min_axis);
    hit->position = smad3(Rd, mint, Ro);
    hit->uv = cube_uv(hit->position, side);
    if ((flags & ST_IS_TRANSPARENT) && transparent_uv(hit->uv)) {
        if (hit->from_behind) return SIDE_NONE;
        side = box_side(oct, max_axis);
        hit->position = smad3(Rd, maxt, Ro);
        hit->uv = cube_uv(hit->position, side);
        if (transparent_uv(hit->uv)) return SIDE_NONE;
        hit->from_behind = 1;
    }
    hit->normal = v3s(0);
    switch (side) {
        case SIDE_LEFT: hit->normal.x = -1.0f; hit->NdotV = +Rd.x; break;
        case SIDE_RIGHT: hit->normal.x = +1.0f; hit->NdotV = -Rd.x; break;
        case SIDE_BOTTOM: hit->normal.y = -1.0f; hit->NdotV = +Rd.y; break;
        case SIDE_TOP: hit->normal.y = +1.0f; hit->NdotV = -Rd.y; break;
        case SIDE_BACK: hit->normal.z = -1.0f; hit->NdotV = +Rd.z; break;
        case SIDE_FRONT: hit->normal.z = +1.0f; hit->NdotV = -Rd.z; break;
        default: return SIDE_NONE;
    }
    if (hit->from_behind) { hit->normal = neg3(hit->normal); hit->NdotV = -hit->NdotV; }
    hit->area = 4;
    hit->uv_area = 1;
    return side;
}

/* intersection/sphere.h:7-42 */
static inline V2 sphere_uv(V3 d) {
    V2 uv;
    float x = d.x, y = d.y, z = d.z;
    float z_over_x = x ? (z / x) : 2;
    float y_over_x = x ? (y / x) : 2;
    if (z_over_x <= 1 && z_over_x >= -1 && y_over_x <= 1 && y_over_x >= -1) {
        uv.x = z_over_x;
        uv.y = x > 0 ? y_over_x : -y_over_x;
    } else {
        float x_over_z = z ? (x / z) : 2;
        float y_over_z = z ? (y / z) : 2;
        if (x_over_z <= 1 && x_over_z >= -1 && y_over_z <= 1 && y_over_z >= -1) {
            uv.x = -x_over_z;
            uv.y = z > 0 ? y_over_z : -y_over_z;
        } else {
            uv.x = x / (y > 0 ? y : -y);
            uv.y = z / y;
        }
    }
    uv.x += 1; uv.x /= 2;
    uv.y += 1; uv.y /= 2;
    return uv;
}

/* sphere.h:44-59 */
static inline int hit_sphere_simple(V3 Ro, V3 Rd, V3 target, float one_over_radius, OSphereHit *h) {
    V3 rc = scl3(sub3(target, Ro), one_over_radius);
    h->b = dot3(Rd, rc);
    h->c = dot3(rc, rc) - 1;
    float d = h->b * h->b - h->c;
    if (d < 0) return 0;
    d = sqrtf(d);
    h->t_near = h->b - d;
    h->t_far = h->b + d;
    return h->t_far > 0 && h->t_near < h->furthest;
}

/* sphere.h:61-97 */
static inline int hit_sphere(OHit *hit, V3 Ro, V3 Rd, uint8_t flags) {
    float t_to_closest = -dot3(Ro, Rd);
    if (t_to_closest <= 0) return 0;
    hit->distance_squared = sqlen3(Ro) - t_to_closest * t_to_closest;
    float delta_squared = 1 - hit->distance_squared;
    if (delta_squared <= 0) { hit->distance = t_to_closest; return 0; }
    float delta = sqrtf(delta_squared);
    hit->distance = t_to_closest - delta;
    int has_outer_hit = hit->distance > 0;
    if (has_outer_hit) {
        hit->position = smad3(Rd, hit->distance, Ro);
        hit->uv = sphere_uv(hit->position);
        if ((flags & ST_IS_TRANSPARENT) && transparent_uv(hit->uv)) has_outer_hit = 0;
    }
    if (!has_outer_hit) {
        hit->distance = t_to_closest + delta;
        hit->position = smad3(Rd, hit->distance, Ro);
        hit->uv = sphere_uv(hit->position);
        if ((flags & ST_IS_TRANSPARENT) && transparent_uv(hit->uv)) return 0;
    }
    hit->from_behind = !has_outer_hit;
    hit->normal = hit->position;
    hit->NdotV = -dot3(hit->normal, Rd);
    hit->area = hit->uv_area = O_UNIT_SPHERE_AREA_OVER_SIX;
    return 1;
}

/* intersection/tetrahedra.h:9-107: four faces with hard-coded tangent frames. */
static const float TET_FRAMES[4][9] = {
    { O_TET_A, -O_TET_B, -O_TET_A,  O_TET_B,  O_TET_B,  O_TET_A, -O_TET_B,  O_TET_A, -O_TET_A},
    { O_TET_B,  O_TET_B,  O_TET_A, -O_TET_B,  O_TET_A, -O_TET_A,  O_TET_A, -O_TET_B, -O_TET_A},
    {-O_TET_B,  O_TET_A, -O_TET_A,  O_TET_A, -O_TET_B, -O_TET_A,  O_TET_B,  O_TET_B,  O_TET_A},
    {-O_TET_A,  O_TET_B,  O_TET_A,  O_TET_B,  O_TET_B,  O_TET_A,  O_TET_B, -O_TET_A,  O_TET_A},
};
static inline int hit_tetrahedron(OHit *closest_hit, V3 Ro, V3 Rd, uint8_t flags) {
    int found = 0;
    float closest_distance = INFINITY;
    OHit hit;
    memset(&hit, 0, sizeof hit); /* reference leaves it uninitialised */
    for (int t = 0; t < 4; t++) {
        V3 tangent_pos = hit.normal = v3s(t == 3 ? O_TET_A : -O_TET_A);
        switch (t) {
            case 0: hit.normal.y = O_TET_A; break;
            case 1: hit.normal.x = O_TET_A; break;
            case 2: hit.normal.z = O_TET_A; break;
            case 3: tangent_pos.y = -O_TET_A; break;
        }
        if (!hit_plane(tangent_pos, hit.normal, Ro, Rd, &hit)) continue;
        const float *f = TET_FRAMES[t];
        M3 m;
        m.X = v3(f[0], f[1], f[2]);
        m.Y = v3(f[3], f[4], f[5]);
        m.Z = v3(f[6], f[7], f[8]);
        hit.position = smad3(Rd, hit.distance, Ro);
        tangent_pos = sub3(hit.position, tangent_pos);
        tangent_pos = vmat3(tangent_pos, m);
        if (tangent_pos.x < 0 || tangent_pos.y < 0 || tangent_pos.y + tangent_pos.x > 1) continue;
        hit.uv.x = tangent_pos.x;
        hit.uv.y = tangent_pos.y;
        if ((flags & ST_IS_TRANSPARENT) && transparent_uv(hit.uv)) continue;
        if (hit.distance < closest_distance) {
            closest_distance = hit.distance;
            hit.area = O_SQRT3 / 4.0f;
            hit.uv_area = 1;
            hit.NdotV = -dot3(hit.normal, Rd);
            *closest_hit = hit;
            found = 1;
        }
    }
    return found;
}

/* intersection/mesh.h:5-37 */
static inline int hit_triangles(OTrace *tr, const OTri *tris, uint32_t count, int any_hit) {
    ORay *ray = &tr->local;
    OHit *hit = &tr->current_hit, *closest = &tr->closest_mesh_hit;
    int found = 0;
    for (uint32_t i = 0; i < count; i++) {
        const OTri *t = tris + i;
        tr->st.triangle_tests++;
        tr->st.class_events[tr->ray_class][2]++;
        if (hit_plane(t->position, t->normal, ray->origin, ray->direction, hit) && hit->distance < closest->distance) {
            V3 UV = sub3(hit->position, t->position);
            UV = vmat3(UV, t->world_to_tangent);
            if (UV.x < 0 || UV.y < 0 || (UV.x + UV.y) > 1) continue;
            closest->NdotV = hit->NdotV;
            closest->normal = hit->normal;
            closest->uv.x = UV.x;
            closest->uv.y = UV.y;
            closest->object_id = i;
            closest->from_behind = hit->from_behind;
            closest->distance = hit->distance;
            closest->position = hit->position;
            closest->area = t->area_of_parallelogram;
            closest->uv_area = t->area_of_uv;
            found = 1;
            if (any_hit) break;
        }
    }
    return found;
}

/* intersection/mesh.h:38-111 */
static int trace_mesh(OTrace *tr, const OMesh *mesh, int any_hit) {
    ORay *ray = &tr->local;
    OHit *closest = &tr->closest_mesh_hit;
    uint32_t *stack = tr->mesh_stack;
    int found = 0;
    float ld, rd;
    tr->st.root_tests++;
    tr->st.class_events[tr->ray_class][1]++;
    if (!hit_aabb(&mesh->nodes[0].aabb, ray, closest->distance, &ld)) return 0;
    if (mesh->nodes[0].child_count) return hit_triangles(tr, mesh->triangles, mesh->triangle_count, any_hit);
    const ONode *left = mesh->nodes + mesh->nodes[0].first_child_id, *right, *tmp;
    uint32_t stack_size = 0;
    for (;;) {
        right = left + 1;
        tr->st.node_pairs++;
        tr->st.class_events[tr->ray_class][0]++;
        int hl = hit_aabb(&left->aabb, ray, closest->distance, &ld);
        int hr = hit_aabb(&right->aabb, ray, closest->distance, &rd);
        if (hl) {
            if (left->child_count) {
                if (hit_triangles(tr, mesh->triangles + left->first_child_id, left->child_count, any_hit)) {
                    closest->object_id += left->first_child_id;
                    found = 1;
                    if (any_hit) break;
                }
                left = NULL;
            }
        } else
            left = NULL;
        if (hr) {
            if (right->child_count) {
                if (hit_triangles(tr, mesh->triangles + right->first_child_id, right->child_count, any_hit)) {
                    closest->object_id += right->first_child_id;
                    found = 1;
                    if (any_hit) break;
                }
                right = NULL;
            }
        } else
            right = NULL;
        if (left) {
            if (right) {
                if (!any_hit && ld > rd) { tmp = left; left = right; right = tmp; }
                stack[stack_size++] = right->first_child_id;
                if (stack_size == tr->mesh_stack_size) return 0;
            }
            left = mesh->nodes + left->first_child_id;
        } else if (right) {
            left = mesh->nodes + right->first_child_id;
        } else {
            if (stack_size == 0) break;
            left = mesh->nodes + stack[--stack_size];
        }
    }
    return found;
}

/* intersection/primitives.h:9-111 */
static int hit_primitives(const ORay *ray, OTrace *tr, const OScene *sc, const uint32_t *ids, uint32_t count,
                          int any_hit, int check_visibility, uint16_t x, uint16_t y) {
    int found = 0, cur;
    V3 *Rd = &tr->local.direction, *Ro = &tr->local.origin;
    OHit *hit = &tr->current_hit, *closest = &tr->closest_hit;
    float cone_angle = closest->cone_angle;
    const OPrim *hit_prim = NULL;
    uint32_t hit_prim_id = 0;
    for (uint32_t i = 0; i < count; i++) {
        const OPrim *p = sc->prims + ids[i];
        tr->st.primitive_visits++;
        tr->st.class_events[tr->ray_class][3]++;
        if (any_hit && !(p->flags & ST_IS_SHADOWING)) continue;
        if (check_visibility) {
            if (!(p->flags & ST_IS_VISIBLE)) continue;
            if (x < p->rect[0] || x >= p->rect[2] || y < p->rect[1] || y >= p->rect[3]) continue;
        }
        to_object(ray->origin, ray->direction, p, Ro, Rd);
        *Ro = smad3(*Rd, O_TRACE_OFFSET, *Ro);
        switch (p->type) {
            case ST_PRIM_QUAD: cur = hit_quad(hit, *Ro, *Rd, p->flags); break;
            case ST_PRIM_BOX: cur = hit_box(hit, *Ro, *Rd, p->flags) != SIDE_NONE; break;
            case ST_PRIM_SPHERE: cur = hit_sphere(hit, *Ro, *Rd, p->flags); break;
            case ST_PRIM_TETRAHEDRON: cur = hit_tetrahedron(hit, *Ro, *Rd, p->flags); break;
            case ST_PRIM_MESH:
                tr->local.direction_reciprocal = rcp3(*Rd);
                tr->closest_mesh_hit.distance = closest->distance == INFINITY
                    ? INFINITY : len3(sub3(pos_to_object(closest->position, p), *Ro));
                prep_ray(&tr->local);
                cur = trace_mesh(tr, sc->meshes + p->mesh_id, any_hit);
                if (cur) *hit = tr->closest_mesh_hit;
                break;
            default: continue;
        }
        if (cur) {
            hit->position = pos_to_world(hit->position, p);
            hit->distance_squared = sqlen3(sub3(hit->position, ray->origin));
            if (hit->distance_squared < closest->distance_squared) {
                *closest = *hit;
                hit_prim_id = ids[i];
                hit_prim = p;
                closest->object_type = p->type;
                closest->material_id = p->material_id;
                found = 1;
                if (any_hit) break;
            }
        }
    }
    if (found) {
        tr->primary_triangle = -1;
        if (closest->object_type == ST_PRIM_MESH) {
            const OMesh *mesh = sc->meshes + hit_prim->mesh_id;
            const OTri *t = mesh->triangles + closest->object_id;
            tr->primary_triangle = (int32_t)closest->object_id;
            tr->st.mesh_attr_fetches++;
            if (mesh->normals_count | mesh->uvs_count) {
                float u = closest->uv.x, v = closest->uv.y;
                float w = 1 - u - v;
                if (mesh->uvs_count) {
                    closest->uv.x = t->uvs[2].x * u + (t->uvs[1].x * v + t->uvs[0].x * w);
                    closest->uv.y = t->uvs[2].y * u + (t->uvs[1].y * v + t->uvs[0].y * w);
                }
                if (mesh->normals_count) {
                    closest->normal.x = t->vertex_normals[2].x * u + (t->vertex_normals[1].x * v + t->vertex_normals[0].x * w);
                    closest->normal.y = t->vertex_normals[2].y * u + (t->vertex_normals[1].y * v + t->vertex_normals[0].y * w);
                    closest->normal.z = t->vertex_normals[2].z * u + (t->vertex_normals[1].z * v + t->vertex_normals[0].z * w);
                    closest->normal = norm3(closest->normal);
                }
            }
        }
        closest->object_id = hit_prim_id;
        closest->normal = norm3(dir_to_world(closest->normal, hit_prim));
        closest->distance = sqrtf(closest->distance_squared);
        closest->cone_width = 2.0f * cone_angle * closest->distance;
        V3 rotated_scale = qrot(hit_prim->scale, hit_prim->rotation);
        closest->area *= dot3(one_minus3(closest->normal), rotated_scale);
    }
    return found;
}

/* render/shaders/trace.h:8-76 */
static int trace_scene(ORay *ray, OTrace *tr, const OScene *sc, int any_hit) {
    ray->direction_reciprocal = rcp3(ray->direction);
    prep_ray(ray);
    int found = 0;
    float ld, rd;
    tr->st.root_tests++;
    tr->st.class_events[tr->ray_class][1]++;
    if (!hit_aabb(&sc->bvh_nodes[0].aabb, ray, tr->closest_hit.distance, &ld)) return 0;
    if (sc->bvh_nodes[0].child_count)
        return hit_primitives(ray, tr, sc, sc->bvh_leaf_ids, sc->prim_count, any_hit, 0, 0, 0);
    const ONode *left = sc->bvh_nodes + sc->bvh_nodes[0].first_child_id, *right, *tmp;
    uint32_t *stack = tr->scene_stack;
    uint32_t stack_size = 0;
    for (;;) {
        right = left + 1;
        tr->st.node_pairs++;
        tr->st.class_events[tr->ray_class][0]++;
        int hl = hit_aabb(&left->aabb, ray, tr->closest_hit.distance, &ld);
        int hr = hit_aabb(&right->aabb, ray, tr->closest_hit.distance, &rd);
        if (hl) {
            if (left->child_count) {
                if (hit_primitives(ray, tr, sc, sc->bvh_leaf_ids + left->first_child_id, left->child_count, any_hit, 0, 0, 0)) {
                    found = 1;
                    if (any_hit) break;
                }
                left = NULL;
            }
        } else
            left = NULL;
        if (hr) {
            if (right->child_count) {
                if (hit_primitives(ray, tr, sc, sc->bvh_leaf_ids + right->first_child_id, right->child_count, any_hit, 0, 0, 0)) {
                    found = 1;
                    if (any_hit) break;
                }
                right = NULL;
            }
        } else
            right = NULL;
        if (left) {
            if (right) {
                if (!any_hit && ld > rd) { tmp = left; left = right; right = tmp; }
                stack[stack_size++] = right->first_child_id;
            }
            left = sc->bvh_nodes + left->first_child_id;
        } else if (right) {
            left = sc->bvh_nodes + right->first_child_id;
        } else {
            if (stack_size == 0) break;
            left = sc->bvh_nodes + stack[--stack_size];
        }
    }
    return found;
}

/* trace.h:85-92 */
static inline int in_shadow(ORay *ray, OTrace *tr, const OScene *sc) {
    tr->st.shadow_rays++;
    tr->ray_class = tr->level ? ST_ORC_SHADOW_DEEP : ST_ORC_SHADOW0;
    return trace_scene(ray, tr, sc, 1);
}
static inline int trace_ray(ORay *ray, OTrace *tr, const OScene *sc) {
    tr->st.bounce_rays++;
    tr->ray_class = ST_ORC_BOUNCE;
    tr->level++;   /* shadow rays after this bounce belong to the next level */
    tr->closest_hit.distance = tr->closest_hit.distance_squared = INFINITY;
    return trace_scene(ray, tr, sc, 0);
}

/* ---- shading ----------------------------------------------------------- */

/* common.h:86-95 */
static inline float tone_map(float c) {
    float x = c - 0.004f;
    if (x < 0.0f) x = 0.0f;
    float x2s = x * x * 6.2f;
    return (x2s + x * 0.5f) / (x2s + x * 1.7f + 0.06f);
}

/* common.h:97-113 */
static inline float inv_dot(V3 a, V3 b) { return clamp01(-dot3(a, b)); }
static inline V3 reflect_with_dot(V3 V, V3 N, float NdotV) { return smad3(N, -2 * NdotV, V); }
static inline V3 refract_dir(V3 V, V3 N, float n1_over_n2, float NdotV) {
    float c = n1_over_n2 * n1_over_n2 * (1 - (NdotV * NdotV));
    if (c > 1) return v3s(0);
    V = scl3(V, n1_over_n2);
    V = smad3(N, n1_over_n2 * -NdotV - sqrtf(1 - c), V);
    return norm3(V);
}

/* common.h:197-228 */
static V3 area_light_vector(const OPrim *p, V3 P, V3 *v) {
    if (p->scale.x == 0 || p->scale.z == 0) return v3s(0);
    V3 U = v3s(0), W = v3s(0);
    U.x = p->scale.x; if (U.x < 0) U.x = -U.x;
    W.z = p->scale.z; if (W.z < 0) W.z = -W.z;
    U = qrot(U, p->rotation);
    W = qrot(W, p->rotation);
    v[0] = sub3(sub3(p->position, U), W);
    v[1] = sub3(add3(p->position, U), W);
    v[2] = add3(add3(p->position, U), W);
    v[3] = add3(sub3(p->position, U), W);
    V3 u1n = norm3(sub3(v[0], P));
    V3 u2n = norm3(sub3(v[1], P));
    V3 u3n = norm3(sub3(v[2], P));
    V3 u4n = norm3(sub3(v[3], P));
    float a12 = acosf(dot3(u1n, u2n));
    float a23 = acosf(dot3(u2n, u3n));
    float a34 = acosf(dot3(u3n, u4n));
    float a41 = acosf(dot3(u4n, u1n));
    V3 out = scl3(cross3(u1n, u2n), a12);
    out = add3(out, scl3(cross3(u2n, u3n), a23));
    out = add3(out, scl3(cross3(u3n, u4n), a34));
    out = add3(out, scl3(cross3(u4n, u1n), a41));
    return scl3(out, 0.5f);
}

/* common.h:277-334 */
static inline float ggx_ndf(float r2, float NdotH) {
    float demon = NdotH * NdotH * (r2 - 1.0f) + 1.0f;
    return O_ONE_OVER_PI * r2 / (demon * demon);
}
static inline float ggx_smith_schlick(float NdotL, float NdotV, float roughness) {
    float k = roughness * 0.5f;
    float one_minus_k = 1.0f - k;
    float denom = NdotV * one_minus_k + k;
    float result = NdotV / fmaxf(denom, O_EPS);
    denom = NdotL * one_minus_k + k;
    return result * NdotL / fmaxf(denom, O_EPS);
}
static inline V3 ggx_fresnel(V3 refl, float LdotH) { return smad3(one_minus3(refl), powf(1.0f - LdotH, 5.0f), refl); }
static inline V3 shade_classic(V3 diffuse, float NdotL, V3 specular, float spec_factor, float exponent, float roughness) {
    float shininess = 1.0f - roughness;
    spec_factor = NdotL * powf(spec_factor, exponent * shininess) * shininess;
    return smad3(specular, spec_factor, diffuse);
}
static V3 shade_point(const OShaded *s, float NdotL) {
    const OMat *M = s->material;
    V3 diffuse = scl3(s->albedo, M->roughness * NdotL * O_ONE_OVER_PI);
    if (M->brdf == ST_BRDF_LAMBERT) return diffuse;
    if (M->brdf == ST_BRDF_PHONG) {
        float RdotL = dotc3(s->reflected_direction, s->light_direction);
        return RdotL ? shade_classic(diffuse, NdotL, M->reflectivity, RdotL, 4.0f, M->roughness) : diffuse;
    }
    V3 N = s->normal, L = s->light_direction, V = neg3(s->viewing_direction);
    V3 H = norm3(add3(L, V));
    float NdotH = dotc3(N, H);
    if (M->brdf == ST_BRDF_BLINN)
        return NdotH ? shade_classic(diffuse, NdotL, M->reflectivity, NdotH, 16.0f, M->roughness) : diffuse;
    V3 F = v3s(0);
    diffuse = scl3(s->albedo, (1.0f - M->metallic) * NdotL * O_ONE_OVER_PI);
    float LdotH = dotc3(L, H);
    if (LdotH != 1.0f) {
        F = ggx_fresnel(M->reflectivity, LdotH);
        diffuse = mul3(diffuse, one_minus3(F));
    }
    if (!M->roughness && NdotH == 1.0f) return diffuse;
    float NdotV = dotc3(N, V);
    if (!NdotV) return diffuse;
    float D = ggx_ndf(M->roughness * M->roughness, NdotH);
    float G = ggx_smith_schlick(NdotL, NdotV, M->roughness);
    return smad3(F, D * G / (4.0f * NdotV), diffuse);
}

/* common.h:336-339 */
static inline float duv_ray_cone(float NdotV, float cone_width, float area, float uv_area) {
    float pcw = cone_width / fabsf(NdotV);
    return sqrtf((pcw * pcw) * (uv_area / area));
}

/* scene/texture.h:5-22 */
static V3 sample_mip(OTrace *tr, const OMip *mip, V2 UV) {
    float u = UV.x, v = UV.y;
    if (u > 1) u -= (float)((uint32_t)u);
    if (v > 1) v -= (float)((uint32_t)v);
    const float U = u * (float)mip->width + 0.5f;
    const float Vv = v * (float)mip->height + 0.5f;
    const uint32_t x = (uint32_t)U, y = (uint32_t)Vv;
    const float r = U - (float)x, b = Vv - (float)y;
    const float l = 1 - r, t = 1 - b;
    const float tl = t * l * O_COLOR_TO_FLOAT, trr = t * r * O_COLOR_TO_FLOAT;
    const float bl = b * l * O_COLOR_TO_FLOAT, br = b * r * O_COLOR_TO_FLOAT;
    const uint8_t *q = mip->quads + 12 * ((size_t)y * (mip->width + 1) + x);
    tr->st.texel_fetches++;
    /* quad byte order: R{TL,TR,BL,BR}, G{...}, B{...} (core/types.h:56-62) */
    return v3((float)q[3] * br + ((float)q[2] * bl + ((float)q[1] * trr + (float)q[0] * tl)),
              (float)q[7] * br + ((float)q[6] * bl + ((float)q[5] * trr + (float)q[4] * tl)),
              (float)q[11] * br + ((float)q[10] * bl + ((float)q[9] * trr + (float)q[8] * tl)));
}
/* texture.h:24-40 */
static V3 sample_texture(OTrace *tr, const OTex *tex, V2 UV, V2 dUV) {
    uint8_t mip_level = 0;
    if (tex->mipmap) {
        const float pw = dUV.x * (float)tex->width;
        const float ph = dUV.y * (float)tex->height;
        float pixel_size = (float)(pw + ph) * 0.5f;
        const uint8_t last_mip = (uint8_t)(tex->mip_count - 1);
        while (pixel_size > 1 && mip_level < last_mip) { pixel_size /= 2; mip_level += 1; }
    }
    return sample_mip(tr, tex->mips + mip_level, UV);
}
/* closest_hit/surface.h:131-145 */
static V3 sample_normal(OTrace *tr, const OTex *tex, V2 uv, V2 dUV) {
    V3 n = sample_texture(tr, tex, uv, dUV);
    float y = n.z;
    n.z = n.y;
    n.y = y;
    return norm3(smad3(n, 2.0f, v3s(-1.0f)));
}
static Q normal_rotation(V3 n) {
    Q q;
    q.v.x = n.z * 0.25f;
    q.v.y = 0;
    q.v.z = n.x * 0.25f;
    q.w = n.y * 0.25f;
    return qnorm(q);
}

/* surface.h:9-35 */
static V3 shade_lights_direct(OShaded *s, ORay *ray, OTrace *tr, const OScene *sc, V3 color) {
    OHit *hit = &tr->closest_hit;
    for (uint32_t i = 0; i < sc->light_count; i++) {
        const OLight *light = sc->lights + i;
        tr->st.light_reads++;
        hit->position = light->position;
        s->light_direction = sub3(hit->position, s->position);
        hit->distance_squared = sqlen3(s->light_direction);
        hit->distance = sqrtf(hit->distance_squared);
        s->light_direction = scl3(s->light_direction, 1.0f / hit->distance);
        float NdotL = dot3(s->normal, s->light_direction);
        if (NdotL <= 0) continue;
        float intensity = light->intensity / hit->distance_squared;
        ray->origin = s->position;
        ray->direction = s->light_direction;
        if (in_shadow(ray, tr, sc)) continue;
        color = mad3(shade_point(s, NdotL), scl3(light->color, intensity), color);
    }
    return color;
}

/* surface.h:37-129 */
static int shade_emissive_quads(OShaded *s, ORay *ray, OTrace *tr, const OScene *sc, V3 *color) {
    OHit *hit = &tr->current_hit;
    V3 *Rd = &tr->local.direction, *Ro = &tr->local.origin;
    int found = 0;
    for (uint32_t i = 0; i < sc->prim_count; i++) {
        const OPrim *quad = sc->prims + i;
        const OMat *em = sc->mats + quad->material_id;
        if (quad->type != ST_PRIM_QUAD || !(em->is & ST_EMISSIVE)) continue;
        to_object(s->viewing_origin, s->viewing_direction, quad, Ro, Rd);
        if (hit_quad(hit, *Ro, *Rd, quad->flags)) {
            hit->position = pos_to_world(hit->position, quad);
            hit->distance_squared = sqlen3(sub3(hit->position, s->viewing_origin));
            if (hit->distance_squared < tr->closest_hit.distance_squared) {
                hit->object_id = i;
                hit->object_type = ST_PRIM_QUAD;
                hit->material_id = quad->material_id;
                tr->closest_hit = *hit;
                found = 1;
            }
        }
        V3 qn = qrot(v3(0, 1, 0), quad->rotation);
        s->light_direction = ray->direction = sub3(quad->position, s->position);
        if (dot3(qn, s->light_direction) >= 0) continue;
        float emission_intensity = dot3(s->normal, area_light_vector(quad, s->position, s->quad_vertices));
        if (emission_intensity > 0) {
            int skip = 1;
            for (int j = 0; j < 4; j++)
                if (dot3(s->normal, sub3(s->quad_vertices[j], s->position)) >= 0) { skip = 0; break; }
            if (skip) continue;
            float shaded_light = 1;
            for (uint32_t k = 0; k < sc->prim_count; k++) {
                const OPrim *sp = sc->prims + k;
                if (quad == s->primitive || quad == sp || dot3(qn, sub3(sp->position, quad->position)) <= 0) continue;
                to_object(s->position, s->light_direction, sp, Ro, Rd);
                *Ro = smad3(*Rd, O_TRACE_OFFSET, *Ro);
                float d = 1;
                if (sp->type == ST_PRIM_SPHERE) {
                    if (hit_sphere(hit, *Ro, *Rd, sp->flags))
                        d -= (1.0f - sqrtf(hit->distance_squared)) / (hit->distance * emission_intensity * 3);
                } else if (sp->type == ST_PRIM_QUAD) {
                    if (hit_quad(hit, *Ro, *Rd, sp->flags)) {
                        hit->position.y = 0;
                        hit->position.x = hit->position.x < 0 ? -hit->position.x : hit->position.x;
                        hit->position.z = hit->position.z < 0 ? -hit->position.z : hit->position.z;
                        if (hit->position.x > hit->position.z) {
                            hit->position.y = hit->position.z;
                            hit->position.z = hit->position.x;
                            hit->position.x = hit->position.y;
                            hit->position.y = 0;
                        }
                        d -= (1.0f - hit->position.z) / (hit->distance * emission_intensity);
                    }
                }
                if (d < shaded_light) shaded_light = d;
            }
            if (shaded_light > 0) {
                float NdotL = dotc3(s->normal, s->light_direction);
                if (NdotL)
                    *color = mad3(shade_point(s, NdotL), scl3(em->emission, emission_intensity * shaded_light), *color);
            }
        }
    }
    if (found) tr->closest_hit.distance = sqrtf(tr->closest_hit.distance_squared);
    return found;
}

/* closest_hit/lights.h:7-39 */
static inline float volume_density(const OSphereHit *h) {
    float t1 = h->t_near > 0 ? h->t_near : 0;
    float t2 = h->t_far < h->furthest ? h->t_far : h->furthest;
    return (h->c * t1 - h->b * t1 * t1 + t1 * t1 * t1 / 3.0f - (h->c * t2 - h->b * t2 * t2 + t2 * t2 * t2 / 3.0f)) * (3.0f / 4.0f);
}
static int shade_light_glow(const OScene *sc, V3 Ro, V3 Rd, float max_distance, OSphereHit *sh, V3 *color) {
    int hit_light = 0;
    sh->closest_hit_distance = INFINITY;
    for (uint32_t i = 0; i < sc->light_count; i++) {
        const OLight *light = sc->lights + i;
        float size = light->intensity * (1.0f / 16.0f);
        float one_over_radius = 8.0f / size;
        sh->furthest = max_distance * one_over_radius;
        if (hit_sphere_simple(Ro, Rd, light->position, one_over_radius, sh)) {
            float distance = sh->t_far - size * 8;
            float density = powf(volume_density(sh), 8) * 8;
            if (distance < sh->closest_hit_distance) {
                sh->closest_hit_distance = distance;
                sh->closest_hit_density = density;
            }
            *color = smad3(light->color, density, *color);
            hit_light = 1;
        }
    }
    return hit_light;
}

/* Texture lookups at a hit: surface.h:159-168 (both dUV components set) and
   :227-236 (only dUV.u set; dUV.v keeps its earlier value). */
static void apply_textures(OTrace *tr, const OScene *sc, OShaded *s, const OMat *M, V2 *dUV, int set_both) {
    OHit *hit = &tr->closest_hit;
    s->uv = v2(hit->uv.x * M->uv_repeat.x, hit->uv.y * M->uv_repeat.y);
    hit->uv_area /= M->uv_repeat.x / M->uv_repeat.y;
    float d = duv_ray_cone(hit->NdotV, hit->cone_width, hit->area, hit->uv_area);
    dUV->x = d;
    if (set_both) dUV->y = d;
    if (M->use & ST_ALBEDO_MAP) s->albedo = sample_texture(tr, sc->textures + M->texture_ids[0], s->uv, *dUV);
    if ((M->use & ST_NORMAL_MAP) && M->texture_count > 1) {
        Q rot = normal_rotation(sample_normal(tr, sc->textures + M->texture_ids[1], s->uv, *dUV));
        hit->normal = qrot(hit->normal, rot);
    }
}

/* surface.h:146-255 */
static V3 shade_surface(ORay *ray, OTrace *tr, const OScene *sc, int *lights_shaded) {
    OHit *hit = &tr->closest_hit;
    V3 color = v3s(0);
    const OMat *M = sc->mats + hit->material_id;
    tr->st.material_reads++;
    if (M->is & ST_EMISSIVE) return hit->from_behind ? color : M->emission;
    V2 dUV = v2(0, 0);
    OShaded s;
    memset(&s, 0, sizeof s);
    s.viewing_origin = ray->origin;
    s.viewing_direction = ray->direction;
    s.primitive = sc->prims + hit->object_id;
    s.material = M;
    s.albedo = M->albedo;
    if (M->use && M->texture_count) apply_textures(tr, sc, &s, M, &dUV, 1);

    float NdotRd = 0, ior, max_distance = hit->distance;
    int has_emissive_quads = 0;
    for (uint32_t i = 0; i < sc->prim_count; i++)
        if (sc->prims[i].type == ST_PRIM_QUAD && (sc->mats[sc->prims[i].material_id].is & ST_EMISSIVE)) {
            has_emissive_quads = 1;
            break;
        }
    V3 current_color, throughput = v3s(1);
    uint32_t depth = tr->depth;
    while (depth) {
        s.position = ray->origin = hit->position;
        s.normal = hit->from_behind ? neg3(hit->normal) : hit->normal;
        ior = hit->from_behind ? s.material->n2_over_n1 : s.material->n1_over_n2;
        int is_ref = (M->is & ST_REFLECTIVE) || (M->is & ST_REFRACTIVE);
        if (is_ref || M->brdf == ST_BRDF_PHONG) {
            NdotRd = -inv_dot(s.normal, s.viewing_direction);
            s.reflected_direction = reflect_with_dot(s.viewing_direction, s.normal, NdotRd);
        }
        current_color = is_ref ? v3s(0) : sc->ambient;
        if (sc->light_count) current_color = shade_lights_direct(&s, ray, tr, sc, current_color);
        if (has_emissive_quads)
            if (shade_emissive_quads(&s, ray, tr, sc, &current_color)) max_distance = hit->distance;
        color = mad3(current_color, throughput, color);
        if (sc->light_count && shade_light_glow(sc, s.viewing_origin, s.viewing_direction, max_distance, &tr->sphere_hit, &color))
            *lights_shaded = 1;
        if (is_ref && --depth) {
            if (M->is & ST_REFLECTIVE) {
                ray->direction = s.reflected_direction;
            } else {
                ray->direction = refract_dir(s.viewing_direction, s.normal, ior, NdotRd);
                if (ray->direction.x == 0 && ray->direction.y == 0 && ray->direction.z == 0)
                    ray->direction = s.reflected_direction;
            }
            if (trace_ray(ray, tr, sc)) {
                s.primitive = sc->prims + hit->object_id;
                s.material = M = sc->mats + hit->material_id;
                tr->st.material_reads++;
                s.albedo = M->albedo;
                if (M->use && M->texture_count) apply_textures(tr, sc, &s, M, &dUV, 0);
                if (hit->object_type == ST_PRIM_QUAD && (M->is & ST_EMISSIVE)) {
                    color = hit->from_behind ? v3s(0) : M->emission;
                    break;
                }
                if (M->brdf != ST_BRDF_COOK_TORRANCE) throughput = mul3(throughput, M->reflectivity);
                s.viewing_origin = ray->origin;
                s.viewing_direction = ray->direction;
                continue;
            }
        }
        break;
    }
    return color;
}

/* render/raytracer.h:81-134 (setPixel, core/types.h:422-452, reduces to an
   overwrite for opacity 1 over a cleared pixel). */
static void ray_trace(ORay *ray, OTrace *tr, const OScene *sc, uint32_t mode, StFloatPixel *pixel, StHitRec *hrec,
                      uint16_t x, uint16_t y, V3 cam_pos, Q cam_rot) {
    OHit *hit = &tr->closest_hit;
    V3 Ro = ray->origin, Rd = ray->direction;
    V3 color = v3s(0);
    int lights_shaded = 0;
    tr->st.primary_rays++;
    tr->ray_class = ST_ORC_PRIMARY;
    tr->level = 0;
    tr->primary_triangle = -1;
    int hit_found = hit_primitives(ray, tr, sc, sc->bvh_leaf_ids, sc->prim_count, 0, 1, x, y);
    if (hrec) {
        hrec->prim_id = hit_found ? (int32_t)hit->object_id : -1;
        hrec->object_type = hit_found ? (int32_t)hit->object_type : 0;
        hrec->triangle_id = hit_found ? tr->primary_triangle : -1;
        hrec->distance = hit_found ? hit->distance : INFINITY;
    }
    float closest_distance = hit_found ? hit->distance : INFINITY;
    float z = INFINITY;
    if (hit_found) {
        z = qrot(sub3(hit->position, cam_pos), cam_rot).z;
        switch (mode) {
            case ST_MODE_BEAUTY: color = shade_surface(ray, tr, sc, &lights_shaded); break;
            case ST_MODE_DEPTH: color = v3s(4 / hit->distance); break;
            case ST_MODE_NORMALS: {
                const OMat *M = sc->mats + hit->material_id;
                if (M->texture_count > 1 && (M->use & ST_NORMAL_MAP)) {
                    V2 uv = v2(hit->uv.x * M->uv_repeat.x, hit->uv.y * M->uv_repeat.y);
                    hit->uv_area /= M->uv_repeat.x / M->uv_repeat.y;
                    V2 dUV = v2(0, 0);
                    dUV.x = duv_ray_cone(hit->NdotV, hit->cone_width, hit->area, hit->uv_area);
                    Q rot = normal_rotation(sample_normal(tr, sc->textures + M->texture_ids[1], uv, dUV));
                    hit->normal = qrot(hit->normal, rot);
                }
                color = v3(0.5f * (hit->normal.x + 1), 0.5f * (hit->normal.y + 1), 0.5f * (hit->normal.z + 1));
            } break;
            case ST_MODE_UVS: color = v3(0.5f * hit->uv.x, 0.5f * hit->uv.y, 0.5f); break;
        }
    }
    if (mode == ST_MODE_BEAUTY && sc->light_count && !lights_shaded) {
        if (shade_light_glow(sc, Ro, Rd, closest_distance, &tr->sphere_hit, &color)) {
            hit_found = 1;
            z = tr->sphere_hit.closest_hit_distance;
        }
    }
    if (hit_found) {
        if (mode == ST_MODE_BEAUTY) {
            color.x = tone_map(color.x);
            color.y = tone_map(color.y);
            color.z = tone_map(color.z);
        }
        color = scl3(color, 255.0f);
        color = mul3(color, color);
        pixel->r = color.x;
        pixel->g = color.y;
        pixel->b = color.z;
        pixel->opacity = 1;
        pixel->depth = (double)z;
    }
}

/* ---- camera and screen-space bounds ------------------------------------ */

/* core/base.h:404-411 */
static inline V2 unit_circle(float t) {
    float t2 = t * t;
    float f = 1 / (1 + t2);
    return v2(f - f * t2, f * 2 * t);
}

static M3 mat3_identity(void) { M3 m; m.X = v3(1, 0, 0); m.Y = v3(0, 1, 0); m.Z = v3(0, 0, 1); return m; }

/* math/mat3.h:79-91 */
static M3 mat3_mul(M3 a, M3 b) {
    M3 o;
    o.X.x = a.X.x * b.X.x + a.X.y * b.Y.x + a.X.z * b.Z.x;
    o.X.y = a.X.x * b.X.y + a.X.y * b.Y.y + a.X.z * b.Z.y;
    o.X.z = a.X.x * b.X.z + a.X.y * b.Y.z + a.X.z * b.Z.z;
    o.Y.x = a.Y.x * b.X.x + a.Y.y * b.Y.x + a.Y.z * b.Z.x;
    o.Y.y = a.Y.x * b.X.y + a.Y.y * b.Y.y + a.Y.z * b.Z.y;
    o.Y.z = a.Y.x * b.X.z + a.Y.y * b.Y.z + a.Y.z * b.Z.z;
    o.Z.x = a.Z.x * b.X.x + a.Z.y * b.Y.x + a.Z.z * b.Z.x;
    o.Z.y = a.Z.x * b.X.y + a.Z.y * b.Y.y + a.Z.z * b.Z.y;
    o.Z.z = a.Z.x * b.X.z + a.Z.y * b.Y.z + a.Z.z * b.Z.z;
    return o;
}
static M3 mat3_transpose(M3 m) {
    M3 o;
    o.X = v3(m.X.x, m.Y.x, m.Z.x);
    o.Y = v3(m.X.y, m.Y.y, m.Z.y);
    o.Z = v3(m.X.z, m.Y.z, m.Z.z);
    return o;
}
/* math/mat3.h:93-117 */
static M3 mat3_inverse(M3 m) {
    M3 o;
    float inv_det = 1.0f / (+m.X.x * (m.Y.y * m.Z.z - m.Z.y * m.Y.z)
                            - m.Y.x * (m.X.y * m.Z.z - m.Z.y * m.X.z)
                            + m.Z.x * (m.X.y * m.Y.z - m.Y.y * m.X.z));
    o.X.x = +(m.Y.y * m.Z.z - m.Z.y * m.Y.z) * inv_det;
    o.Y.x = -(m.Y.x * m.Z.z - m.Z.x * m.Y.z) * inv_det;
    o.Z.x = +(m.Y.x * m.Z.y - m.Z.x * m.Y.y) * inv_det;
    o.X.y = -(m.X.y * m.Z.z - m.Z.y * m.X.z) * inv_det;
    o.Y.y = +(m.X.x * m.Z.z - m.Z.x * m.X.z) * inv_det;
    o.Z.y = -(m.X.x * m.Z.y - m.Z.x * m.X.y) * inv_det;
    o.X.z = +(m.X.y * m.Y.z - m.Y.y * m.X.z) * inv_det;
    o.Y.z = -(m.X.x * m.Y.z - m.Y.x * m.X.z) * inv_det;
    o.Z.z = +(m.X.x * m.Y.y - m.Y.x * m.X.y) * inv_det;
    return o;
}
/* math/quat.h:53-107 */
static Q quat_from_matrix(M3 r) {
    Q o;
    float fx = r.X.x - r.Y.y - r.Z.z;
    float fy = r.Y.y - r.X.x - r.Z.z;
    float fz = r.Z.z - r.X.x - r.Y.y;
    float fw = r.X.x + r.Y.y + r.Z.z;
    int bi = 0;
    float fb = fw;
    if (fx > fb) { fb = fx; bi = 1; }
    if (fy > fb) { fb = fy; bi = 2; }
    if (fz > fb) { fb = fz; bi = 3; }
    float bv = sqrtf(fb + 1.0f) * 0.5f;
    float mult = 0.25f / bv;
    switch (bi) {
        case 0: o.w = bv; o.v.x = (r.Y.z - r.Z.y) * mult; o.v.y = (r.Z.x - r.X.z) * mult; o.v.z = (r.X.y - r.Y.x) * mult; break;
        case 1: o.w = (r.Y.z - r.Z.y) * mult; o.v.x = bv; o.v.y = (r.X.y + r.Y.x) * mult; o.v.z = (r.Z.x + r.X.z) * mult; break;
        case 2: o.w = (r.Z.x - r.X.z) * mult; o.v.x = (r.X.y + r.Y.x) * mult; o.v.y = bv; o.v.z = (r.Y.z + r.Z.y) * mult; break;
        default: o.w = (r.X.y - r.Y.x) * mult; o.v.x = (r.Z.x + r.X.z) * mult; o.v.y = (r.Y.z + r.Z.y) * mult; o.v.z = bv; break;
    }
    return o;
}

typedef struct {
    V3 position, forward, right, up;
    M3 rotation_matrix, rotation_matrix_inverted;
    Q rotation, rotation_inverted;
    float focal_length;
} OCamera;

typedef struct {
    uint16_t width, height; /* traced grid (QCAA: window + 1) */
    float f_width, f_height, h_width, h_height, width_over_height, height_over_width;
} ODims;

/* scene/xform.h:8-20 with the yaw/pitch/roll matrix updates of math/mat3.h:137-183. */
static void camera_init(OCamera *c, const StView *v) {
    M3 yaw = mat3_identity(), pitch = mat3_identity(), roll = mat3_identity();
    if (v->yaw) {
        V2 xy = unit_circle(v->yaw);
        M3 m = yaw;
        yaw.X.x = xy.x * m.X.x - xy.y * m.X.z;
        yaw.Y.x = xy.x * m.Y.x - xy.y * m.Y.z;
        yaw.Z.x = xy.x * m.Z.x - xy.y * m.Z.z;
        yaw.X.z = xy.x * m.X.z + xy.y * m.X.x;
        yaw.Y.z = xy.x * m.Y.z + xy.y * m.Y.x;
        yaw.Z.z = xy.x * m.Z.z + xy.y * m.Z.x;
    }
    if (v->pitch) {
        V2 xy = unit_circle(v->pitch);
        M3 m = pitch;
        pitch.X.y = xy.x * m.X.y + xy.y * m.X.z;
        pitch.Y.y = xy.x * m.Y.y + xy.y * m.Y.z;
        pitch.Z.y = xy.x * m.Z.y + xy.y * m.Z.z;
        pitch.X.z = xy.x * m.X.z - xy.y * m.X.y;
        pitch.Y.z = xy.x * m.Y.z - xy.y * m.Y.y;
        pitch.Z.z = xy.x * m.Z.z - xy.y * m.Z.y;
    }
    if (v->roll) {
        V2 xy = unit_circle(v->roll);
        M3 m = roll;
        roll.X.x = xy.x * m.X.x + xy.y * m.X.y;
        roll.Y.x = xy.x * m.Y.x + xy.y * m.Y.y;
        roll.Z.x = xy.x * m.Z.x + xy.y * m.Z.y;
        roll.X.y = xy.x * m.X.y - xy.y * m.X.x;
        roll.Y.y = xy.x * m.Y.y - xy.y * m.Y.x;
        roll.Z.y = xy.x * m.Z.y - xy.y * m.Z.x;
    }
    c->rotation_matrix = mat3_mul(mat3_mul(pitch, yaw), roll);
    c->rotation_matrix_inverted = mat3_transpose(c->rotation_matrix);
    c->rotation = quat_from_matrix(c->rotation_matrix);
    c->rotation_inverted = quat_from_matrix(c->rotation_matrix_inverted);
    c->right = c->rotation_matrix.X;
    c->up = c->rotation_matrix.Y;
    c->forward = c->rotation_matrix.Z;
    c->position = v3(v->position[0], v->position[1], v->position[2]);
    c->focal_length = v->focal_length;
}

/* core/base.h:344-358 with QCAA on */
static void dims_init(ODims *d, uint16_t w, uint16_t h) {
    d->width = (uint16_t)(w + 1);
    d->height = (uint16_t)(h + 1);
    d->f_width = (float)d->width;
    d->f_height = (float)d->height;
    d->h_width = d->f_width / 2;
    d->h_height = d->f_height / 2;
    d->width_over_height = d->f_width / d->f_height;
    d->height_over_width = d->f_height / d->f_width;
}

/* core/init.h:545-559 with the QCAA offset 0 */
static void projection_plane(const OCamera *c, const ODims *d, V3 *start, V3 *right, V3 *down) {
    const float offset = 0.0f;
    *start = scl3(c->forward, d->f_height * c->focal_length);
    *right = scl3(c->right, offset - (float)d->width);
    *down = scl3(c->up, d->f_height - offset);
    *start = add3(*start, *right);
    *start = add3(*start, *down);
    *right = scl3(c->right, 2);
    *down = scl3(c->up, -2);
}

/* render/SSB.h:5-135; (u16) casts truncate through 32-bit int. */
static int compute_ssb(int32_t *b, V3 pos, float r, float fl, const ODims *d) {
    b[0] = d->width + 1;
    b[2] = d->width + 1;
    b[1] = d->height + 1;
    b[3] = d->height + 1;
    if (pos.z <= -r) return 0;
    float x = pos.x, y = pos.y, z = pos.z;
    float left, right, top, bottom, factor;
    if (z <= r) {
        float h = fabsf(z) / fl;
        top = y + r;
        bottom = y - r;
        if (h < bottom || top < -h) return 0;
        float w = h * d->width_over_height;
        right = x + r;
        left = x - r;
        if (w < left || right < -w) return 0;
        b[0] = -w < left ? (uint16_t)(int32_t)(d->f_width * (left + w) / (2 * w)) : 0;
        b[2] = right < w ? (uint16_t)(int32_t)(d->f_width * (right + w) / (2 * w)) : d->width;
        b[1] = top < h ? (uint16_t)(int32_t)(d->f_height * (h - top) / (2 * h)) : 0;
        b[3] = -h < bottom ? (uint16_t)(int32_t)(d->f_height * (h - bottom) / (2 * h)) : d->height;
        return 1;
    }
    float den = z * z - r * r;
    factor = fl / den;
    float yz = y * z;
    float sqr = y * y + den;
    sqr = r * sqrtf(sqr);
    bottom = factor * (yz - sqr);
    top = factor * (yz + sqr);
    if (bottom < 1 && top > -1) {
        factor *= d->height_over_width;
        float xz = x * z;
        sqr = r * sqrtf(x * x + den);
        left = factor * (xz - sqr);
        right = factor * (xz + sqr);
        if (left < 1 && right > -1) {
            bottom = bottom > -1 ? bottom : -1; bottom += 1;
            top = top < 1 ? top : 1; top += 1;
            left = left > -1 ? left : -1; left += 1;
            right = right < 1 ? right : 1; right += 1;
            top = 2 - top;
            bottom = 2 - bottom;
            b[0] = (uint16_t)(int32_t)(d->h_width * left);
            b[2] = (uint16_t)(int32_t)(d->h_width * right);
            b[3] = (uint16_t)(int32_t)(d->h_height * bottom);
            b[1] = (uint16_t)(int32_t)(d->h_height * top);
            return 1;
        }
    }
    return 0;
}

/* SSB.h:137-173 */
static void update_ssb(OScene *sc, const OCamera *c, const ODims *d) {
    for (uint32_t i = 0; i < sc->prim_count; i++) {
        OPrim *p = sc->prims + i;
        V3 vsp = vmat3(sub3(p->position, c->position), c->rotation_matrix_inverted);
        float radius;
        switch (p->type) {
            case ST_PRIM_QUAD:
            case ST_PRIM_BOX: radius = len3(p->scale); break;
            case ST_PRIM_TETRAHEDRON: radius = maxc3(p->scale) * O_SQRT2; break;
            case ST_PRIM_SPHERE: radius = maxc3(p->scale); break;
            case ST_PRIM_MESH: {
                const Box3 *bb = &sc->meshes[p->mesh_id].aabb;
                float min_r = len3(mul3(p->scale, bb->min));
                float max_r = len3(mul3(p->scale, bb->max));
                radius = max_r > min_r ? max_r : min_r;
            } break;
            default: continue;
        }
        p->flags &= (uint8_t)~ST_IS_VISIBLE;
        if (compute_ssb(p->rect, vsp, radius, c->focal_length, d)) p->flags |= ST_IS_VISIBLE;
    }
}

/* ---- SAH builder (render/acceleration_structures/builder_top_down.h) ---- */

typedef struct {
    Box3 *leaf_boxes;       /* per leaf: box; leaf_first[i] = payload id */
    uint32_t *leaf_first;
    uint32_t *ids;          /* builder leaf ids */
    uint32_t *sorted[3];
    Box3 *lbox[3], *rbox[3];
    float *lsa[3], *rsa[3];
    uint32_t left_count[3];
    float smallest[3];
    int32_t *sort_stack;
    uint32_t *iter_start, *iter_end, *iter_node;
} OBuilder;

static float key_of(const OBuilder *b, uint32_t id, int axis) {
    const Box3 *bx = b->leaf_boxes + id;
    return axis == 0 ? bx->max.x : axis == 1 ? bx->max.y : bx->max.z;
}

/* builder_top_down.h:66-123: iterative Lomuto quicksort, pivot = last. */
static void sort_by_axis(OBuilder *b, int axis, int32_t start, int32_t end, uint32_t *ids) {
    int32_t *stack = b->sort_stack;
    int32_t top = -1;
    stack[++top] = start;
    stack[++top] = end;
    while (top >= 0) {
        end = stack[top--];
        start = stack[top--];
        int32_t p = start;
        uint32_t pivot = ids[end];
        float pk = key_of(b, pivot, axis);
        for (int32_t i = start; i < end; i++) {
            if (key_of(b, ids[i], axis) < pk) {
                uint32_t t = ids[i]; ids[i] = ids[p]; ids[p] = t;
                p++;
            }
        }
        uint32_t t = ids[end]; ids[end] = ids[p]; ids[p] = t;
        if (p - 1 > start) { stack[++top] = start; stack[++top] = p - 1; }
        if (p + 1 < end) { stack[++top] = p + 1; stack[++top] = end; }
    }
}

/* builder_top_down.h:125-166 */
static void partition_axis(OBuilder *b, int axis, uint32_t N) {
    uint32_t *sorted = b->sorted[axis];
    sort_by_axis(b, axis, 0, (int32_t)N - 1, sorted);
    uint32_t li = 0, ri = N - 1;
    Box3 L = b->leaf_boxes[sorted[li]], R = b->leaf_boxes[sorted[ri]];
    b->lbox[axis][li] = L;
    b->rbox[axis][ri] = R;
    for (li = 0; li < N; li++, ri--) {
        if (li) {
            L = merge_boxes(L, b->leaf_boxes[sorted[li]]);
            R = merge_boxes(R, b->leaf_boxes[sorted[ri]]);
            b->lbox[axis][li] = L;
            b->rbox[axis][ri] = R;
        }
        b->lsa[axis][li] = box_area(L);
        b->rsa[axis][ri] = box_area(R);
    }
    uint32_t last = N - 1, next = 1, lcount = 1, rcount = N - 1;
    b->left_count[axis] = 1;
    b->smallest[axis] = INFINITY;
    for (uint32_t cur = 0; cur < last; cur++, next++, lcount++, rcount--) {
        float sa = b->lsa[axis][cur] * (float)lcount + b->rsa[axis][next] * (float)rcount;
        if (b->smallest[axis] > sa) {
            b->smallest[axis] = sa;
            b->left_count[axis] = next;
        }
    }
}

static void node_reset(ONode *n) { memset(n, 0, sizeof *n); }

/* builder_top_down.h:168-202 */
static uint32_t split_node(OBuilder *b, ONode *nodes, uint32_t *node_count, ONode *node, uint32_t start, uint32_t end) {
    uint32_t N = end - start;
    uint32_t *ids = b->ids + start;
    node->first_child_id = *node_count;
    ONode *left = nodes + (*node_count)++;
    ONode *right = nodes + (*node_count)++;
    node_reset(left);
    node_reset(right);
    float smallest = INFINITY;
    int chosen = 0;
    for (int axis = 0; axis < 3; axis++) {
        memcpy(b->sorted[axis], ids, N * sizeof(uint32_t));
        partition_axis(b, axis, N);
        if (b->smallest[axis] < smallest) { smallest = b->smallest[axis]; chosen = axis; }
    }
    uint32_t lc = b->left_count[chosen];
    left->aabb = b->lbox[chosen][lc - 1];
    right->aabb = b->rbox[chosen][lc];
    memcpy(ids, b->sorted[chosen], N * sizeof(uint32_t));
    return start + lc;
}

static OBuilder *builder_new(uint32_t n) {
    OBuilder *b = (OBuilder *)calloc(1, sizeof(OBuilder));
    size_t m = n + 2;
    b->leaf_boxes = (Box3 *)calloc(m, sizeof(Box3));
    b->leaf_first = (uint32_t *)calloc(m, 4);
    b->ids = (uint32_t *)calloc(m, 4);
    for (int a = 0; a < 3; a++) {
        b->sorted[a] = (uint32_t *)calloc(m, 4);
        b->lbox[a] = (Box3 *)calloc(m, sizeof(Box3));
        b->rbox[a] = (Box3 *)calloc(m, sizeof(Box3));
        b->lsa[a] = (float *)calloc(m, 4);
        b->rsa[a] = (float *)calloc(m, 4);
    }
    b->sort_stack = (int32_t *)calloc(2 * m + 4, 4);
    b->iter_start = (uint32_t *)calloc(m, 4);
    b->iter_end = (uint32_t *)calloc(m, 4);
    b->iter_node = (uint32_t *)calloc(m, 4);
    return b;
}
static void builder_free(OBuilder *b) {
    if (!b) return;
    free(b->leaf_boxes); free(b->leaf_first); free(b->ids);
    for (int a = 0; a < 3; a++) { free(b->sorted[a]); free(b->lbox[a]); free(b->rbox[a]); free(b->lsa[a]); free(b->rsa[a]); }
    free(b->sort_stack); free(b->iter_start); free(b->iter_end); free(b->iter_node);
    free(b);
}

/* builder_top_down.h:204-270. leaf_boxes/leaf_first/ids already filled. */
static uint32_t build_bvh(OBuilder *b, uint32_t N, uint32_t max_leaf, ONode *nodes, uint32_t *leaf_ids, uint32_t *height) {
    uint32_t node_count = 1;
    node_reset(nodes);
    if (N <= max_leaf) {
        nodes->aabb.min = v3s(INFINITY);
        nodes->aabb.max = v3s(-INFINITY);
        for (uint32_t i = 0; i < N; i++) {
            leaf_ids[i] = b->leaf_first[i];
            nodes->aabb = merge_boxes(nodes->aabb, b->leaf_boxes[i]);
        }
        nodes->depth = 0;
        nodes->child_count = (uint16_t)N;
        nodes->first_child_id = 0;
        *height = 1;
        return node_count;
    }
    uint32_t middle = split_node(b, nodes, &node_count, nodes, 0, N);
    b->iter_start[0] = 0; b->iter_end[0] = middle; b->iter_node[0] = 1;
    b->iter_start[1] = middle; b->iter_end[1] = N; b->iter_node[1] = 2;
    int32_t depth = 1;
    uint32_t leaf_count = 0;
    *height = 1;
    while (depth >= 0) {
        uint32_t ls = b->iter_start[depth], le = b->iter_end[depth], ln = b->iter_node[depth];
        ONode *node = nodes + ln;
        uint32_t n = le - ls;
        if (n <= max_leaf) {
            node->depth = (uint16_t)depth;
            node->child_count = (uint16_t)n;
            node->first_child_id = leaf_count;
            for (uint32_t i = 0; i < n; i++) leaf_ids[leaf_count + i] = b->leaf_first[b->ids[ls + i]];
            leaf_count += n;
            depth--;
        } else {
            middle = split_node(b, nodes, &node_count, node, ls, le);
            nodes[node_count - 1].depth = nodes[node_count - 2].depth = (uint16_t)depth;
            b->iter_start[depth] = ls; b->iter_end[depth] = middle; b->iter_node[depth] = node->first_child_id;
            depth++;
            b->iter_start[depth] = middle; b->iter_end[depth] = le; b->iter_node[depth] = node->first_child_id + 1;
            if (depth > (int32_t)*height) *height = (uint32_t)depth;
        }
    }
    nodes->aabb = merge_boxes(nodes[1].aabb, nodes[2].aabb);
    return node_count;
}

/* Mesh builder used by scene_load: 0 = the reference's SAH (leaves <= 4),
   1 = the Morton LBVH with one triangle per leaf (the device mesh
   builder's specification), 2 = the binned SAH builder with one triangle
   per leaf (the library's host builder for large meshes). */
static int g_mesh_builder = 0;
static uint32_t lbvh_nodes(const Box3 *bx, uint32_t n, uint32_t max_leaf, ONode *on, uint32_t *leaf_ids,
                           uint32_t *height);
static uint32_t binned_nodes(const Box3 *bx, uint32_t n, uint32_t max_leaf, ONode *on, uint32_t *leaf_ids,
                             uint32_t *height);

/* builder_top_down.h:272-364 */
static int build_mesh(OMesh *mesh, const float *pos, const uint32_t *pidx, const float *nrm, const uint32_t *nidx,
                      const float *uvs, const uint32_t *uidx) {
    uint32_t T = mesh->triangle_count;
    OBuilder *b = builder_new(T);
    for (uint32_t i = 0; i < T; i++) {
        const float *v1 = pos + 3 * pidx[3 * i], *v2 = pos + 3 * pidx[3 * i + 1], *v3p = pos + 3 * pidx[3 * i + 2];
        Box3 *bx = b->leaf_boxes + i;
        bx->min.x = v2[0] < v3p[0] ? v2[0] : v3p[0];
        bx->min.y = v2[1] < v3p[1] ? v2[1] : v3p[1];
        bx->min.z = v2[2] < v3p[2] ? v2[2] : v3p[2];
        bx->max.x = v2[0] > v3p[0] ? v2[0] : v3p[0];
        bx->max.y = v2[1] > v3p[1] ? v2[1] : v3p[1];
        bx->max.z = v2[2] > v3p[2] ? v2[2] : v3p[2];
        bx->min.x = bx->min.x < v1[0] ? bx->min.x : v1[0];
        bx->min.y = bx->min.y < v1[1] ? bx->min.y : v1[1];
        bx->min.z = bx->min.z < v1[2] ? bx->min.z : v1[2];
        bx->max.x = bx->max.x > v1[0] ? bx->max.x : v1[0];
        bx->max.y = bx->max.y > v1[1] ? bx->max.y : v1[1];
        bx->max.z = bx->max.z > v1[2] ? bx->max.z : v1[2];
        float diff;
        diff = bx->max.x - bx->min.x; if (diff < 0) diff = -diff;
        if (diff < O_EPS) { bx->min.x -= O_EPS; bx->max.x += O_EPS; }
        diff = bx->max.y - bx->min.y; if (diff < 0) diff = -diff;
        if (diff < O_EPS) { bx->min.y -= O_EPS; bx->max.y += O_EPS; }
        diff = bx->max.z - bx->min.z; if (diff < 0) diff = -diff;
        if (diff < O_EPS) { bx->min.z -= O_EPS; bx->max.z += O_EPS; }
        b->leaf_first[i] = b->ids[i] = i;
    }
    mesh->nodes = (ONode *)calloc(2 * (size_t)T + 2, sizeof(ONode));
    mesh->leaf_ids = (uint32_t *)calloc(T + 1, 4);
    mesh->triangles = (OTri *)calloc(T + 1, sizeof(OTri));
    if (g_mesh_builder == 1) mesh->node_count = lbvh_nodes(b->leaf_boxes, T, 1, mesh->nodes, mesh->leaf_ids, &mesh->height);
    else if (g_mesh_builder == 2)
        mesh->node_count = binned_nodes(b->leaf_boxes, T, 1, mesh->nodes, mesh->leaf_ids, &mesh->height);
    else mesh->node_count = build_bvh(b, T, 4, mesh->nodes, mesh->leaf_ids, &mesh->height);
    uint32_t *order = b->ids;
    memcpy(order, mesh->leaf_ids, T * sizeof(uint32_t));
    for (uint32_t i = 0; i < T; i++) {
        uint32_t tid = order[i];
        mesh->leaf_ids[tid] = i;
        OTri *t = mesh->triangles + i;
        const float *a = pos + 3 * pidx[3 * tid], *bb = pos + 3 * pidx[3 * tid + 1], *c = pos + 3 * pidx[3 * tid + 2];
        V3 v1 = v3(a[0], a[1], a[2]), v2p = v3(bb[0], bb[1], bb[2]), v3v = v3(c[0], c[1], c[2]);
        M3 m;
        m.X = sub3(v3v, v1);
        m.Y = sub3(v2p, v1);
        m.Z = cross3(m.X, m.Y);
        t->area_of_parallelogram = len3(m.Z);
        m.Z = scl3(m.Z, 1.0f / t->area_of_parallelogram);
        t->world_to_tangent = mat3_inverse(m);
        t->normal = m.Z;
        t->position = v1;
        if (mesh->normals_count) {
            for (int k = 0; k < 3; k++) {
                const float *n = nrm + 3 * nidx[3 * tid + k];
                t->vertex_normals[k] = v3(n[0], n[1], n[2]);
            }
        }
        if (mesh->uvs_count) {
            V2 ua, ub, uc;
            const float *p0 = uvs + 2 * uidx[3 * tid], *p1 = uvs + 2 * uidx[3 * tid + 1], *p2 = uvs + 2 * uidx[3 * tid + 2];
            ua = v2(p0[0], p0[1]); ub = v2(p1[0], p1[1]); uc = v2(p2[0], p2[1]);
            t->uvs[0] = ua; t->uvs[1] = ub; t->uvs[2] = uc;
            t->area_of_uv = fabsf((ub.x - ua.x) * (uc.y - ua.y) - (uc.x - ua.x) * (ub.y - ua.y));
        }
    }
    builder_free(b);
    return 0;
}

/* scene/primitive.h:53-59 and render/AABB.h:44-84 */
static Box3 prim_world_box(const OScene *sc, const OPrim *p) {
    Box3 bx;
    if (p->type == ST_PRIM_MESH) bx = sc->meshes[p->mesh_id].aabb;
    else {
        bx.max = v3s(p->type == ST_PRIM_TETRAHEDRON ? (O_SQRT3 / 3) : 1);
        bx.min = neg3(bx.max);
    }
    float x = bx.min.x, y = bx.min.y, z = bx.min.z, X = bx.max.x, Y = bx.max.y, Z = bx.max.z;
    V3 corners[8] = {v3(x, y, z), v3(x, y, Z), v3(x, Y, z), v3(x, Y, Z), v3(X, y, z), v3(X, y, Z), v3(X, Y, z), v3(X, Y, Z)};
    x = y = z = +INFINITY;
    X = Y = Z = -INFINITY;
    for (int i = 0; i < 8; i++) {
        V3 q = pos_to_world(corners[i], p);
        if (q.x > X) X = q.x;
        if (q.y > Y) Y = q.y;
        if (q.z > Z) Z = q.z;
        if (q.x < x) x = q.x;
        if (q.y < y) y = q.y;
        if (q.z < z) z = q.z;
    }
    bx.min = v3(x, y, z);
    bx.max = v3(X, Y, Z);
    return bx;
}

/* builder_top_down.h:366-381 */
static void build_scene_bvh(OScene *sc) {
    uint32_t P = sc->prim_count;
    OBuilder *b = builder_new(P);
    for (uint32_t i = 0; i < P; i++) {
        b->leaf_boxes[i] = prim_world_box(sc, sc->prims + i);
        b->leaf_first[i] = b->ids[i] = i;
    }
    sc->bvh_nodes = (ONode *)calloc(2 * (size_t)P + 2, sizeof(ONode));
    sc->bvh_leaf_ids = (uint32_t *)calloc(P + 1, 4);
    sc->bvh_node_count = build_bvh(b, P, 2, sc->bvh_nodes, sc->bvh_leaf_ids, &sc->bvh_height);
    builder_free(b);
}

/* ---- blob loading ------------------------------------------------------ */
typedef struct { const uint8_t *p, *end; int bad; } OReader;
static const void *rd_take(OReader *r, size_t n) {
    if (r->bad || (size_t)(r->end - r->p) < n) { r->bad = 1; return NULL; }
    const void *o = r->p;
    size_t adv = (n + 3) & ~(size_t)3;
    r->p = (size_t)(r->end - r->p) < adv ? r->end : r->p + adv;
    return o;
}

static void scene_free(OScene *sc) {
    if (!sc) return;
    for (uint32_t m = 0; m < sc->mesh_count && sc->meshes; m++) {
        free(sc->meshes[m].nodes); free(sc->meshes[m].triangles); free(sc->meshes[m].leaf_ids);
    }
    for (uint32_t t = 0; t < sc->texture_count && sc->textures; t++) {
        for (uint32_t i = 0; i < sc->textures[t].mip_count; i++) free(sc->textures[t].mips[i].quads);
        free(sc->textures[t].mips);
    }
    free(sc->prims); free(sc->mats); free(sc->lights); free(sc->meshes); free(sc->textures);
    free(sc->bvh_nodes); free(sc->bvh_leaf_ids);
    free(sc);
}

static void node_from_rec(ONode *n, const StBVHNodeRec *q) {
    n->aabb.min = v3(q->min[0], q->min[1], q->min[2]);
    n->aabb.max = v3(q->max[0], q->max[1], q->max[2]);
    n->first_child_id = q->first_child_id;
    n->child_count = q->child_count;
    n->depth = q->depth;
}

static OScene *scene_load(const void *blob, size_t size) {
    OReader r = {(const uint8_t *)blob, (const uint8_t *)blob + size, 0};
    const StBlobHeader *h = (const StBlobHeader *)rd_take(&r, sizeof *h);
    if (!h || h->magic != ST_BLOB_MAGIC || h->version != ST_BLOB_VERSION) return NULL;
    OScene *sc = (OScene *)calloc(1, sizeof(OScene));
    sc->prim_count = h->prim_count;
    sc->material_count = h->material_count;
    sc->light_count = h->light_count;
    sc->mesh_count = h->mesh_count;
    sc->texture_count = h->texture_count;
    sc->ambient = v3(h->ambient[0], h->ambient[1], h->ambient[2]);
    const StPrimRec *pr = (const StPrimRec *)rd_take(&r, sizeof(StPrimRec) * h->prim_count);
    const StMaterialRec *mr = (const StMaterialRec *)rd_take(&r, sizeof(StMaterialRec) * h->material_count);
    const StLightRec *lr = (const StLightRec *)rd_take(&r, sizeof(StLightRec) * h->light_count);
    if (r.bad) { scene_free(sc); return NULL; }
    sc->prims = (OPrim *)calloc(h->prim_count + 1, sizeof(OPrim));
    for (uint32_t i = 0; i < h->prim_count; i++) {
        OPrim *p = sc->prims + i;
        p->position = v3(pr[i].position[0], pr[i].position[1], pr[i].position[2]);
        p->scale = v3(pr[i].scale[0], pr[i].scale[1], pr[i].scale[2]);
        p->rotation.v = v3(pr[i].rotation[0], pr[i].rotation[1], pr[i].rotation[2]);
        p->rotation.w = pr[i].rotation[3];
        p->type = pr[i].type;
        p->mesh_id = pr[i].mesh_id;
        p->material_id = (uint8_t)pr[i].material_id;
        p->flags = (uint8_t)pr[i].flags;
    }
    sc->mats = (OMat *)calloc(h->material_count + 1, sizeof(OMat));
    for (uint32_t i = 0; i < h->material_count; i++) {
        OMat *m = sc->mats + i;
        const StMaterialRec *q = mr + i;
        m->albedo = v3(q->albedo[0], q->albedo[1], q->albedo[2]);
        m->reflectivity = v3(q->reflectivity[0], q->reflectivity[1], q->reflectivity[2]);
        m->emission = v3(q->emission[0], q->emission[1], q->emission[2]);
        m->uv_repeat = v2(q->uv_repeat[0], q->uv_repeat[1]);
        m->metallic = q->metallic;
        m->roughness = q->roughness;
        m->n1_over_n2 = q->n1_over_n2;
        m->n2_over_n1 = q->n2_over_n1;
        m->brdf = q->brdf;
        m->is = q->is;
        m->use = q->use;
        m->texture_count = q->texture_count;
        memcpy(m->texture_ids, q->texture_ids, 16);
    }
    sc->lights = (OLight *)calloc(h->light_count + 1, sizeof(OLight));
    for (uint32_t i = 0; i < h->light_count; i++) {
        sc->lights[i].position = v3(lr[i].position[0], lr[i].position[1], lr[i].position[2]);
        sc->lights[i].color = v3(lr[i].color[0], lr[i].color[1], lr[i].color[2]);
        sc->lights[i].intensity = lr[i].intensity;
    }
    sc->meshes = (OMesh *)calloc(h->mesh_count + 1, sizeof(OMesh));
    for (uint32_t m = 0; m < h->mesh_count; m++) {
        const StMeshRec *q = (const StMeshRec *)rd_take(&r, sizeof *q);
        if (!q) break;
        OMesh *mesh = sc->meshes + m;
        mesh->aabb.min = v3(q->aabb[0], q->aabb[1], q->aabb[2]);
        mesh->aabb.max = v3(q->aabb[3], q->aabb[4], q->aabb[5]);
        mesh->vertex_count = q->vertex_count;
        mesh->triangle_count = q->triangle_count;
        mesh->normals_count = q->normals_count;
        mesh->uvs_count = q->uvs_count;
        const float *pos = (const float *)rd_take(&r, 12u * q->vertex_count);
        const uint32_t *pidx = (const uint32_t *)rd_take(&r, 12u * q->triangle_count);
        const float *nrm = NULL, *uvs = NULL;
        const uint32_t *nidx = NULL, *uidx = NULL;
        if (q->normals_count) {
            nrm = (const float *)rd_take(&r, 12u * q->normals_count);
            nidx = (const uint32_t *)rd_take(&r, 12u * q->triangle_count);
        }
        if (q->uvs_count) {
            uvs = (const float *)rd_take(&r, 8u * q->uvs_count);
            uidx = (const uint32_t *)rd_take(&r, 12u * q->triangle_count);
        }
        if (r.bad) break;
        if (q->prebuilt) {
            const StBVHNodeRec *nodes = (const StBVHNodeRec *)rd_take(&r, sizeof(StBVHNodeRec) * q->node_count);
            const StTriangleRec *tris = (const StTriangleRec *)rd_take(&r, sizeof(StTriangleRec) * q->triangle_count);
            const uint32_t *leaf = (const uint32_t *)rd_take(&r, 4u * q->triangle_count);
            if (r.bad) break;
            mesh->node_count = q->node_count;
            mesh->height = q->height;
            mesh->nodes = (ONode *)calloc(q->node_count + 1, sizeof(ONode));
            mesh->triangles = (OTri *)calloc(q->triangle_count + 1, sizeof(OTri));
            mesh->leaf_ids = (uint32_t *)calloc(q->triangle_count + 1, 4);
            for (uint32_t n = 0; n < q->node_count; n++) node_from_rec(mesh->nodes + n, nodes + n);
            for (uint32_t t = 0; t < q->triangle_count; t++) {
                const StTriangleRec *s = tris + t;
                OTri *o = mesh->triangles + t;
                o->world_to_tangent.X = v3(s->world_to_tangent[0], s->world_to_tangent[1], s->world_to_tangent[2]);
                o->world_to_tangent.Y = v3(s->world_to_tangent[3], s->world_to_tangent[4], s->world_to_tangent[5]);
                o->world_to_tangent.Z = v3(s->world_to_tangent[6], s->world_to_tangent[7], s->world_to_tangent[8]);
                o->position = v3(s->position[0], s->position[1], s->position[2]);
                o->normal = v3(s->normal[0], s->normal[1], s->normal[2]);
                for (int k = 0; k < 3; k++) {
                    o->vertex_normals[k] = v3(s->vertex_normals[3 * k], s->vertex_normals[3 * k + 1], s->vertex_normals[3 * k + 2]);
                    o->uvs[k] = v2(s->uvs[2 * k], s->uvs[2 * k + 1]);
                }
                o->area_of_parallelogram = s->area_of_parallelogram;
                o->area_of_uv = s->area_of_uv;
                mesh->leaf_ids[t] = leaf[t];
            }
        } else {
            build_mesh(mesh, pos, pidx, nrm, nidx, uvs, uidx);
        }
    }
    sc->textures = (OTex *)calloc(h->texture_count + 1, sizeof(OTex));
    for (uint32_t t = 0; t < h->texture_count && !r.bad; t++) {
        const StTextureRec *q = (const StTextureRec *)rd_take(&r, sizeof *q);
        if (!q) break;
        OTex *tex = sc->textures + t;
        tex->width = q->width;
        tex->height = q->height;
        tex->mip_count = q->mip_count;
        tex->wrap = q->wrap;
        tex->mipmap = q->mipmap;
        tex->mips = (OMip *)calloc(q->mip_count + 1, sizeof(OMip));
        for (uint32_t i = 0; i < q->mip_count; i++) {
            const uint16_t *dims = (const uint16_t *)rd_take(&r, 4);
            if (!dims) break;
            size_t nq = (size_t)(dims[0] + 1) * (dims[1] + 1);
            const uint8_t *src = (const uint8_t *)rd_take(&r, 12 * nq);
            if (!src) break;
            tex->mips[i].width = dims[0];
            tex->mips[i].height = dims[1];
            tex->mips[i].quads = (uint8_t *)malloc(12 * nq);
            memcpy(tex->mips[i].quads, src, 12 * nq);
        }
    }
    if (!r.bad && (h->flags & ST_BLOB_HAS_SCENE_BVH)) {
        const uint32_t *cnt = (const uint32_t *)rd_take(&r, 8);
        if (cnt) {
            const StBVHNodeRec *nodes = (const StBVHNodeRec *)rd_take(&r, sizeof(StBVHNodeRec) * cnt[0]);
            const uint32_t *leaf = (const uint32_t *)rd_take(&r, 4u * h->prim_count);
            if (!r.bad) {
                sc->bvh_node_count = cnt[0];
                sc->bvh_height = cnt[1];
                sc->bvh_nodes = (ONode *)calloc(cnt[0] + 1, sizeof(ONode));
                sc->bvh_leaf_ids = (uint32_t *)calloc(h->prim_count + 1, 4);
                for (uint32_t n = 0; n < cnt[0]; n++) node_from_rec(sc->bvh_nodes + n, nodes + n);
                memcpy(sc->bvh_leaf_ids, leaf, 4u * h->prim_count);
            }
        }
    }
    if (r.bad) { scene_free(sc); return NULL; }
    if (!sc->bvh_nodes && sc->prim_count) build_scene_bvh(sc);
    return sc;
}

/* ---- frame driver (render/raytracer.h:136-171) -------------------------- */

typedef struct {
    const OScene *sc;
    const OCamera *cam;
    const ODims *dims;
    V3 start, right, down;
    uint32_t mode, depth;
    uint8_t mesh_stack_size;
    float cone_angle;
    StFloatPixel *pixels;
    StHitRec *hits;
    uint64_t *steps;       /* optional: per pixel, node pairs + triangle tests + primitive visits */
    int tid, nthreads;
    StOracleStats st;
} OJob;

static void *render_rows(void *arg) {
    OJob *j = (OJob *)arg;
    OTrace *tr = (OTrace *)calloc(1, sizeof(OTrace));
    tr->closest_mesh_hit.object_type = ST_PRIM_MESH;
    tr->depth = (uint8_t)j->depth;
    tr->mesh_stack_size = j->mesh_stack_size;
    V3 start = j->start;
    ORay ray;
    memset(&ray, 0, sizeof ray);
    for (uint16_t y = 0; y < j->dims->height; y++) {
        if (y % j->nthreads == j->tid) {
            V3 current = start;
            size_t row = (size_t)y * j->dims->width;
            for (uint16_t x = 0; x < j->dims->width; x++) {
                ray.origin = j->cam->position;
                ray.direction = norm3(current);
                ray.direction_reciprocal = rcp3(ray.direction);
                tr->closest_hit.distance = tr->closest_hit.distance_squared = INFINITY;
                tr->closest_hit.cone_angle = j->cone_angle;
                tr->closest_hit.cone_width = 0;
                const uint64_t e0 = tr->st.node_pairs + tr->st.triangle_tests + tr->st.primitive_visits;
                ray_trace(&ray, tr, j->sc, j->mode, j->pixels + row + x, j->hits ? j->hits + row + x : NULL, x, y,
                          j->cam->position, j->cam->rotation_inverted);
                if (j->steps)
                    j->steps[row + x] = tr->st.node_pairs + tr->st.triangle_tests + tr->st.primitive_visits - e0;
                current = add3(current, j->right);
            }
        }
        start = add3(start, j->down);
    }
    j->st = tr->st;
    free(tr);
    return NULL;
}

static uint8_t mesh_stack_size_of(const OScene *sc) {
    if (!sc->mesh_count) return 0;
    uint32_t max_depth = 0;
    for (uint32_t m = 0; m < sc->mesh_count; m++)
        if (sc->meshes[m].height > max_depth) max_depth = sc->meshes[m].height;
    return (uint8_t)(max_depth + 2); /* core/init.h:523-527 */
}

static void add_stats(StOracleStats *a, const StOracleStats *b) {
    uint64_t *x = (uint64_t *)a;
    const uint64_t *y = (const uint64_t *)b;
    for (size_t i = 0; i < sizeof(StOracleStats) / 8; i++) x[i] += y[i];
}

static void render_frame(OScene *sc, const StView *v, StFloatPixel *pixels, StHitRec *hits, int threads, StOracleStats *stats,
                         uint64_t *steps) {
    OCamera cam;
    ODims dims;
    camera_init(&cam, v);
    dims_init(&dims, v->width, v->height);
    update_ssb(sc, &cam, &dims);
    size_t n = (size_t)dims.width * dims.height;
    for (size_t i = 0; i < n; i++) {
        pixels[i].r = pixels[i].g = pixels[i].b = 0;
        pixels[i].opacity = 0;
        pixels[i].depth = INFINITY;
    }
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    OJob jobs[256];
    pthread_t th[256];
    for (int t = 0; t < threads; t++) {
        OJob *j = jobs + t;
        memset(j, 0, sizeof *j);
        j->sc = sc; j->cam = &cam; j->dims = &dims;
        projection_plane(&cam, &dims, &j->start, &j->right, &j->down);
        j->mode = v->render_mode; j->depth = v->depth; j->cone_angle = v->cone_angle;
        j->mesh_stack_size = mesh_stack_size_of(sc);
        j->pixels = pixels; j->hits = hits; j->steps = steps; j->tid = t; j->nthreads = threads;
    }
    if (threads == 1) render_rows(jobs);
    else {
        for (int t = 0; t < threads; t++) pthread_create(th + t, NULL, render_rows, jobs + t);
        for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    }
    if (stats) {
        memset(stats, 0, sizeof *stats);
        for (int t = 0; t < threads; t++) add_stats(stats, &jobs[t].st);
    }
}

/* ---- exported API ------------------------------------------------------ */

int oracle_render(const void *blob, size_t size, const StView *view, StFloatPixel *pixels, StHitRec *hits,
                  int threads, StOracleStats *stats) {
    OScene *sc = scene_load(blob, size);
    if (!sc) return -1;
    render_frame(sc, view, pixels, hits, threads, stats, NULL);
    scene_free(sc);
    return 0;
}

int oracle_render_steps(const void *blob, size_t size, const StView *view, StFloatPixel *pixels, StHitRec *hits,
                        uint64_t *steps, int threads, StOracleStats *stats) {
    OScene *sc = scene_load(blob, size);
    if (!sc) return -1;
    render_frame(sc, view, pixels, hits, threads, stats, steps);
    scene_free(sc);
    return 0;
}

/* A scene loaded once and rendered from many views (the batched and
   served workloads render hundreds of requests over a few large scenes). */
void *oracle_scene_open(const void *blob, size_t size, uint32_t flags) {
    /* meshes without a prebuilt BVH: the reference SAH build, or with
       flags & 1 the Morton LBVH the device builds for large meshes, or with
       flags & 2 the binned SAH tree of the library's host builder */
    g_mesh_builder = (flags & 1u) ? 1 : (flags & 2u) ? 2 : 0;
    OScene *sc = scene_load(blob, size);
    g_mesh_builder = 0;
    return sc;
}
void oracle_scene_close(void *scene) {
    if (scene) scene_free((OScene *)scene);
}
int oracle_scene_render(void *scene, const StView *view, StFloatPixel *pixels, StHitRec *hits, int threads,
                        StOracleStats *stats, double *seconds) {
    if (!scene) return -1;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    render_frame((OScene *)scene, view, pixels, hits, threads, stats, NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (seconds) *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    return 0;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int oracle_render_timed(const void *blob, size_t size, const StView *view, StFloatPixel *pixels, int threads,
                        int frames, double *seconds, StOracleStats *stats) {
    OScene *sc = scene_load(blob, size);
    if (!sc) return -1;
    for (int f = 0; f < frames; f++) {
        double t0 = now_s();
        render_frame(sc, view, pixels, NULL, threads, stats, NULL);
        seconds[f] = now_s() - t0;
    }
    scene_free(sc);
    return 0;
}

static void node_to_rec(StBVHNodeRec *o, const ONode *n) {
    o->min[0] = n->aabb.min.x; o->min[1] = n->aabb.min.y; o->min[2] = n->aabb.min.z;
    o->max[0] = n->aabb.max.x; o->max[1] = n->aabb.max.y; o->max[2] = n->aabb.max.z;
    o->first_child_id = n->first_child_id;
    o->child_count = n->child_count;
    o->depth = n->depth;
}

int oracle_scene_bvh(const void *blob, size_t size, StBVHNodeRec *nodes, uint32_t *leaf_ids, uint32_t *node_count,
                     uint32_t *height) {
    OScene *sc = scene_load(blob, size);
    if (!sc) return -1;
    for (uint32_t n = 0; n < sc->bvh_node_count; n++) node_to_rec(nodes + n, sc->bvh_nodes + n);
    memcpy(leaf_ids, sc->bvh_leaf_ids, 4u * sc->prim_count);
    *node_count = sc->bvh_node_count;
    *height = sc->bvh_height;
    scene_free(sc);
    return 0;
}

static int mesh_bvh_export(const void *blob, size_t size, uint32_t m, StBVHNodeRec *nodes, StTriangleRec *tris,
                           uint32_t *leaf_ids, uint32_t *node_count, uint32_t *height, int builder) {
    g_mesh_builder = builder;
    OScene *sc = scene_load(blob, size);
    g_mesh_builder = 0;
    if (!sc || m >= sc->mesh_count) { scene_free(sc); return -1; }
    const OMesh *mesh = sc->meshes + m;
    for (uint32_t n = 0; n < mesh->node_count; n++) node_to_rec(nodes + n, mesh->nodes + n);
    for (uint32_t t = 0; t < mesh->triangle_count; t++) {
        const OTri *s = mesh->triangles + t;
        StTriangleRec *o = tris + t;
        float w[9] = {s->world_to_tangent.X.x, s->world_to_tangent.X.y, s->world_to_tangent.X.z,
                      s->world_to_tangent.Y.x, s->world_to_tangent.Y.y, s->world_to_tangent.Y.z,
                      s->world_to_tangent.Z.x, s->world_to_tangent.Z.y, s->world_to_tangent.Z.z};
        memcpy(o->world_to_tangent, w, sizeof w);
        o->position[0] = s->position.x; o->position[1] = s->position.y; o->position[2] = s->position.z;
        o->normal[0] = s->normal.x; o->normal[1] = s->normal.y; o->normal[2] = s->normal.z;
        for (int k = 0; k < 3; k++) {
            o->vertex_normals[3 * k] = s->vertex_normals[k].x;
            o->vertex_normals[3 * k + 1] = s->vertex_normals[k].y;
            o->vertex_normals[3 * k + 2] = s->vertex_normals[k].z;
            o->uvs[2 * k] = s->uvs[k].x;
            o->uvs[2 * k + 1] = s->uvs[k].y;
        }
        o->area_of_parallelogram = s->area_of_parallelogram;
        o->area_of_uv = s->area_of_uv;
        leaf_ids[t] = mesh->leaf_ids[t];
    }
    *node_count = mesh->node_count;
    *height = mesh->height;
    scene_free(sc);
    return 0;
}

int oracle_mesh_bvh(const void *blob, size_t size, uint32_t m, StBVHNodeRec *nodes, StTriangleRec *tris,
                    uint32_t *leaf_ids, uint32_t *node_count, uint32_t *height) {
    return mesh_bvh_export(blob, size, m, nodes, tris, leaf_ids, node_count, height, 0);
}

int oracle_mesh_lbvh(const void *blob, size_t size, uint32_t m, StBVHNodeRec *nodes, StTriangleRec *tris,
                     uint32_t *leaf_ids, uint32_t *node_count, uint32_t *height) {
    return mesh_bvh_export(blob, size, m, nodes, tris, leaf_ids, node_count, height, 1);
}

int oracle_mesh_binned(const void *blob, size_t size, uint32_t m, StBVHNodeRec *nodes, StTriangleRec *tris,
                       uint32_t *leaf_ids, uint32_t *node_count, uint32_t *height) {
    return mesh_bvh_export(blob, size, m, nodes, tris, leaf_ids, node_count, height, 2);
}

int oracle_ssb(const void *blob, size_t size, const StView *view, int32_t *rects, uint8_t *flags) {
    OScene *sc = scene_load(blob, size);
    if (!sc) return -1;
    OCamera cam;
    ODims dims;
    camera_init(&cam, view);
    dims_init(&dims, view->width, view->height);
    update_ssb(sc, &cam, &dims);
    for (uint32_t i = 0; i < sc->prim_count; i++) {
        memcpy(rects + 4 * i, sc->prims[i].rect, 16);
        flags[i] = sc->prims[i].flags;
    }
    scene_free(sc);
    return 0;
}

/* manipulateSelection's first press (viewport/manipulation.h:55-124):
   the ray of setRayFromCoords (render/shaders/common.h:170-175, closed
   form start + right*x + down*y), hitPrimitives over all primitives with the
   SSB gate at the point, then every light's glow sphere (radius I/128) in
   front of the hit; the view-space z of the hit through the inverted
   rotation matrix (manipulation.h:117-118). */
int oracle_pick(const void *blob, size_t size, const StView *view, uint32_t n, const uint32_t *xy, StPick *out) {
    OScene *sc = scene_load(blob, size);
    if (!sc) return -1;
    OCamera cam;
    ODims dims;
    camera_init(&cam, view);
    dims_init(&dims, view->width, view->height);
    update_ssb(sc, &cam, &dims);
    V3 start, right, down;
    projection_plane(&cam, &dims, &start, &right, &down);
    OTrace *tr = (OTrace *)calloc(1, sizeof(OTrace));
    if (!tr) { scene_free(sc); return -1; }
    for (uint32_t k = 0; k < n; k++) {
        memset(tr, 0, sizeof *tr);
        tr->closest_mesh_hit.object_type = ST_PRIM_MESH;
        tr->depth = (uint8_t)view->depth;
        tr->mesh_stack_size = mesh_stack_size_of(sc);
        const int32_t x = (int32_t)xy[2 * k], y = (int32_t)xy[2 * k + 1];
        ORay ray;
        memset(&ray, 0, sizeof ray);
        ray.origin = cam.position;
        ray.direction = smad3(right, (float)x, start);
        ray.direction = smad3(down, (float)y, ray.direction);
        ray.direction = norm3(ray.direction);
        ray.direction_reciprocal = rcp3(ray.direction);
        OHit *hit = &tr->closest_hit;
        hit->distance = hit->distance_squared = INFINITY;
        int found = hit_primitives(&ray, tr, sc, sc->bvh_leaf_ids, sc->prim_count, 0, 1, (uint16_t)x, (uint16_t)y);
        for (uint32_t i = 0; i < sc->light_count; i++) {
            const OLight *l = sc->lights + i;
            const float light_radius = l->intensity * (1.0f / (8.0f * 16.0f));
            tr->sphere_hit.furthest = hit->distance / light_radius;
            if (hit_sphere_simple(ray.origin, ray.direction, l->position, 1.0f / light_radius, &tr->sphere_hit)) {
                hit->distance = tr->sphere_hit.t_near * light_radius;
                hit->position = smad3(ray.direction, hit->distance, ray.origin);
                hit->object_type = ST_PRIM_LIGHT;
                hit->object_id = i;
                found = 1;
            }
        }
        StPick *o = out + k;
        memset(o, 0, sizeof *o);
        if (found) {
            o->object_type = (int32_t)hit->object_type;
            o->object_id = hit->object_id;
            o->position[0] = hit->position.x;
            o->position[1] = hit->position.y;
            o->position[2] = hit->position.z;
            o->object_distance = vmat3(sub3(hit->position, ray.origin), cam.rotation_matrix_inverted).z;
        }
    }
    free(tr);
    scene_free(sc);
    return 0;
}

int oracle_camera(const StView *view, float *out) {
    OCamera cam;
    ODims dims;
    V3 s, r, d;
    camera_init(&cam, view);
    dims_init(&dims, view->width, view->height);
    projection_plane(&cam, &dims, &s, &r, &d);
    float v[17] = {s.x, s.y, s.z, r.x, r.y, r.z, d.x, d.y, d.z,
                   cam.rotation_inverted.v.x, cam.rotation_inverted.v.y, cam.rotation_inverted.v.z, cam.rotation_inverted.w,
                   cam.rotation.v.x, cam.rotation.v.y, cam.rotation.v.z, cam.rotation.w};
    memcpy(out, v, sizeof v);
    return 0;
}

uint32_t oracle_sah_build(const float *boxes, uint32_t n, uint32_t max_leaf, StBVHNodeRec *nodes, uint32_t *leaf_ids,
                          uint32_t *height) {
    OBuilder *b = builder_new(n);
    for (uint32_t i = 0; i < n; i++) {
        b->leaf_boxes[i].min = v3(boxes[6 * i], boxes[6 * i + 1], boxes[6 * i + 2]);
        b->leaf_boxes[i].max = v3(boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5]);
        b->leaf_first[i] = b->ids[i] = i;
    }
    ONode *on = (ONode *)calloc(2 * (size_t)n + 2, sizeof(ONode));
    uint32_t count = build_bvh(b, n, max_leaf, on, leaf_ids, height);
    for (uint32_t i = 0; i < count; i++) node_to_rec(nodes + i, on + i);
    free(on);
    builder_free(b);
    return count;
}

/* ---- Morton LBVH (no live reference: the reference's Z-curve builder,
   builder_bottom_up.h, is never compiled into the renderer) -------------
   Specification shared with the device builder:
     centroid c = (min + max) * 0.5f + 0.0f (the + 0 folds -0 to +0);
     bounds = componentwise min/max of the centroids;
     q = clamp((c - cmin) * (extent > 0 ? 1023 / extent : 0), 0, 1023) -> u32;
     code = interleave(qx, qy, qz) with x in the highest of each bit triple;
     primitives sorted by (code, original index);
     binary radix tree over the unique keys (code << 32 | sorted position);
     output: root = node 0, the two children of radix-tree internal node k
     are nodes 2k+1 and 2k+2, leaves hold one primitive (child_count 1,
     first_child_id = sorted position), boxes merged bottom-up as
     merge(left, right), depth = tree depth, height = levels.
   This restatement builds the radix tree top-down by splitting each range
   at the highest differing key bit; the device builds it bottom-up with
   Karras's per-node search (Karras 2012). The tree is unique, so both must
   produce the same nodes. */
static uint32_t expand10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

typedef struct { uint32_t code, index; } LKey;
static int lkey_cmp(const void *a, const void *b) {
    const LKey *x = (const LKey *)a, *y = (const LKey *)b;
    if (x->code != y->code) return x->code < y->code ? -1 : 1;
    return x->index < y->index ? -1 : x->index > y->index;
}

typedef struct {
    const LKey *keys;
    const Box3 *boxes;      /* by original index */
    ONode *nodes;
    uint32_t max_depth;
    uint32_t max_leaf;      /* ranges of at most this many become one leaf */
} LBuild;

static uint64_t lkey64(const LBuild *b, uint32_t pos) { return ((uint64_t)b->keys[pos].code << 32) | pos; }

/* last position in [first, last) whose key shares more than the common
   prefix of keys[first] and keys[last] with keys[first] */
static uint32_t lsplit(const LBuild *b, uint32_t first, uint32_t last) {
    uint64_t kf = lkey64(b, first), kl = lkey64(b, last);
    int common = __builtin_clzll(kf ^ kl);
    uint32_t split = first, step = last - first;
    do {
        step = (step + 1) >> 1;
        uint32_t cand = split + step;
        if (cand < last && __builtin_clzll(kf ^ lkey64(b, cand)) > common) split = cand;
    } while (step > 1);
    return split;
}

/* box of the radix-tree node over [first, last]: merge(left, right) down to
   single leaves, also inside ranges that become one leaf */
static Box3 lbox(const LBuild *b, uint32_t first, uint32_t last) {
    if (first == last) return b->boxes[b->keys[first].index];
    uint32_t g = lsplit(b, first, last);
    Box3 l = lbox(b, first, g), r = lbox(b, g + 1, last);
    return merge_boxes(l, r);
}

static Box3 lbuild(LBuild *b, uint32_t k, uint32_t first, uint32_t last, uint32_t out, uint32_t depth) {
    ONode *node = b->nodes + out;
    if (last - first + 1 <= b->max_leaf) {
        node->aabb = lbox(b, first, last);
        node->child_count = (uint16_t)(last - first + 1);
        node->first_child_id = first;
        node->depth = (uint16_t)depth;
        if (depth + 1 > b->max_depth) b->max_depth = depth + 1;
        return node->aabb;
    }
    uint32_t g = lsplit(b, first, last);
    node->first_child_id = 2 * k + 1;
    node->child_count = 0;
    node->depth = (uint16_t)depth;
    if (depth + 1 > b->max_depth) b->max_depth = depth + 1;
    Box3 bl, br;
    uint32_t lo = 2 * k + 1, ro = 2 * k + 2;
    if (first == g) {
        ONode *l = b->nodes + lo;
        l->aabb = bl = b->boxes[b->keys[g].index];
        l->child_count = 1; l->first_child_id = g; l->depth = (uint16_t)(depth + 1);
        if (depth + 2 > b->max_depth) b->max_depth = depth + 2;
    } else
        bl = lbuild(b, g, first, g, lo, depth + 1);
    if (g + 1 == last) {
        ONode *r = b->nodes + ro;
        r->aabb = br = b->boxes[b->keys[g + 1].index];
        r->child_count = 1; r->first_child_id = g + 1; r->depth = (uint16_t)(depth + 1);
        if (depth + 2 > b->max_depth) b->max_depth = depth + 2;
    } else
        br = lbuild(b, g + 1, g + 1, last, ro, depth + 1);
    node->aabb = merge_boxes(bl, br);
    return node->aabb;
}

/* Morton LBVH over n boxes; leaves hold up to max_leaf sorted positions.
   Writes 2n-1 node slots (slots under a range that became one leaf stay
   zero) and the sorted order; returns the slot count. */
static uint32_t lbvh_nodes(const Box3 *bx, uint32_t n, uint32_t max_leaf, ONode *on, uint32_t *leaf_ids,
                           uint32_t *height) {
    LKey *keys = (LKey *)calloc(n, sizeof(LKey));
    V3 cmin = v3s(INFINITY), cmax = v3s(-INFINITY);
    V3 *cent = (V3 *)calloc(n, sizeof(V3));
    for (uint32_t i = 0; i < n; i++) {
        V3 c = v3((bx[i].min.x + bx[i].max.x) * 0.5f + 0.0f, (bx[i].min.y + bx[i].max.y) * 0.5f + 0.0f,
                  (bx[i].min.z + bx[i].max.z) * 0.5f + 0.0f);
        cent[i] = c;
        cmin.x = c.x < cmin.x ? c.x : cmin.x; cmin.y = c.y < cmin.y ? c.y : cmin.y; cmin.z = c.z < cmin.z ? c.z : cmin.z;
        cmax.x = c.x > cmax.x ? c.x : cmax.x; cmax.y = c.y > cmax.y ? c.y : cmax.y; cmax.z = c.z > cmax.z ? c.z : cmax.z;
    }
    V3 ext = sub3(cmax, cmin);
    V3 sc = v3(ext.x > 0 ? 1023.0f / ext.x : 0.0f, ext.y > 0 ? 1023.0f / ext.y : 0.0f, ext.z > 0 ? 1023.0f / ext.z : 0.0f);
    for (uint32_t i = 0; i < n; i++) {
        float t[3] = {(cent[i].x - cmin.x) * sc.x, (cent[i].y - cmin.y) * sc.y, (cent[i].z - cmin.z) * sc.z};
        uint32_t q[3];
        for (int a = 0; a < 3; a++) {
            float v = t[a] < 0 ? 0 : t[a];
            v = v > 1023.0f ? 1023.0f : v;
            q[a] = (uint32_t)v;
        }
        keys[i].code = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
        keys[i].index = i;
    }
    qsort(keys, n, sizeof(LKey), lkey_cmp);
    for (uint32_t j = 0; j < n; j++) leaf_ids[j] = keys[j].index;
    uint32_t count;
    if (n == 1) {
        on[0].aabb = bx[0];
        on[0].child_count = 1;
        on[0].first_child_id = 0;
        *height = 1;
        count = 1;
    } else {
        LBuild b = {keys, bx, on, 0, max_leaf ? max_leaf : 1};
        lbuild(&b, 0, 0, n - 1, 0, 0);
        *height = b.max_depth;
        count = 2 * n - 1;
    }
    free(keys); free(cent);
    return count;
}

/* ---- binned SAH builder --------------------------------------------------
 * No counterpart in the reference: the specification of the library's host
 * builder for large meshes (a tree over leaf boxes in the reference's BVH
 * layout), written here from that specification as a recursion:
 *   centroid c = (min + max) * 0.5 per axis; a range of at most max_leaf
 *   boxes is a leaf (first_child_id = range start, child_count = size);
 *   per axis with centroid extent e > 0, 32 bins of width e / 32, bin =
 *   (int)((c - cmin) * (32 / e)) clamped to 31; candidate plane k (1..31)
 *   splits bins < k from bins >= k; cost = half_area(left) * n_left +
 *   half_area(right) * n_right with both sides non-empty (the bins' boxes
 *   merged upward for the left side, downward for the right); the first
 *   strictly smallest cost over axes x, y, z and planes 1..31 wins; the
 *   range is partitioned stably; with no positive extent it is halved in
 *   range order; children take the next pair of slots, left subtree first;
 *   node boxes merge the range in range order; depth 0 at the root, height
 *   = deepest node + 1.
 * Pinned by tree invariants, by the reference renderer's hits over it
 * equalling its hits over the reference SAH tree, and by SAH cost against
 * that tree (tests/test_binned_bvh.py). */
#define O_BINS 32

typedef struct {
    const Box3 *bx;
    V3 *cen;
    uint32_t *ids, *tmp;
    ONode *nodes;
    uint32_t count, max_leaf, max_depth;
} BinBuild;

static float v3_axis(V3 v, int a) { return a == 0 ? v.x : a == 1 ? v.y : v.z; }

static Box3 bin_range_box(const BinBuild *b, uint32_t s, uint32_t e) {
    Box3 r = b->bx[b->ids[s]];
    for (uint32_t i = s + 1; i < e; i++) r = merge_boxes(r, b->bx[b->ids[i]]);
    return r;
}

static uint32_t bin_index(float c, float lo, float scale) {
    int k = (int)((c - lo) * scale);
    return k > O_BINS - 1 ? O_BINS - 1 : (uint32_t)k;
}

static void binned_node(BinBuild *b, uint32_t node, uint32_t s, uint32_t e, uint32_t depth) {
    ONode *nd = b->nodes + node;
    nd->depth = (uint16_t)depth;
    if (depth > b->max_depth) b->max_depth = depth;
    if (e - s <= b->max_leaf) {
        nd->child_count = (uint16_t)(e - s);
        nd->first_child_id = s;
        return;
    }
    float lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = INFINITY;
        hi[a] = -INFINITY;
        for (uint32_t i = s; i < e; i++) {
            float c = v3_axis(b->cen[b->ids[i]], a);
            lo[a] = c < lo[a] ? c : lo[a];
            hi[a] = c > hi[a] ? c : hi[a];
        }
    }
    float best = INFINITY, best_scale = 0;
    int axis = -1;
    uint32_t plane = 0;
    for (int a = 0; a < 3; a++) {
        float ext = hi[a] - lo[a];
        if (!(ext > 0)) continue;
        float scale = 32.0f / ext;
        Box3 bins[O_BINS];
        uint32_t cnt[O_BINS] = {0};
        memset(bins, 0, sizeof bins);
        for (uint32_t i = s; i < e; i++) {
            uint32_t id = b->ids[i], k = bin_index(v3_axis(b->cen[id], a), lo[a], scale);
            bins[k] = cnt[k] ? merge_boxes(bins[k], b->bx[id]) : b->bx[id];
            cnt[k]++;
        }
        /* below[k]: bins 0..k-1 merged upward, above[k]: bins 31..k merged
           downward, with their box counts */
        Box3 below[O_BINS + 1], above[O_BINS + 1];
        uint32_t n_below[O_BINS + 1], n_above[O_BINS + 1];
        n_below[0] = 0;
        below[0] = bins[0];
        for (uint32_t j = 0; j < O_BINS; j++) {
            below[j + 1] = !cnt[j] ? below[j] : n_below[j] ? merge_boxes(below[j], bins[j]) : bins[j];
            n_below[j + 1] = n_below[j] + cnt[j];
        }
        n_above[O_BINS] = 0;
        above[O_BINS] = bins[0];
        for (uint32_t j = O_BINS; j-- > 0;) {
            above[j] = !cnt[j] ? above[j + 1] : n_above[j + 1] ? merge_boxes(above[j + 1], bins[j]) : bins[j];
            n_above[j] = n_above[j + 1] + cnt[j];
        }
        for (uint32_t k = 1; k < O_BINS; k++) {
            if (!n_below[k] || !n_above[k]) continue;
            float cost = box_area(below[k]) * (float)n_below[k] + box_area(above[k]) * (float)n_above[k];
            if (cost < best) { best = cost; axis = a; plane = k; best_scale = scale; }
        }
    }
    uint32_t mid;
    if (axis < 0) {
        mid = s + (e - s) / 2;
    } else {
        uint32_t nl = 0, nr = 0;
        for (uint32_t i = s; i < e; i++) {
            uint32_t id = b->ids[i];
            if (bin_index(v3_axis(b->cen[id], axis), lo[axis], best_scale) < plane) b->tmp[s + nl++] = id;
        }
        for (uint32_t i = s; i < e; i++) {
            uint32_t id = b->ids[i];
            if (bin_index(v3_axis(b->cen[id], axis), lo[axis], best_scale) >= plane) b->tmp[s + nl + nr++] = id;
        }
        memcpy(b->ids + s, b->tmp + s, (e - s) * sizeof(uint32_t));
        mid = s + nl;
    }
    uint32_t lc = b->count;
    b->count += 2;
    nd->first_child_id = lc;
    nd->child_count = 0;
    b->nodes[lc].aabb = bin_range_box(b, s, mid);
    b->nodes[lc + 1].aabb = bin_range_box(b, mid, e);
    binned_node(b, lc, s, mid, depth + 1);
    binned_node(b, lc + 1, mid, e, depth + 1);
}

static uint32_t binned_nodes(const Box3 *bx, uint32_t n, uint32_t max_leaf, ONode *on, uint32_t *leaf_ids,
                             uint32_t *height) {
    BinBuild b = {bx, (V3 *)calloc(n + 1, sizeof(V3)), leaf_ids, (uint32_t *)calloc(n + 1, 4), on, 1,
                  max_leaf ? max_leaf : 1, 0};
    for (uint32_t i = 0; i < n; i++) {
        b.cen[i] = v3((bx[i].min.x + bx[i].max.x) * 0.5f, (bx[i].min.y + bx[i].max.y) * 0.5f,
                      (bx[i].min.z + bx[i].max.z) * 0.5f);
        leaf_ids[i] = i;
    }
    on[0].aabb = bin_range_box(&b, 0, n);
    binned_node(&b, 0, 0, n, 0);
    *height = b.max_depth + 1;
    free(b.cen);
    free(b.tmp);
    return b.count;
}

uint32_t oracle_binned_build(const float *boxes, uint32_t n, uint32_t max_leaf, StBVHNodeRec *nodes,
                             uint32_t *leaf_ids, uint32_t *height) {
    if (n == 0) { *height = 0; return 0; }
    Box3 *bx = (Box3 *)calloc(n, sizeof(Box3));
    ONode *on = (ONode *)calloc(2 * (size_t)n + 1, sizeof(ONode));
    for (uint32_t i = 0; i < n; i++) {
        bx[i].min = v3(boxes[6 * i], boxes[6 * i + 1], boxes[6 * i + 2]);
        bx[i].max = v3(boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5]);
    }
    uint32_t count = binned_nodes(bx, n, max_leaf, on, leaf_ids, height);
    for (uint32_t i = 0; i < count; i++) node_to_rec(nodes + i, on + i);
    free(bx);
    free(on);
    return count;
}

uint32_t oracle_lbvh_build(const float *boxes, uint32_t n, StBVHNodeRec *nodes, uint32_t *leaf_ids, uint32_t *height) {
    if (n == 0) { *height = 0; return 0; }
    Box3 *bx = (Box3 *)calloc(n, sizeof(Box3));
    for (uint32_t i = 0; i < n; i++) {
        bx[i].min = v3(boxes[6 * i], boxes[6 * i + 1], boxes[6 * i + 2]);
        bx[i].max = v3(boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5]);
    }
    ONode *on = (ONode *)calloc(2 * (size_t)n, sizeof(ONode));
    uint32_t count = lbvh_nodes(bx, n, 1, on, leaf_ids, height);
    for (uint32_t i = 0; i < count; i++) node_to_rec(nodes + i, on + i);
    free(bx); free(on);
    return count;
}

int oracle_prim_boxes(const void *blob, size_t size, float *boxes) {
    OScene *sc = scene_load(blob, size);
    if (!sc) return -1;
    for (uint32_t i = 0; i < sc->prim_count; i++) {
        Box3 b = prim_world_box(sc, sc->prims + i);
        float v[6] = {b.min.x, b.min.y, b.min.z, b.max.x, b.max.y, b.max.z};
        memcpy(boxes + 6 * i, v, sizeof v);
    }
    scene_free(sc);
    return 0;
}

/* ---- display resolve --------------------------------------------------------
 * drawViewportToWindowContent (viewport/viewport.h:12-66), QCAA branch: each
 * window pixel averages the 4 traced grid corners around it,
 *   color = (TL.c*TL.o + (TR.c*TR.o + (BL.c*BL.o + BR.c*BR.o))) * 0.25
 * (scaleAddVec3 / scaleVec3 with fast_mul_add = multiply then add), and each
 * channel becomes (u8)(c > 255*255 ? 255 : sqrt(c)) with a double sqrt.
 * A pixel whose 4 corners all have depth INFINITY is the background
 * (0, 0, 0, 255). Bytes are B, G, R, A (core/base.h:206).
 * Defined here where the reference is not: alpha is 255 for every pixel (the
 * reference keeps the previous pixel's alpha, uninitialised until the first
 * background pixel), and a NaN or negative channel gives 0 (the reference's
 * (u8) cast of a NaN double is undefined; x86-64 yields 0).
 */
static uint8_t resolve_channel(float c) {
    if (c > 65025.0f) return 255;
    double s = sqrt((double)c);
    if (!(s >= 0.0)) return 0;
    return (uint8_t)s;
}

void oracle_resolve(const StFloatPixel *grid, uint32_t grid_w, uint32_t grid_h, uint32_t *out) {
    if (grid_w < 2 || grid_h < 2) return;
    uint32_t w = grid_w - 1, h = grid_h - 1;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            const StFloatPixel *tl = grid + (size_t)y * grid_w + x, *tr = tl + 1, *bl = tl + grid_w, *br = bl + 1;
            uint32_t px;
            if (tl->depth == INFINITY && tr->depth == INFINITY && bl->depth == INFINITY && br->depth == INFINITY) {
                px = 0xFF000000u;
            } else {
                float c[3];
                const float *a = &tl->r, *b = &tr->r, *d = &bl->r, *e = &br->r;
                for (int k = 0; k < 3; k++) {
                    float inner = e[k] * br->opacity;
                    inner = d[k] * bl->opacity + inner;
                    inner = b[k] * tr->opacity + inner;
                    inner = a[k] * tl->opacity + inner;
                    c[k] = inner * 0.25f;
                }
                px = (uint32_t)resolve_channel(c[2]) | ((uint32_t)resolve_channel(c[1]) << 8) |
                     ((uint32_t)resolve_channel(c[0]) << 16) | 0xFF000000u;
            }
            out[(size_t)y * w + x] = px;
        }
}
