/*
 * st_trace.h - per-ray device code of the tracing path for gfx950.
 *
 * The building blocks of one lane's ray: primary visibility over all
 * primitives in scene-BVH leaf order with the screen-space-bounds gate
 * (SlimTracin render/raytracer.h:81-134,
 * render/shaders/intersection/primitives.h:9-111), two-level BVH traversal
 * (render/shaders/trace.h:8-76, intersection/mesh.h:38-111), the implicit and
 * triangle intersection shaders, and the shading functions of
 * closest_hit/surface.h, lights.h and debug.h (the stages that sequence them
 * are in st_wave.h). Hit IDs must equal the reference CPU renderer's, so
 * every arithmetic step keeps the reference's operation order and the
 * reference's hit-state data flow (shadow and area-light passes overwrite
 * the closest hit, exactly as the reference's shared `Trace` does). The
 * values the reference leaves uninitialised (hitTetrahedron's local hit, an
 * unset dUV.v) are zero here, as in the CPU oracle.
 *
 * Memory: primitives, materials and lights are tiny and stay in L2/scalar
 * caches; BVH node pairs are one aligned 64-byte segment (four 16-byte
 * loads); a triangle test reads one aligned 64-byte record; per-hit
 * attributes come from a separate cold record.
 */
#ifndef ST_TRACE_H
#define ST_TRACE_H

#include <hip/hip_runtime.h>

#include "../st_math.h"
#include "../st_layout.h"
#include "st_blob.h"

#define D_EPS 0.0001f
#define D_SQRT3 1.73205080757f
#define D_PI 3.14159265f
#define D_ONE_OVER_PI (1.0f / D_PI)
#define D_UNIT_SPHERE_AREA_OVER_SIX ((4.0f * D_PI) / 6.0f)
#define D_COLOR_TO_FLOAT 0.00392156862f
#define D_TRACE_OFFSET 0.0001f
#define D_TET_A 0.577350259f
#define D_TET_B 0.288675159f

/* Everything is inlined into the stage kernels so the per-ray state stays in
   registers (a call boundary would force the shared trace state to scratch). */
#define DFN __device__ static inline __attribute__((always_inline))
#ifndef DNOINLINE
#define DNOINLINE DFN
#endif

struct DHit {
    sv3 position, normal;
    sv2 uv, dUV;
    float distance, distance_squared, cone_angle, cone_width, NdotV, area, uv_area;
    uint32_t material_id, object_id, object_type;
    uint32_t from_behind;
};

struct DSphereHit { float b, c, t_near, t_far, furthest, closest_hit_distance, closest_hit_density; };

struct DRay {
    sv3 origin, scaled_origin, direction, direction_reciprocal;
    uint32_t neg_x, neg_y, neg_z;   /* octant bits: direction sign */
};

struct DPrim {
    sq4 rot;
    sv3 pos, scale, inv_scale;
    uint32_t flags, material_id, type, mesh_id;
};

struct DTrace {
    DSphereHit sphere_hit;
    DHit closest_hit, closest_mesh_hit, current_hit;
    DRay local;
    uint32_t mesh_stack_size;
    int32_t primary_triangle;
    uint32_t *scene_stack, *mesh_stack;   /* lane columns of LDS: entry i at [i * ST_STACK_STRIDE] */
    uint32_t *aux_stack;                  /* wave-cooperative traversal: the path maxima beside mesh_stack */
    uint32_t abandoned;                   /* a mesh traversal gave up (StDScene.main_step_limit): this ray's
                                             result is not the reference's */
};

#define ST_STACK_STRIDE 64u

struct DShaded {
    sv3 albedo, position, normal, viewing_direction, viewing_origin, reflected_direction, light_direction;
    sv3 quad_vertices[4];
    sv2 uv;
    uint32_t prim_index;
};

struct DMat {
    sv3 albedo, reflectivity, emission;
    sv2 uv_repeat;
    float metallic, roughness, n1_over_n2, n2_over_n1;
    uint32_t brdf, is, use, texture_count, tex0, tex1;
};

/* ---- loads -------------------------------------------------------------- */

/* Scene data is read through global-address-space pointers. The scene
   record holds plain C pointers, and a load through a plain (generic)
   pointer is a flat load, which also counts against the LDS wait counter:
   every wait for a traversal-stack read would then wait for all outstanding
   node and triangle fetches too. */
#define ST_GLOBAL __attribute__((address_space(1)))
typedef float st_vf4 __attribute__((ext_vector_type(4)));
typedef int st_vi4 __attribute__((ext_vector_type(4)));

/* i-th 16-byte element at base */
DFN float4 ldg4(const void *base, size_t i) {
    st_vf4 v = ((const ST_GLOBAL st_vf4 *)base)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
DFN int4 ldgi4(const void *base, size_t i) {
    st_vi4 v = ((const ST_GLOBAL st_vi4 *)base)[i];
    return make_int4(v.x, v.y, v.z, v.w);
}
DFN uint32_t ldgu(const uint32_t *base, size_t i) { return ((const ST_GLOBAL uint32_t *)base)[i]; }
/* Loads through the constant address space: scalar loads when the index is
   wave-uniform (scene data is not written while a tracing kernel runs). */
#define ST_CONST __attribute__((address_space(4)))
DFN uint32_t ldcu(const uint32_t *base, size_t i) { return ((const ST_CONST uint32_t *)base)[i]; }
DFN float4 ldc4(const void *base, size_t i) {
    st_vf4 v = ((const ST_CONST st_vf4 *)base)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
typedef uint32_t st_vu2 __attribute__((ext_vector_type(2)));
/* i-th 8-byte element at base */
DFN void stg2(void *base, size_t i, uint32_t a, uint32_t b) {
    st_vu2 v = {a, b};
    ((ST_GLOBAL st_vu2 *)base)[i] = v;
}
/* i-th 16-byte element at base */
DFN void stgi4(void *base, size_t i, int a, int b, int c, int d) {
    st_vi4 v = {a, b, c, d};
    ((ST_GLOBAL st_vi4 *)base)[i] = v;
}
DFN StDMesh load_mesh(const StDScene &sc, uint32_t i) {
    int4 a = ldgi4(sc.meshes + i, 0);
    StDMesh m = {};
    m.node_offset = (uint32_t)a.x;
    m.tri_offset = (uint32_t)a.y;
    m.triangle_count = (uint32_t)a.z;
    m.attrs = (uint32_t)a.w;
    return m;
}
DFN StDTex load_tex(const StDScene &sc, uint32_t i) {
    const void *p = sc.textures + i;
    int4 a = ldgi4(p, 0), b = ldgi4(p, 1);
    StDTex t = {};
    t.width = (uint32_t)a.x;
    t.height = (uint32_t)a.y;
    t.mip_count = (uint32_t)a.z;
    t.mipmap = (uint32_t)a.w;
    t.mip_base = (uint32_t)b.x;
    return t;
}
DFN StDMip load_mip(const StDScene &sc, uint32_t i) {
    int4 a = ldgi4(sc.mips + i, 0);
    StDMip m;
    m.width = (uint32_t)a.x;
    m.height = (uint32_t)a.y;
    m.quad_offset = (uint64_t)(uint32_t)a.z | ((uint64_t)(uint32_t)a.w << 32);
    return m;
}

DFN DPrim prim_of(float4 a, float4 b, float4 c, float4 d) {
    DPrim o;
    o.rot.v = V3(a.x, a.y, a.z);
    o.rot.w = a.w;
    o.pos = V3(b.x, b.y, b.z);
    uint32_t fm = __float_as_uint(b.w);
    o.flags = fm & 0xFFu;
    o.material_id = fm >> 8;
    o.scale = V3(c.x, c.y, c.z);
    o.type = __float_as_uint(c.w);
    o.inv_scale = V3(d.x, d.y, d.z);
    o.mesh_id = __float_as_uint(d.w);
    return o;
}
DFN DPrim load_prim(const StDScene &sc, uint32_t i) {
    const void *p = sc.prims + i;
    return prim_of(ldg4(p, 0), ldg4(p, 1), ldg4(p, 2), ldg4(p, 3));
}
/* load_prim for a wave-uniform index */
DFN DPrim load_prim_c(const StDScene &sc, uint32_t i) {
    const void *p = sc.prims + i;
    return prim_of(ldc4(p, 0), ldc4(p, 1), ldc4(p, 2), ldc4(p, 3));
}

DFN DMat load_mat(const StDScene &sc, uint32_t i) {
    const void *p = sc.mats + i;
    float4 a = ldg4(p, 0), b = ldg4(p, 1), c = ldg4(p, 2), d = ldg4(p, 3), e = ldg4(p, 4);
    DMat m;
    m.albedo = V3(a.x, a.y, a.z);
    m.reflectivity = V3(a.w, b.x, b.y);
    m.emission = V3(b.z, b.w, c.x);
    m.uv_repeat = V2(c.y, c.z);
    m.metallic = c.w;
    m.roughness = d.x;
    m.n1_over_n2 = d.y;
    m.n2_over_n1 = d.z;
    m.brdf = __float_as_uint(d.w);
    uint32_t bits = __float_as_uint(e.x);
    m.is = bits & 0xFFu;
    m.use = (bits >> 8) & 0xFFu;
    m.texture_count = (bits >> 16) & 0xFFu;
    m.tex0 = __float_as_uint(e.y);
    m.tex1 = __float_as_uint(e.z);
    return m;
}

DFN uint32_t mat_is(const StDScene &sc, uint32_t i) { return ldgu(&sc.mats[i].bits, 0) & 0xFFu; }

struct DNodePair {
    sbox lb, rb;
    uint32_t lfirst, lcount, rfirst, rcount;
};

DFN void load_node(const uint32_t *nodes, uint32_t i, sbox &b, uint32_t &first, uint32_t &count) {
    const void *p = nodes + 8 * (size_t)i;
    float4 a = ldg4(p, 0), c = ldg4(p, 1);
    b.min = V3(a.x, a.y, a.z);
    b.max = V3(a.w, c.x, c.y);
    first = __float_as_uint(c.z);
    count = __float_as_uint(c.w) & 0xFFFFu;
}

DFN DNodePair load_pair(const uint32_t *nodes, uint32_t left) {
    const void *p = nodes + 8 * (size_t)left;
    float4 a = ldg4(p, 0), c = ldg4(p, 1), d = ldg4(p, 2), e = ldg4(p, 3);
    DNodePair o;
    o.lb.min = V3(a.x, a.y, a.z);
    o.lb.max = V3(a.w, c.x, c.y);
    o.lfirst = __float_as_uint(c.z);
    o.lcount = __float_as_uint(c.w) & 0xFFFFu;
    o.rb.min = V3(d.x, d.y, d.z);
    o.rb.max = V3(d.w, e.x, e.y);
    o.rfirst = __float_as_uint(e.z);
    o.rcount = __float_as_uint(e.w) & 0xFFFFu;
    return o;
}

/* ---- ray set-up and slab test (core/init.h:536-543, render/AABB.h:6-23) -- */

DFN void prep_ray(DRay &r) {
    r.neg_x = __builtin_signbit(r.direction.x) ? 1u : 0u;
    r.neg_y = __builtin_signbit(r.direction.y) ? 1u : 0u;
    r.neg_z = __builtin_signbit(r.direction.z) ? 1u : 0u;
    r.scaled_origin.x = -r.origin.x * r.direction_reciprocal.x;
    r.scaled_origin.y = -r.origin.y * r.direction_reciprocal.y;
    r.scaled_origin.z = -r.origin.z * r.direction_reciprocal.z;
}

DFN bool hit_aabb(const sbox &b, const DRay &r, float closest, float &dist) {
    sv3 lo = V3(r.neg_x ? b.max.x : b.min.x, r.neg_y ? b.max.y : b.min.y, r.neg_z ? b.max.z : b.min.z);
    sv3 hi = V3(r.neg_x ? b.min.x : b.max.x, r.neg_y ? b.min.y : b.max.y, r.neg_z ? b.min.z : b.max.z);
    lo = v_mad(lo, r.direction_reciprocal, r.scaled_origin);
    hi = v_mad(hi, r.direction_reciprocal, r.scaled_origin);
    float h = hi.x < hi.y ? hi.x : hi.y;
    h = h < hi.z ? h : hi.z;
    h = h < closest ? h : closest;
    float l = lo.x > lo.y ? lo.x : lo.y;
    l = l > lo.z ? l : lo.z;
    l = l > 0 ? l : 0;
    dist = l;
    return l <= h;
}

/* ---- instance transforms (scene/primitive.h:8-51) --------------------- */

DFN void to_object(sv3 pos, sv3 dir, const DPrim &p, sv3 &opos, sv3 &odir) {
    opos = (p.flags & ST_IS_TRANSLATED) ? v_sub(pos, p.pos) : pos;
    if (p.flags & ST_IS_ROTATED) {
        sq4 inv = q_conj(p.rot);
        opos = q_rot(opos, inv);
        odir = q_rot(dir, inv);
    } else
        odir = dir;
    if (p.flags & ST_IS_SCALED) {
        opos = v_mul(opos, p.inv_scale);
        if (p.flags & ST_IS_SCALED_NON_UNIFORMLY) odir = v_norm(v_mul(odir, p.inv_scale));
    }
}
DFN sv3 pos_to_world(sv3 pos, const DPrim &p) {
    if (p.flags & ST_IS_SCALED) pos = v_mul(pos, p.scale);
    if (p.flags & ST_IS_ROTATED) pos = q_rot(pos, p.rot);
    if (p.flags & ST_IS_TRANSLATED) pos = v_add(pos, p.pos);
    return pos;
}
DFN sv3 pos_to_object(sv3 pos, const DPrim &p) {
    if (p.flags & ST_IS_TRANSLATED) pos = v_sub(pos, p.pos);
    if (p.flags & ST_IS_ROTATED) pos = q_rot(pos, q_conj(p.rot));
    if (p.flags & ST_IS_SCALED) pos = v_mul(pos, p.inv_scale);
    return pos;
}
DFN sv3 dir_to_world(sv3 dir, const DPrim &p) {
    if (p.flags & ST_IS_SCALED_NON_UNIFORMLY) dir = v_mul(dir, p.inv_scale);
    if (p.flags & ST_IS_ROTATED) dir = q_rot(dir, p.rot);
    return dir;
}

/* ---- intersection shaders --------------------------------------------- */

/* common.h:115-119 (float -> u8 through a 32-bit integer) */
DFN bool transparent_uv(sv2 uv) {
    uint32_t v = (uint32_t)(uint8_t)(int32_t)(uv.y * 4);
    uint32_t u = (uint32_t)(uint8_t)(int32_t)(uv.x * 4);
    return (v % 2 != 0) == (u % 2);
}

/* common.h:177-195 */
DFN bool hit_plane(sv3 po, sv3 pn, sv3 ro, sv3 rd, DHit &hit) {
    hit.NdotV = -v_dot(rd, pn);
    if (hit.NdotV == 0) return false;
    uint32_t facing = hit.NdotV > 0;
    sv3 RP = v_sub(po, ro);
    float RPdotN = v_dot(RP, pn);
    hit.from_behind = RPdotN > 0;
    if (hit.from_behind == facing) return false;
    hit.distance = fabsf(RPdotN / hit.NdotV);
    hit.position = v_scale(rd, hit.distance);
    hit.position = v_add(ro, hit.position);
    hit.normal = pn;
    return true;
}

/* intersection/quad.h:7-42 */
DFN bool hit_quad(DHit &hit, sv3 Ro, sv3 Rd, uint32_t flags) {
    if (Rd.y == 0) return false;
    if (Ro.y == 0) return false;
    hit.from_behind = Ro.y < 0;
    if (hit.from_behind == (uint32_t)(Rd.y < 0)) return false;
    hit.distance = fabsf(Ro.y / Rd.y);
    hit.position = v_smad(Rd, hit.distance, Ro);
    if (hit.position.x < -1 || hit.position.x > +1 || hit.position.z < -1 || hit.position.z > +1) return false;
    hit.uv.x = (hit.position.x + 1.0f) * 0.5f;
    hit.uv.y = (hit.position.z + 1.0f) * 0.5f;
    if ((flags & ST_IS_TRANSPARENT) && transparent_uv(hit.uv)) return false;
    hit.normal = V3(0, 1, 0);
    hit.NdotV = -Rd.y;
    hit.area = 4;
    hit.uv_area = 1;
    return true;
}

enum { D_SIDE_NONE = 0, D_TOP = 1, D_BOTTOM = 2, D_LEFT = 4, D_RIGHT = 8, D_FRONT = 16, D_BACK = 32 };

/* common.h:121-130 */
DFN int box_side(sv3 oct, int axis) {
    switch (axis) {
        case 0: return oct.x > 0 ? D_RIGHT : D_LEFT;
        case 3: return oct.x > 0 ? D_LEFT : D_RIGHT;
        case 1: return oct.y > 0 ? D_TOP : D_BOTTOM;
        case 4: return oct.y > 0 ? D_BOTTOM : D_TOP;
        case 2: return oct.z > 0 ? D_FRONT : D_BACK;
        default: return oct.z > 0 ? D_BACK : D_FRONT;
    }
}
/* common.h:132-168 */
DFN sv2 cube_uv(sv3 p, int side) {
    sv2 uv;
    switch (side) {
        case D_TOP: uv.x = p.x; uv.y = p.z; break;
        case D_BOTTOM: uv.x = -p.x; uv.y = -p.z; break;
        case D_LEFT: uv.x = -p.z; uv.y = p.y; break;
        case D_RIGHT: uv.x = p.z; uv.y = p.y; break;
        case D_FRONT: uv.x = p.x; uv.y = p.y; break;
        default: uv.x = -p.x; uv.y = p.y; break;
    }
    uv.x += 1;
    uv.y += 1;
    uv.x *= 0.5f;
    uv.y *= 0.5f;
    return uv;
}

/* intersection/box.h:7-72 */
DFN bool hit_box(DHit &hit, sv3 Ro, sv3 Rd, uint32_t flags) {
    sv3 rcp = v_rcp(Rd), oct;
    oct.x = __builtin_signbit(Rd.x) ? 1.0f : -1.0f;
    oct.y = __builtin_signbit(Rd.y) ? 1.0f : -1.0f;
    oct.z = __builtin_signbit(Rd.z) ? 1.0f : -1.0f;
    float t0 = (oct.x - Ro.x) * rcp.x;
    float t1 = (oct.y - Ro.y) * rcp.y;
    float t2 = (oct.z - Ro.z) * rcp.z;
    float t3 = (-oct.x - Ro.x) * rcp.x;
    float t4 = (-oct.y - Ro.y) * rcp.y;
    float t5 = (-oct.z - Ro.z) * rcp.z;
    int max_axis = t3 < t4 ? 3 : 4;
    float tmax_ax = max_axis == 3 ? t3 : t4;
    if (t5 < tmax_ax) { max_axis = 5; tmax_ax = t5; }
    float maxt = tmax_ax;
    if (maxt < 0) return false;
    int min_axis = t0 > t1 ? 0 : 1;
    float tmin_ax = min_axis == 0 ? t0 : t1;
    if (t2 > tmin_ax) { min_axis = 2; tmin_ax = t2; }
    float mint = tmin_ax;
    if (maxt < (mint > 0 ? mint : 0)) return false;
    hit.from_behind = mint < 0;
    if (hit.from_behind) { mint = maxt; min_axis = max_axis; }
    int side = box_side(oct, min_axis);
    hit.position = v_smad(Rd, mint, Ro);
    hit.uv = cube_uv(hit.position, side);
    if ((flags & ST_IS_TRANSPARENT) && transparent_uv(hit.uv)) {
        if (hit.from_behind) return false;
        side = box_side(oct, max_axis);
        hit.position = v_smad(Rd, maxt, Ro);
        hit.uv = cube_uv(hit.position, side);
        if (transparent_uv(hit.uv)) return false;
        hit.from_behind = 1;
    }
    hit.normal = V3s(0);
    switch (side) {
        case D_LEFT: hit.normal.x = -1.0f; hit.NdotV = +Rd.x; break;
        case D_RIGHT: hit.normal.x = +1.0f; hit.NdotV = -Rd.x; break;
        case D_BOTTOM: hit.normal.y = -1.0f; hit.NdotV = +Rd.y; break;
        case D_TOP: hit.normal.y = +1.0f; hit.NdotV = -Rd.y; break;
        case D_BACK: hit.normal.z = -1.0f; hit.NdotV = +Rd.z; break;
        case D_FRONT: hit.normal.z = +1.0f; hit.NdotV = -Rd.z; break;
        default: return false;
    }
    if (hit.from_behind) { hit.normal = v_neg(hit.normal); hit.NdotV = -hit.NdotV; }
    hit.area = 4;
    hit.uv_area = 1;
    return true;
}

/* intersection/sphere.h:7-42 */
DFN sv2 sphere_uv(sv3 d) {
    sv2 uv;
    float x = d.x, y = d.y, z = d.z;
    float z_over_x = x != 0 ? (z / x) : 2;
    float y_over_x = x != 0 ? (y / x) : 2;
    if (z_over_x <= 1 && z_over_x >= -1 && y_over_x <= 1 && y_over_x >= -1) {
        uv.x = z_over_x;
        uv.y = x > 0 ? y_over_x : -y_over_x;
    } else {
        float x_over_z = z != 0 ? (x / z) : 2;
        float y_over_z = z != 0 ? (y / z) : 2;
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
DFN bool hit_sphere_simple(sv3 Ro, sv3 Rd, sv3 target, float one_over_radius, DSphereHit &h) {
    sv3 rc = v_scale(v_sub(target, Ro), one_over_radius);
    h.b = v_dot(Rd, rc);
    h.c = v_dot(rc, rc) - 1;
    float d = h.b * h.b - h.c;
    if (d < 0) return false;
    d = sqrtf(d);
    h.t_near = h.b - d;
    h.t_far = h.b + d;
    return h.t_far > 0 && h.t_near < h.furthest;
}

/* sphere.h:61-97 */
DFN bool hit_sphere(DHit &hit, sv3 Ro, sv3 Rd, uint32_t flags) {
    float t_to_closest = -v_dot(Ro, Rd);
    if (t_to_closest <= 0) return false;
    hit.distance_squared = v_sqlen(Ro) - t_to_closest * t_to_closest;
    float delta_squared = 1 - hit.distance_squared;
    if (delta_squared <= 0) { hit.distance = t_to_closest; return false; }
    float delta = sqrtf(delta_squared);
    hit.distance = t_to_closest - delta;
    bool has_outer_hit = hit.distance > 0;
    if (has_outer_hit) {
        hit.position = v_smad(Rd, hit.distance, Ro);
        hit.uv = sphere_uv(hit.position);
        if ((flags & ST_IS_TRANSPARENT) && transparent_uv(hit.uv)) has_outer_hit = false;
    }
    if (!has_outer_hit) {
        hit.distance = t_to_closest + delta;
        hit.position = v_smad(Rd, hit.distance, Ro);
        hit.uv = sphere_uv(hit.position);
        if ((flags & ST_IS_TRANSPARENT) && transparent_uv(hit.uv)) return false;
    }
    hit.from_behind = !has_outer_hit;
    hit.normal = hit.position;
    hit.NdotV = -v_dot(hit.normal, Rd);
    hit.area = hit.uv_area = D_UNIT_SPHERE_AREA_OVER_SIX;
    return true;
}

/* intersection/tetrahedra.h:9-107 */
DFN sm3 tet_frame(int t) {
    const float A = D_TET_A, B = D_TET_B;
    sm3 m;
    if (t == 0) { m.X = V3(A, -B, -A); m.Y = V3(B, B, A); m.Z = V3(-B, A, -A); }
    else if (t == 1) { m.X = V3(B, B, A); m.Y = V3(-B, A, -A); m.Z = V3(A, -B, -A); }
    else if (t == 2) { m.X = V3(-B, A, -A); m.Y = V3(A, -B, -A); m.Z = V3(B, B, A); }
    else { m.X = V3(-A, B, A); m.Y = V3(B, B, A); m.Z = V3(B, -A, A); }
    return m;
}
DFN bool hit_tetrahedron(DHit &closest_hit, sv3 Ro, sv3 Rd, uint32_t flags) {
    bool found = false;
    float closest_distance = INFINITY;
    DHit hit = {};
    for (int t = 0; t < 4; t++) {
        sv3 tangent_pos = V3s(t == 3 ? D_TET_A : -D_TET_A);
        hit.normal = tangent_pos;
        if (t == 0) hit.normal.y = D_TET_A;
        else if (t == 1) hit.normal.x = D_TET_A;
        else if (t == 2) hit.normal.z = D_TET_A;
        else tangent_pos.y = -D_TET_A;
        if (!hit_plane(tangent_pos, hit.normal, Ro, Rd, hit)) continue;
        sm3 m = tet_frame(t);
        hit.position = v_smad(Rd, hit.distance, Ro);
        tangent_pos = v_sub(hit.position, tangent_pos);
        tangent_pos = v_mat(tangent_pos, m);
        if (tangent_pos.x < 0 || tangent_pos.y < 0 || tangent_pos.y + tangent_pos.x > 1) continue;
        hit.uv.x = tangent_pos.x;
        hit.uv.y = tangent_pos.y;
        if ((flags & ST_IS_TRANSPARENT) && transparent_uv(hit.uv)) continue;
        if (hit.distance < closest_distance) {
            closest_distance = hit.distance;
            hit.area = D_SQRT3 / 4.0f;
            hit.uv_area = 1;
            hit.NdotV = -v_dot(hit.normal, Rd);
            closest_hit = hit;
            found = true;
        }
    }
    return found;
}

/* intersection/mesh.h:5-37 for one 64-byte hot triangle record a, b, c, d
   (leaf-order index i): the plane hit into tr.current_hit, then, when it is
   nearer than the closest mesh hit and inside the triangle, the closest
   mesh hit. True when taken. */
DFN bool tri_test(DTrace &tr, float4 a, float4 b, float4 c, float4 d, uint32_t i) {
    DRay &ray = tr.local;
    DHit &hit = tr.current_hit;
    DHit &closest = tr.closest_mesh_hit;
    sv3 tpos = V3(c.y, c.z, c.w);
    sv3 tnrm = V3(d.x, d.y, d.z);
    if (!(hit_plane(tpos, tnrm, ray.origin, ray.direction, hit) && hit.distance < closest.distance)) return false;
    sm3 w;
    w.X = V3(a.x, a.y, a.z);
    w.Y = V3(a.w, b.x, b.y);
    w.Z = V3(b.z, b.w, c.x);
    sv3 UV = v_sub(hit.position, tpos);
    UV = v_mat(UV, w);
    if (UV.x < 0 || UV.y < 0 || (UV.x + UV.y) > 1) return false;
    closest.NdotV = hit.NdotV;
    closest.normal = hit.normal;
    closest.uv.x = UV.x;
    closest.uv.y = UV.y;
    closest.object_id = i;
    closest.from_behind = hit.from_behind;
    closest.distance = hit.distance;
    closest.position = hit.position;
    closest.area = 0;       /* areas are fetched from the cold record once the hit is final */
    closest.uv_area = 0;
    return true;
}

/* The triangles of the leaves one traversal step reached: the left leaf's
   range [f0, f0 + c0), then the right leaf's [f1, f1 + c1), the order in
   which the reference calls its per-leaf test (mesh.h:5-37). One loop covers
   both so that the test is emitted once per kernel; object_id is the
   triangle's leaf-order index. */
DFN bool hit_triangles(DTrace &tr, const StDTriHot *tris, uint32_t f0, uint32_t c0, uint32_t f1, uint32_t c1,
                       bool any_hit) {
    bool found = false;
    const uint32_t n = c0 + c1;
#pragma clang loop unroll(disable)
    for (uint32_t j = 0; j < n; j++) {
        const uint32_t i = j < c0 ? f0 + j : f1 + (j - c0);
        const void *p = tris + i;
        if (tri_test(tr, ldg4(p, 0), ldg4(p, 1), ldg4(p, 2), ldg4(p, 3), i)) {
            found = true;
            if (any_hit) break;
        }
    }
    return found;
}

/* hit_triangles with the records of up to four triangles loaded together
   before they are tested in order (the cooperative traversal, where a
   thread's memory latency is not hidden by other waves). */
#define ST_TRI_BATCH 4
DFN bool hit_triangles_batched(DTrace &tr, const StDTriHot *tris, uint32_t f0, uint32_t c0, uint32_t f1, uint32_t c1,
                               bool any_hit) {
    bool found = false;
    const uint32_t n = c0 + c1;
    for (uint32_t j = 0; j < n; j += ST_TRI_BATCH) {
        float4 r[ST_TRI_BATCH][4];
        uint32_t id[ST_TRI_BATCH];
#pragma unroll
        for (int k = 0; k < ST_TRI_BATCH; k++) {
            const uint32_t jj = j + (uint32_t)k;
            id[k] = jj < c0 ? f0 + jj : f1 + (jj - c0);
            if (jj < n) {
                const void *p = tris + id[k];
                r[k][0] = ldg4(p, 0);
                r[k][1] = ldg4(p, 1);
                r[k][2] = ldg4(p, 2);
                r[k][3] = ldg4(p, 3);
            }
        }
#pragma unroll
        for (int k = 0; k < ST_TRI_BATCH; k++) {
            if (j + (uint32_t)k >= n) break;
            if (tri_test(tr, r[k][0], r[k][1], r[k][2], r[k][3], id[k])) {
                found = true;
                if (any_hit) return true;
            }
        }
    }
    return found;
}

/* A traversal ran past a bound that no valid tree reaches: record it for
   the host (the launch then fails with an error) and stop. */
DFN void trav_fault(const StDScene &sc, uint32_t code) {
    if (sc.fault) atomicOr(sc.fault, code);
}

/* "not written" marker for a current-hit distance: a signalling-NaN
   pattern, which no arithmetic result can have */
#define ST_DIST_UNSET 0x7FA5A5A5u

/* Block-shared state of one cooperative mesh traversal. */
struct DCoopShared {
    unsigned long long best;   /* closest hit: least (distance bits << 32 | key) over the units' hits;
                                  any hit: least key */
    uint32_t next;             /* next unit to claim */
    uint32_t status;           /* OR of the units' failures: 1 stack limit, 2 step bound */
    uint32_t first_key;        /* any hit: least key of a unit with a hit so far */
    uint32_t last_key;         /* nothing hit: greatest key + 1 of a unit with a plane hit */
    uint32_t tri;              /* the winning unit's triangle */
    float pm;                  /* its path maximum */
    float last_dist;           /* distance of the last plane hit of the latest unit with one */
};

/* The reference's mesh traversal of one ray (mesh.h:38-111), run by every
   thread of a block, all on the same ray (the redo stage). The top D levels
   of the tree (D = log2(triangles) - 5, clamped to 6..15) split it into
   2^D units: unit u descends along the bits of u (left = 0, from the root)
   and is the subtree it reaches, or a leaf its path ends at (then only the
   unit whose lower bits are zero holds it). Threads claim units from a
   block counter, which skips the rest of a range of units found empty (a
   box on their common path not admitted, or the leaf their path ends at),
   and run the reference's loop over each unit's subtree (same step
   structure and child order as trace_mesh) with the box bound fixed at the
   traversal's start, a superset of the boxes the sequential traversal
   admits, and a triangle test that starts afresh in each unit.
     Each loop iteration handles one node pair per thread, of a descent to a
   unit or of a unit's traversal, and a thread that ends its unit claims the
   next one in the same loop, so the threads of a wave do not wait for each
   other's units. With the bound fixed, a step's successor follows from its
   box tests alone: its node pair is loaded before the step's triangles are
   tested, and those are loaded together (hit_triangles_batched).
     A unit's order key is, per pair on its path, where the reference takes
   that child: leaf children in that step (left 0, right 1), then the inner
   child it enters first (2, the nearer one for a closest hit) and last (3);
   the reference visits the units in key order.
     any hit: the sequential traversal's bound stays fixed until its first
       hit, so its hit is the first hit of the unit with the least key
       (units whose key exceeds a key already hit are skipped or left);
     closest hit: t* = the least hit distance; among the hits at t*, the one
       of the least key is first in the reference's order. The reference
       reaches it if every box on its path admits it at bound t* (its bound
       is still above t* there), i.e. every entry distance on the path is at
       most t* (the boxes were admitted at the larger starting bound), and it
       then keeps it: nothing is nearer, and ties lose to the first.
   The winner's hit record is recomputed from its triangle with the same
   test. Nothing hit: the bound never moved, so the units are exactly what
   the reference visits, and its current-hit distance (the only field of
   that record a later result reads) is that of the last plane hit of the
   unit latest in key order that has one.
   Returns 1 (hit in tr.closest_mesh_hit), 0 (none), or -1 when the check
   fails, or a unit's stack would reach the reference's limit (never for a
   valid tree), and the caller traces sequentially (start state restored). */
DFN int coop_trace_mesh(DTrace &tr, const StDScene &sc, const StDMesh &mesh, const uint32_t *nodes,
                        const StDTriHot *tris, uint32_t root_first, bool any_hit) {
    __shared__ DCoopShared cs;
    volatile DCoopShared &vs = cs;
    DRay &ray = tr.local;
    DHit &closest = tr.closest_mesh_hit;
    uint32_t *stack = tr.mesh_stack, *pstack = tr.aux_stack;
    const float c0 = closest.distance, cur0 = tr.current_hit.distance;
    const uint32_t T = mesh.triangle_count;
    const uint32_t lg = 31u - (uint32_t)__builtin_clz(T);
    const uint32_t D = lg < 11u ? 6u : (lg > 20u ? 15u : lg - 5u);
    const uint32_t units = 1u << D;
    uint32_t max_steps = 2u * T + 2u;
    if (sc.step_cap && sc.step_cap < max_steps) max_steps = sc.step_cap;
    __syncthreads();  /* the previous traversal's readers of cs are done */
    if (threadIdx.x == 0) {
        cs.best = ~0ull;
        cs.next = 0;
        cs.status = 0;
        cs.first_key = 0xFFFFFFFFu;
        cs.last_key = 0;
    }
    __syncthreads();
    /* this thread's best unit result */
    bool found = false;
    uint32_t b_key = 0xFFFFFFFFu, b_tri = 0, l_key = 0;
    float b_t = INFINITY, b_pm = 0.0f, l_dist = 0.0f;
    /* its current unit: claiming (0), descending to it (1), traversing it (2) */
    enum { CLAIM, DESCEND, TRAVERSE };
    uint32_t phase = CLAIM, u = 0, d = 0, key = 0, left = 0, sp = 0, steps = 0;
    float pm = 0.0f, u_pm = 0.0f;
    bool u_found = false;
    DNodePair np = {};
    for (;;) {
        if (phase == CLAIM) {
            u = atomicAdd(&cs.next, 1u);
            if (u >= units || vs.status) break;
            phase = DESCEND;
            d = key = 0;
            pm = 0.0f;
            left = root_first;
            np = load_pair(nodes, left);
        }
        float ld, rd;
        const bool hl = hit_aabb(np.lb, ray, c0, ld);
        const bool hr = hit_aabb(np.rb, ray, c0, rd);
        const float pml = ld > pm ? ld : pm, pmr = rd > pm ? rd : pm;
        uint32_t f0 = 0, n0 = 0, f1 = 0, n1 = 0;
        bool next = false, unit_end = false;
        if (phase == DESCEND) {
            const uint32_t below = D - 1u - d;
            const uint32_t bit = (u >> below) & 1u;
            const uint32_t cnt = bit ? np.rcount : np.lcount, first = bit ? np.rfirst : np.lfirst;
            if (!(bit ? hr : hl)) {
                /* every unit below this box is empty */
                atomicMax(&cs.next, ((u >> below) + 1u) << below);
                phase = CLAIM;
                continue;
            }
            const uint32_t first_inner = (!any_hit && ld > rd) ? 1u : 0u;
            key = key * 4u + (cnt ? bit : (bit == first_inner ? 2u : 3u));
            pm = bit ? pmr : pml;
            if (cnt) {
                /* the leaf belongs to the first unit of its range */
                atomicMax(&cs.next, ((u >> below) + 1u) << below);
                if (u & ((1u << below) - 1u)) {
                    phase = CLAIM;
                    continue;
                }
                key <<= 2u * below;
                closest.distance = c0;
                tr.current_hit.distance = __uint_as_float(ST_DIST_UNSET);
                u_found = false;
                f0 = first;
                n0 = cnt;
                unit_end = true;
            } else {
                left = first;
                next = true;
                if (++d == D) {
                    phase = TRAVERSE;
                    sp = steps = 0;
                    closest.distance = c0;
                    tr.current_hit.distance = __uint_as_float(ST_DIST_UNSET);
                    u_found = false;
                }
            }
        } else {
            if (++steps > max_steps) {
                atomicOr(&cs.status, 2u);
                break;
            }
            if (any_hit && (steps & 31u) == 0 && vs.first_key < key) {
                /* a unit earlier in key order has a hit: this one is moot */
                phase = CLAIM;
                continue;
            }
            if (hl && np.lcount) { f0 = np.lfirst; n0 = np.lcount; }
            if (hr && np.rcount) { f1 = np.rfirst; n1 = np.rcount; }
            const bool il = hl && !np.lcount, ir = hr && !np.rcount;
            if (il && ir) {
                uint32_t a = np.lfirst, b = np.rfirst;
                float pa = pml, pb = pmr;
                if (!any_hit && ld > rd) {
                    a = np.rfirst; b = np.lfirst;
                    pa = pmr; pb = pml;
                }
                if (sp + 1u >= tr.mesh_stack_size) {
                    atomicOr(&cs.status, 1u);
                    break;
                }
                stack[ST_STACK_STRIDE * sp] = b;
                pstack[ST_STACK_STRIDE * sp] = __float_as_uint(pb);
                sp++;
                left = a;
                pm = pa;
                next = true;
            } else if (il) {
                left = np.lfirst;
                pm = pml;
                next = true;
            } else if (ir) {
                left = np.rfirst;
                pm = pmr;
                next = true;
            } else if (sp) {
                sp--;
                left = stack[ST_STACK_STRIDE * sp];
                pm = __uint_as_float(pstack[ST_STACK_STRIDE * sp]);
                next = true;
            } else {
                unit_end = true;
            }
        }
        /* the successor's pair is in flight while this step's triangles are tested */
        if (next) np = load_pair(nodes, left);
        if (n0 + n1 && hit_triangles_batched(tr, tris, f0, n0, f1, n1, any_hit)) {
            u_found = true;
            u_pm = phase == DESCEND ? pm : (closest.object_id - f0 < n0 ? pml : pmr);
            if (any_hit) unit_end = true;
        }
        if (unit_end) {
            if (u_found) {
                const float t = closest.distance;
                if (any_hit) {
                    if (key < b_key) {
                        b_key = key;
                        b_tri = closest.object_id;
                    }
                    atomicMin(&cs.first_key, key);
                } else if (t < b_t || (t == b_t && key < b_key)) {
                    b_t = t;
                    b_key = key;
                    b_tri = closest.object_id;
                    b_pm = u_pm;
                }
                found = true;
            } else {
                const float cd = tr.current_hit.distance;
                if (__float_as_uint(cd) != ST_DIST_UNSET && key + 1u > l_key) {
                    l_key = key + 1u;
                    l_dist = cd;
                }
            }
            phase = CLAIM;
        }
    }
    __syncthreads();
    closest.distance = c0;
    tr.current_hit.distance = cur0;
    const uint32_t status = vs.status;
    if (status) {
        if (status & 2u) {
            trav_fault(sc, ST_FAULT_MESH_STEPS);
            return 0;
        }
        return -1;
    }
    const unsigned long long mine =
        any_hit ? (unsigned long long)b_key : ((unsigned long long)__float_as_uint(b_t) << 32) | b_key;
    if (found) atomicMin(&cs.best, mine);
    if (l_key) atomicMax(&cs.last_key, l_key);
    __syncthreads();
    const unsigned long long best = vs.best;
    const uint32_t lk = vs.last_key;
    if (found && mine == best) {
        cs.tri = b_tri;
        cs.pm = b_pm;
    } else if (best == ~0ull && lk && l_key == lk) {
        cs.last_dist = l_dist;
    }
    __syncthreads();
    if (best != ~0ull) {
        const uint32_t tri = vs.tri;
        const float pm = vs.pm;
        if (!any_hit && !(pm <= __uint_as_float((uint32_t)(best >> 32)))) return -1;
        if (hit_triangles(tr, tris, tri, 1u, 0, 0, any_hit)) return 1;
        closest.distance = c0;
        tr.current_hit.distance = cur0;
        return -1;
    }
    if (lk) tr.current_hit.distance = vs.last_dist;
    return 0;
}

/* intersection/mesh.h:38-111. A leaf root holds every triangle; otherwise
   each step tests both children's boxes, then the triangles of the children
   that are leaves (left first), then descends. The leaf-root case runs
   through the same loop body (one step, no children) so that the triangle
   test and the traversal are emitted once. COOP (the redo stage, every
   thread of the block on the same ray): large meshes are traversed by the
   whole block (coop_trace_mesh), with the same result. */
template <bool COOP = false>
DNOINLINE bool trace_mesh(DTrace &tr, const StDScene &sc, uint32_t mesh_id, bool any_hit) {
    const StDMesh mesh = load_mesh(sc, mesh_id);
    const uint32_t *nodes = sc.mesh_nodes + 8 * (size_t)mesh.node_offset;
    const StDTriHot *tris = sc.tri_hot + mesh.tri_offset;
    DRay &ray = tr.local;
    DHit &closest = tr.closest_mesh_hit;
    uint32_t *stack = tr.mesh_stack;
    bool found = false;
    float ld, rd;
    sbox root;
    uint32_t root_first, root_count;
    load_node(nodes, 0, root, root_first, root_count);
    if (!hit_aabb(root, ray, closest.distance, ld)) return false;
    const bool root_leaf = root_count != 0;
    if (COOP && !root_leaf && mesh.triangle_count >= sc.coop_min_tris) {
        const int r = coop_trace_mesh(tr, sc, mesh, nodes, tris, root_first, any_hit);
        if (r >= 0) return r != 0;
        /* r < 0: the cooperative result could not be shown equal to the
           sequential one; trace sequentially */
    }
    uint32_t left = root_first;
    uint32_t stack_size = 0;
    /* each step loads one node pair, and a traversal loads a pair at most
       once: more steps than the mesh has nodes (at most 2T - 1) means a
       malformed tree, which would otherwise loop for ever */
    /* The main stages give up after main_step_limit steps (the pixel is
       redone). Each step loads one node pair, and a traversal loads a pair
       at most once, so the redo stage's bound, the mesh's node count, is
       never reached in a valid tree: a malformed one fails the launch
       instead of looping for ever. */
    uint32_t limit = sc.main_step_limit;
    if (COOP) {
        limit = 2u * mesh.triangle_count + 2u;
        if (sc.step_cap && sc.step_cap < limit) limit = sc.step_cap;
    }
    uint32_t steps = 0;
    for (;;) {
        if (++steps > limit) {
            if (COOP) trav_fault(sc, ST_FAULT_MESH_STEPS);
            else tr.abandoned = 1;
            found = false;
            break;
        }
        uint32_t f0 = 0, c0 = 0, f1 = 0, c1 = 0, lnext = 0, rnext = 0;
        bool has_left = false, has_right = false;
        if (root_leaf) {
            c0 = mesh.triangle_count;
        } else {
            DNodePair np = load_pair(nodes, left);
            const bool hl = hit_aabb(np.lb, ray, closest.distance, ld);
            const bool hr = hit_aabb(np.rb, ray, closest.distance, rd);
            lnext = np.lfirst;
            rnext = np.rfirst;
            if (hl && np.lcount) { f0 = np.lfirst; c0 = np.lcount; } else has_left = hl;
            if (hr && np.rcount) { f1 = np.rfirst; c1 = np.rcount; } else has_right = hr;
        }
        if (c0 + c1 && hit_triangles(tr, tris, f0, c0, f1, c1, any_hit)) {
            found = true;
            if (any_hit) break;
        }
        if (has_left) {
            if (has_right) {
                if (!any_hit && ld > rd) { uint32_t t = lnext; lnext = rnext; rnext = t; }
                stack[ST_STACK_STRIDE * stack_size++] = rnext;
                if (stack_size == tr.mesh_stack_size) return false;
            }
            left = lnext;
        } else if (has_right) {
            left = rnext;
        } else {
            if (stack_size == 0) break;
            left = stack[ST_STACK_STRIDE * --stack_size];
        }
    }
    return found;
}

/* Pixel box [x0, x1) x [y0, y1) of one primary-ray tile. */
struct DTileBox { int32_t x0, y0, x1, y1; };

/* intersection/primitives.h:9-111. `ids` are scene-BVH leaf ids, visited in
   that order. With `rects` (the primary pass) a primitive is tested only at
   pixels inside its screen-space rect. With `box` as well (called by the
   whole wave, every lane active, one tile per wave) the wave first tests 64
   primitives at a time against the tile's box and visits only those whose
   rect meets it, in the same order: the per-pixel test is unchanged, so
   the result is too, and a visited primitive's record is one scalar load
   per wave instead of a load per pixel for every primitive. */
template <bool COOP = false>
DNOINLINE bool hit_primitives(const DRay &ray, DTrace &tr, const StDScene &sc, const uint32_t *ids, uint32_t count,
                              bool any_hit, const int32_t *rects, uint32_t x, uint32_t y,
                              const DTileBox *box = nullptr) {
    bool found = false, cur = false;
    DHit &hit = tr.current_hit;
    DHit &closest = tr.closest_hit;
    float cone_angle = closest.cone_angle;
    uint32_t hit_prim_id = 0;
    /* one primitive; true ends the walk (any-hit search satisfied) */
    auto visit = [&](uint32_t pid, const DPrim &p) -> bool {
        if (any_hit && !(p.flags & ST_IS_SHADOWING)) return false;
        if (rects) {
            /* IS_VISIBLE is folded into the per-view rect: an invisible
               primitive has an empty rect. */
            int4 r = ldgi4(rects, pid);
            if ((int32_t)x < r.x || (int32_t)x >= r.z || (int32_t)y < r.y || (int32_t)y >= r.w) return false;
        }
        sv3 Ro, Rd;
        to_object(ray.origin, ray.direction, p, Ro, Rd);
        Ro = v_smad(Rd, D_TRACE_OFFSET, Ro);
        tr.local.origin = Ro;
        tr.local.direction = Rd;
        switch (p.type) {
            case ST_PRIM_QUAD: cur = hit_quad(hit, Ro, Rd, p.flags); break;
            case ST_PRIM_BOX: cur = hit_box(hit, Ro, Rd, p.flags); break;
            case ST_PRIM_SPHERE: cur = hit_sphere(hit, Ro, Rd, p.flags); break;
            case ST_PRIM_TETRAHEDRON: cur = hit_tetrahedron(hit, Ro, Rd, p.flags); break;
            case ST_PRIM_MESH:
                tr.local.direction_reciprocal = v_rcp(Rd);
                tr.closest_mesh_hit.distance = closest.distance == INFINITY
                    ? INFINITY : v_len(v_sub(pos_to_object(closest.position, p), Ro));
                prep_ray(tr.local);
                cur = trace_mesh<COOP>(tr, sc, p.mesh_id, any_hit);
                if (cur) hit = tr.closest_mesh_hit;
                break;
            default: return false;
        }
        if (cur) {
            hit.position = pos_to_world(hit.position, p);
            hit.distance_squared = v_sqlen(v_sub(hit.position, ray.origin));
            if (hit.distance_squared < closest.distance_squared) {
                closest = hit;
                hit_prim_id = pid;
                closest.object_type = p.type;
                closest.material_id = p.material_id;
                found = true;
                if (any_hit) return true;
            }
        }
        return false;
    };
    if (box) {
        bool done = false;
        for (uint32_t base = 0; base < count && !done; base += 64) {
            const uint32_t j = base + __lane_id();
            bool meets = false;
            if (j < count) {
                const int4 r = ldgi4(rects, ldgu(ids, j));
                meets = r.x < r.z && r.y < r.w && r.x < box->x1 && box->x0 < r.z && r.y < box->y1 && box->y0 < r.w;
            }
            for (uint64_t cand = __ballot(meets); cand; cand &= cand - 1) {
                const uint32_t pid = ldcu(ids, base + (uint32_t)__builtin_ctzll(cand));
                if (visit(pid, load_prim_c(sc, pid))) { done = true; break; }
            }
        }
    } else {
        for (uint32_t i = 0; i < count; i++) {
            const uint32_t pid = ldgu(ids, i);
            if (visit(pid, load_prim(sc, pid))) break;
        }
    }
    if (found) {
        DPrim hp = load_prim(sc, hit_prim_id);
        tr.primary_triangle = -1;
        if (closest.object_type == ST_PRIM_MESH) {
            const StDMesh mesh = load_mesh(sc, hp.mesh_id);
            uint32_t tri = closest.object_id;
            tr.primary_triangle = (int32_t)tri;
            const void *q = sc.tri_cold + mesh.tri_offset + tri;
            float4 c0 = ldg4(q, 0), c1 = ldg4(q, 1), c2 = ldg4(q, 2), c3 = ldg4(q, 3), c4 = ldg4(q, 4);
            /* the hot record carried no areas: restore the reference's
               closest_hit->area / uv_area (mesh.h:26-27). Cold record
               floats: vn 0-8, uv 9-14, area 15 (c3.w), uv_area 16 (c4.x) */
            closest.area = c3.w;
            closest.uv_area = c4.x;
            if (mesh.attrs) {
                float u = closest.uv.x, v = closest.uv.y;
                float w = 1 - u - v;
                if (mesh.attrs & 2u) {
                    /* uvs[0] = (c2.y, c2.z), uvs[1] = (c2.w, c3.x), uvs[2] = (c3.y, c3.z) */
                    closest.uv.x = c3.y * u + (c2.w * v + c2.y * w);
                    closest.uv.y = c3.z * u + (c3.x * v + c2.z * w);
                }
                if (mesh.attrs & 1u) {
                    /* vn[0] = (c0.x,c0.y,c0.z), vn[1] = (c0.w,c1.x,c1.y), vn[2] = (c1.z,c1.w,c2.x) */
                    closest.normal.x = c1.z * u + (c0.w * v + c0.x * w);
                    closest.normal.y = c1.w * u + (c1.x * v + c0.y * w);
                    closest.normal.z = c2.x * u + (c1.y * v + c0.z * w);
                    closest.normal = v_norm(closest.normal);
                }
            }
        }
        closest.object_id = hit_prim_id;
        closest.normal = v_norm(dir_to_world(closest.normal, hp));
        closest.distance = sqrtf(closest.distance_squared);
        closest.cone_width = 2.0f * cone_angle * closest.distance;
        sv3 rotated_scale = q_rot(hp.scale, hp.rot);
        closest.area *= v_dot(v_one_minus(closest.normal), rotated_scale);
    }
    return found;
}

/* trace.h:8-76. Same step structure as trace_mesh: both children's boxes,
   then the leaves among them (left first, each one primitive-test call as
   the reference makes it), then descent; the leaf-root case is one step of
   the same loop, so the primitive tests are emitted once. */
template <bool COOP = false>
DNOINLINE bool trace_scene(DRay &ray, DTrace &tr, const StDScene &sc, bool any_hit) {
    ray.direction_reciprocal = v_rcp(ray.direction);
    prep_ray(ray);
    bool found = false;
    float ld, rd;
    sbox root;
    uint32_t root_first, root_count;
    load_node(sc.scene_nodes, 0, root, root_first, root_count);
    if (!hit_aabb(root, ray, tr.closest_hit.distance, ld)) return false;
    const bool root_leaf = root_count != 0;
    uint32_t left = root_first;
    uint32_t *stack = tr.scene_stack;
    uint32_t stack_size = 0;
    /* bounds no valid tree reaches (see trace_mesh); the stack holds at
       most the tree's height, scene_stack_need - 1 entries */
    uint32_t max_steps = 2u * sc.prim_count + 2u;
    if (sc.step_cap && sc.step_cap < max_steps) max_steps = sc.step_cap;
    uint32_t steps = 0;
    for (;;) {
        if (++steps > max_steps) {
            trav_fault(sc, ST_FAULT_SCENE_STEPS);
            break;
        }
        uint32_t f0 = 0, c0 = 0, f1 = 0, c1 = 0, nleaf = 0, lnext = 0, rnext = 0;
        bool has_left = false, has_right = false;
        if (root_leaf) {
            c0 = sc.prim_count;
            nleaf = 1;
        } else {
            DNodePair np = load_pair(sc.scene_nodes, left);
            const bool hl = hit_aabb(np.lb, ray, tr.closest_hit.distance, ld);
            const bool hr = hit_aabb(np.rb, ray, tr.closest_hit.distance, rd);
            lnext = np.lfirst;
            rnext = np.rfirst;
            if (hl && np.lcount) { f0 = np.lfirst; c0 = np.lcount; nleaf = 1; } else has_left = hl;
            if (hr && np.rcount) {
                if (nleaf) { f1 = np.rfirst; c1 = np.rcount; } else { f0 = np.rfirst; c0 = np.rcount; }
                nleaf++;
            } else has_right = hr;
        }
        bool stop = false;
#pragma clang loop unroll(disable)
        for (uint32_t k = 0; k < nleaf; k++) {
            const uint32_t first = k ? f1 : f0, count = k ? c1 : c0;
            if (hit_primitives<COOP>(ray, tr, sc, sc.scene_leaf_ids + first, count, any_hit, nullptr, 0, 0)) {
                found = true;
                if (any_hit) { stop = true; break; }
            }
        }
        if (stop) break;
        if (has_left) {
            if (has_right) {
                if (!any_hit && ld > rd) { uint32_t t = lnext; lnext = rnext; rnext = t; }
                if (stack_size + 1u >= sc.scene_stack_need) {
                    trav_fault(sc, ST_FAULT_SCENE_STACK);
                    break;
                }
                stack[ST_STACK_STRIDE * stack_size++] = rnext;
            }
            left = lnext;
        } else if (has_right) {
            left = rnext;
        } else {
            if (stack_size == 0) break;
            left = stack[ST_STACK_STRIDE * --stack_size];
        }
    }
    return found;
}


/* ---- shading ----------------------------------------------------------- */

DFN float tone_map(float c) {
    float x = c - 0.004f;
    if (x < 0.0f) x = 0.0f;
    float x2s = x * x * 6.2f;
    return (x2s + x * 0.5f) / (x2s + x * 1.7f + 0.06f);
}
DFN float inv_dot(sv3 a, sv3 b) { return f_clamp01(-v_dot(a, b)); }
DFN sv3 reflect_with_dot(sv3 V, sv3 N, float NdotV) { return v_smad(N, -2 * NdotV, V); }
DFN sv3 refract_dir(sv3 V, sv3 N, float n1_over_n2, float NdotV) {
    float c = n1_over_n2 * n1_over_n2 * (1 - (NdotV * NdotV));
    if (c > 1) return V3s(0);
    V = v_scale(V, n1_over_n2);
    V = v_smad(N, n1_over_n2 * -NdotV - sqrtf(1 - c), V);
    return v_norm(V);
}

/* common.h:197-228 */
DFN sv3 area_light_vector(const DPrim &p, sv3 P, sv3 *v) {
    if (p.scale.x == 0 || p.scale.z == 0) return V3s(0);
    sv3 U = V3s(0), W = V3s(0);
    U.x = p.scale.x; if (U.x < 0) U.x = -U.x;
    W.z = p.scale.z; if (W.z < 0) W.z = -W.z;
    U = q_rot(U, p.rot);
    W = q_rot(W, p.rot);
    v[0] = v_sub(v_sub(p.pos, U), W);
    v[1] = v_sub(v_add(p.pos, U), W);
    v[2] = v_add(v_add(p.pos, U), W);
    v[3] = v_add(v_sub(p.pos, U), W);
    sv3 u1n = v_norm(v_sub(v[0], P));
    sv3 u2n = v_norm(v_sub(v[1], P));
    sv3 u3n = v_norm(v_sub(v[2], P));
    sv3 u4n = v_norm(v_sub(v[3], P));
    float a12 = acosf(v_dot(u1n, u2n));
    float a23 = acosf(v_dot(u2n, u3n));
    float a34 = acosf(v_dot(u3n, u4n));
    float a41 = acosf(v_dot(u4n, u1n));
    sv3 out = v_scale(v_cross(u1n, u2n), a12);
    out = v_add(out, v_scale(v_cross(u2n, u3n), a23));
    out = v_add(out, v_scale(v_cross(u3n, u4n), a34));
    out = v_add(out, v_scale(v_cross(u4n, u1n), a41));
    return v_scale(out, 0.5f);
}

/* common.h:277-334 */
DFN float ggx_ndf(float r2, float NdotH) {
    float demon = NdotH * NdotH * (r2 - 1.0f) + 1.0f;
    return D_ONE_OVER_PI * r2 / (demon * demon);
}
DFN float ggx_smith_schlick(float NdotL, float NdotV, float roughness) {
    float k = roughness * 0.5f;
    float one_minus_k = 1.0f - k;
    float denom = NdotV * one_minus_k + k;
    float result = NdotV / fmaxf(denom, D_EPS);
    denom = NdotL * one_minus_k + k;
    return result * NdotL / fmaxf(denom, D_EPS);
}
DFN sv3 shade_classic(sv3 diffuse, float NdotL, sv3 specular, float f, float exponent, float roughness) {
    float shininess = 1.0f - roughness;
    f = NdotL * powf(f, exponent * shininess) * shininess;
    return v_smad(specular, f, diffuse);
}
DFN sv3 shade_point(const DShaded &s, const DMat &M, float NdotL) {
    sv3 diffuse = v_scale(s.albedo, M.roughness * NdotL * D_ONE_OVER_PI);
    if (M.brdf == ST_BRDF_LAMBERT) return diffuse;
    if (M.brdf == ST_BRDF_PHONG) {
        float RdotL = v_dotc(s.reflected_direction, s.light_direction);
        return RdotL != 0 ? shade_classic(diffuse, NdotL, M.reflectivity, RdotL, 4.0f, M.roughness) : diffuse;
    }
    sv3 N = s.normal, L = s.light_direction, V = v_neg(s.viewing_direction);
    sv3 H = v_norm(v_add(L, V));
    float NdotH = v_dotc(N, H);
    if (M.brdf == ST_BRDF_BLINN)
        return NdotH != 0 ? shade_classic(diffuse, NdotL, M.reflectivity, NdotH, 16.0f, M.roughness) : diffuse;
    sv3 F = V3s(0);
    diffuse = v_scale(s.albedo, (1.0f - M.metallic) * NdotL * D_ONE_OVER_PI);
    float LdotH = v_dotc(L, H);
    if (LdotH != 1.0f) {
        F = v_smad(v_one_minus(M.reflectivity), powf(1.0f - LdotH, 5.0f), M.reflectivity);
        diffuse = v_mul(diffuse, v_one_minus(F));
    }
    if (M.roughness == 0 && NdotH == 1.0f) return diffuse;
    float NdotV = v_dotc(N, V);
    if (NdotV == 0) return diffuse;
    float D = ggx_ndf(M.roughness * M.roughness, NdotH);
    float G = ggx_smith_schlick(NdotL, NdotV, M.roughness);
    return v_smad(F, D * G / (4.0f * NdotV), diffuse);
}

DFN float duv_ray_cone(float NdotV, float cone_width, float area, float uv_area) {
    float pcw = cone_width / fabsf(NdotV);
    return sqrtf((pcw * pcw) * (uv_area / area));
}

/* scene/texture.h:5-40: one 12-byte quad load per bilinear sample. Indices
   are clamped to the quad grid so that out-of-range uvs (undefined in the
   reference) cannot leave the texture. */
DFN sv3 sample_texture(const StDScene &sc, uint32_t tex_id, sv2 UV, sv2 dUV) {
    const StDTex tex = load_tex(sc, tex_id);
    uint32_t mip_level = 0;
    if (tex.mipmap) {
        const float pw = dUV.x * (float)tex.width;
        const float ph = dUV.y * (float)tex.height;
        float pixel_size = (float)(pw + ph) * 0.5f;
        const uint32_t last_mip = (uint8_t)(tex.mip_count - 1);
        while (pixel_size > 1 && mip_level < last_mip) { pixel_size /= 2; mip_level += 1; }
    }
    const StDMip mip = load_mip(sc, tex.mip_base + mip_level);
    float u = UV.x, v = UV.y;
    if (u > 1) u -= (float)((uint32_t)u);
    if (v > 1) v -= (float)((uint32_t)v);
    const float U = u * (float)mip.width + 0.5f;
    const float Vv = v * (float)mip.height + 0.5f;
    uint32_t x = (uint32_t)U, y = (uint32_t)Vv;
    const float r = U - (float)x, b = Vv - (float)y;
    const float l = 1 - r, t = 1 - b;
    const float tl = t * l * D_COLOR_TO_FLOAT, trr = t * r * D_COLOR_TO_FLOAT;
    const float bl = b * l * D_COLOR_TO_FLOAT, br = b * r * D_COLOR_TO_FLOAT;
    x = x < mip.width ? x : mip.width;
    y = y < mip.height ? y : mip.height;
    const uint32_t *q = reinterpret_cast<const uint32_t *>(sc.texels + 12 * (mip.quad_offset + (uint64_t)y * (mip.width + 1) + x));
    uint32_t R = ldgu(q, 0), G = ldgu(q, 1), B = ldgu(q, 2);
    return V3((float)(R >> 24) * br + ((float)((R >> 16) & 255) * bl + ((float)((R >> 8) & 255) * trr + (float)(R & 255) * tl)),
              (float)(G >> 24) * br + ((float)((G >> 16) & 255) * bl + ((float)((G >> 8) & 255) * trr + (float)(G & 255) * tl)),
              (float)(B >> 24) * br + ((float)((B >> 16) & 255) * bl + ((float)((B >> 8) & 255) * trr + (float)(B & 255) * tl)));
}
DFN sv3 sample_normal(const StDScene &sc, uint32_t tex_id, sv2 uv, sv2 dUV) {
    sv3 n = sample_texture(sc, tex_id, uv, dUV);
    float y = n.z;
    n.z = n.y;
    n.y = y;
    return v_norm(v_smad(n, 2.0f, V3s(-1.0f)));
}
DFN sq4 normal_rotation(sv3 n) {
    sq4 q;
    q.v.x = n.z * 0.25f;
    q.v.y = 0;
    q.v.z = n.x * 0.25f;
    q.w = n.y * 0.25f;
    return q_norm(q);
}

/* surface.h:37-129 */
DNOINLINE bool shade_emissive_quads(DShaded &s, const DMat &M, DRay &ray, DTrace &tr, const StDScene &sc, sv3 &color) {
    DHit &hit = tr.current_hit;
    sv3 Ro, Rd;
    bool found = false;
    for (uint32_t i = 0; i < sc.prim_count; i++) {
        DPrim quad = load_prim(sc, i);
        if (quad.type != ST_PRIM_QUAD) continue;
        uint32_t em_is = mat_is(sc, quad.material_id);
        if (!(em_is & ST_EMISSIVE)) continue;
        to_object(s.viewing_origin, s.viewing_direction, quad, Ro, Rd);
        tr.local.origin = Ro;
        tr.local.direction = Rd;
        if (hit_quad(hit, Ro, Rd, quad.flags)) {
            hit.position = pos_to_world(hit.position, quad);
            hit.distance_squared = v_sqlen(v_sub(hit.position, s.viewing_origin));
            if (hit.distance_squared < tr.closest_hit.distance_squared) {
                hit.object_id = i;
                hit.object_type = ST_PRIM_QUAD;
                hit.material_id = quad.material_id;
                tr.closest_hit = hit;
                found = true;
            }
        }
        sv3 qn = q_rot(V3(0, 1, 0), quad.rot);
        s.light_direction = ray.direction = v_sub(quad.pos, s.position);
        if (v_dot(qn, s.light_direction) >= 0) continue;
        float emission_intensity = v_dot(s.normal, area_light_vector(quad, s.position, s.quad_vertices));
        if (emission_intensity > 0) {
            bool skip = true;
            for (int j = 0; j < 4; j++)
                if (v_dot(s.normal, v_sub(s.quad_vertices[j], s.position)) >= 0) { skip = false; break; }
            if (skip) continue;
            float shaded_light = 1;
            for (uint32_t k = 0; k < sc.prim_count; k++) {
                if (i == s.prim_index || i == k) continue;
                DPrim sp = load_prim(sc, k);
                if (v_dot(qn, v_sub(sp.pos, quad.pos)) <= 0) continue;
                to_object(s.position, s.light_direction, sp, Ro, Rd);
                Ro = v_smad(Rd, D_TRACE_OFFSET, Ro);
                tr.local.origin = Ro;
                tr.local.direction = Rd;
                float d = 1;
                if (sp.type == ST_PRIM_SPHERE) {
                    if (hit_sphere(hit, Ro, Rd, sp.flags))
                        d -= (1.0f - sqrtf(hit.distance_squared)) / (hit.distance * emission_intensity * 3);
                } else if (sp.type == ST_PRIM_QUAD) {
                    if (hit_quad(hit, Ro, Rd, sp.flags)) {
                        hit.position.y = 0;
                        hit.position.x = hit.position.x < 0 ? -hit.position.x : hit.position.x;
                        hit.position.z = hit.position.z < 0 ? -hit.position.z : hit.position.z;
                        if (hit.position.x > hit.position.z) {
                            hit.position.y = hit.position.z;
                            hit.position.z = hit.position.x;
                            hit.position.x = hit.position.y;
                            hit.position.y = 0;
                        }
                        d -= (1.0f - hit.position.z) / (hit.distance * emission_intensity);
                    }
                }
                if (d < shaded_light) shaded_light = d;
            }
            if (shaded_light > 0) {
                float NdotL = v_dotc(s.normal, s.light_direction);
                if (NdotL != 0) {
                    const void *mp = sc.mats + quad.material_id;
                    float4 mb = ldg4(mp, 1), mc = ldg4(mp, 2);
                    sv3 emission = V3(mb.z, mb.w, mc.x);
                    color = v_mad(shade_point(s, M, NdotL), v_scale(emission, emission_intensity * shaded_light), color);
                }
            }
        }
    }
    if (found) tr.closest_hit.distance = sqrtf(tr.closest_hit.distance_squared);
    return found;
}

/* closest_hit/lights.h:7-39 */
DFN float volume_density(const DSphereHit &h) {
    float t1 = h.t_near > 0 ? h.t_near : 0;
    float t2 = h.t_far < h.furthest ? h.t_far : h.furthest;
    return (h.c * t1 - h.b * t1 * t1 + t1 * t1 * t1 / 3.0f - (h.c * t2 - h.b * t2 * t2 + t2 * t2 * t2 / 3.0f)) * (3.0f / 4.0f);
}
DNOINLINE bool shade_light_glow(const StDScene &sc, sv3 Ro, sv3 Rd, float max_distance, DSphereHit &sh, sv3 &color) {
    bool hit_light = false;
    sh.closest_hit_distance = INFINITY;
    for (uint32_t i = 0; i < sc.light_count; i++) {
        const void *lp = sc.lights + i;
        float4 la = ldg4(lp, 0), lb = ldg4(lp, 1);
        float size = la.w * (1.0f / 16.0f);
        float one_over_radius = 8.0f / size;
        sh.furthest = max_distance * one_over_radius;
        if (hit_sphere_simple(Ro, Rd, V3(la.x, la.y, la.z), one_over_radius, sh)) {
            float distance = sh.t_far - size * 8;
            float density = powf(volume_density(sh), 8) * 8;
            if (distance < sh.closest_hit_distance) {
                sh.closest_hit_distance = distance;
                sh.closest_hit_density = density;
            }
            color = v_smad(V3(lb.x, lb.y, lb.z), density, color);
            hit_light = true;
        }
    }
    return hit_light;
}

/* surface.h:159-168 (both dUV components) and :227-236 (dUV.u only) */
DFN void apply_textures(DTrace &tr, const StDScene &sc, DShaded &s, const DMat &M, sv2 &dUV, bool set_both) {
    DHit &hit = tr.closest_hit;
    s.uv = V2(hit.uv.x * M.uv_repeat.x, hit.uv.y * M.uv_repeat.y);
    hit.uv_area /= M.uv_repeat.x / M.uv_repeat.y;
    float d = duv_ray_cone(hit.NdotV, hit.cone_width, hit.area, hit.uv_area);
    dUV.x = d;
    if (set_both) dUV.y = d;
    if (M.use & ST_ALBEDO_MAP) s.albedo = sample_texture(sc, M.tex0, s.uv, dUV);
    if ((M.use & ST_NORMAL_MAP) && M.texture_count > 1) {
        sq4 rot = normal_rotation(sample_normal(sc, M.tex1, s.uv, dUV));
        hit.normal = q_rot(hit.normal, rot);
    }
}

#endif
