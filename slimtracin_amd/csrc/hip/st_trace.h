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
/* A traversal stack: its newest entries in the lane's LDS column, used as a
   ring of mask + 1 slots, the older ones in the launch's spill buffer
   (StDScene.spill: per wave of the kernel, 64 lanes of spill_s scene then
   spill_m mesh entries; `first` is 0 for the scene stack, spill_s for the
   mesh stack). Entries [lo, sp) are in LDS. With mask all ones the column
   holds the whole stack and nothing spills. The spill address is computed
   only when an entry moves, so the traversal keeps its registers. */
struct DStk {
    uint32_t *lds;
    uint32_t mask, lo;
};
DFN ST_GLOBAL uint32_t *stk_spill(const StDScene &sc, uint32_t first, uint32_t i) {
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const size_t per_wave = 64u * (size_t)(sc.spill_s + sc.spill_m);
    return (ST_GLOBAL uint32_t *)sc.spill + wave * per_wave + 64u * (size_t)(first + i) + (threadIdx.x & 63u);
}
/* push v as entry sp */
DFN void stk_push(DStk &s, const StDScene &sc, uint32_t first, uint32_t sp, uint32_t v) {
    if (sp - s.lo > s.mask) {
        *stk_spill(sc, first, s.lo) = s.lds[ST_STACK_STRIDE * (s.lo & s.mask)];
        s.lo++;
    }
    s.lds[ST_STACK_STRIDE * (sp & s.mask)] = v;
}
/* entry sp (the stack's size after the pop) */
DFN uint32_t stk_pop(DStk &s, const StDScene &sc, uint32_t first, uint32_t sp) {
    if (sp < s.lo) {
        s.lo = sp;
        return *stk_spill(sc, first, sp);
    }
    return s.lds[ST_STACK_STRIDE * (sp & s.mask)];
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

/* ---- cooperative traversal of one ray by the blocks of a redo group ------

   The redo stage traces each of its pixels with a group of blocks: the
   group's leader block runs the pixel's path (every thread of it computing
   the same values), and every traversal of a large mesh on that path is a
   job that the leader posts to the group's record in device memory; the
   group's other blocks (helpers) wait for jobs and take part in them. No
   block ever waits for a block that has not started: a helper only joins a
   posted job, and the leader waits only for helpers that have joined, so a
   group whose helpers never get a CU still finishes, on the leader alone.

   Work of a job: tasks, each a subtree of the mesh tree that one lane
   traverses with the reference's loop (mesh.h:38-111) with the box bound
   fixed at the traversal's start. The leader's first lane starts with the
   root; a lane with no task takes the oldest entry of another lane's stack
   (the largest subtree that lane has left, visited last), within its wave
   or through the group's queue (co_work). A task's order key is its path
   from the root, two bits per level, most significant first: leaf children
   of a step 0 (left) and 1 (right), the inner child entered first 2 and
   last 3 (the nearer first for a closest hit). The reference visits the
   subtrees in key order, and each task's visited set is a contiguous run
   of that order (an entry taken from a stack is visited after everything
   its lane still visits), so:
     any hit: the reference's hit is the first hit of the task with the
       least key (tasks beyond a key already hit stop early);
     closest hit: t* = the least hit distance; among hits at t*, the one of
       the least key is first in the reference's order. The reference
       reaches it if every box on its path admits it at bound t* (every
       entry distance on the path is at most t*), and then keeps it;
     nothing hit: the bound never moved, so the tasks are exactly what the
       reference visits, and its current-hit distance (the only field of
       that record a later result reads) is that of the last plane hit of
       the task latest in key order that has one.
   All shared words are read and written with device-scope atomics (they
   bypass the non-coherent caches: the group's blocks sit on different
   XCDs); queued tasks carry a tag in each 64-bit word, so a reader never
   needs a fence to know a task is whole. */

#define CO_CLOSED 0x80000000u     /* job word: the leader has ended the job */
#define CO_MAX_JOBS 4095u         /* jobs per group and launch (the job's sequence number's low 12 bits) */
#define CO_MAX_DIGITS 64u         /* order-key digits (tree levels) */
#define CO_NODE_BITS 25u          /* stack entry: node pair index | depth << CO_NODE_BITS */
#define CO_EXPORT_LOW 128u        /* waves hand entries to the queue while it holds fewer than this */
#define CO_SPIN_LIMIT (1u << 22)  /* polls of any wait before it gives up (seconds) */
#define CO_HELPER_IDLE_TICKS 20000u  /* 200 us of the 100 MHz clock without a job: a helper leaves */
/* status bits of a job */
#define CO_ST_DEEP 1u             /* a path deeper than the order key or a stack limit: trace sequentially */
#define CO_ST_STEPS 2u            /* a traversal ran past its step bound (malformed tree) */
#define CO_ST_LOST 4u             /* a wait gave up: the job's result is not known */

typedef unsigned long long u64;
#define CO_G __attribute__((address_space(1)))

DFN uint32_t co_ld(const uint32_t *p) {
    return __hip_atomic_load((const CO_G uint32_t *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DFN u64 co_ld64(const u64 *p) {
    return __hip_atomic_load((const CO_G u64 *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DFN void co_st(uint32_t *p, uint32_t v) {
    __hip_atomic_store((CO_G uint32_t *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DFN void co_st64(u64 *p, u64 v) { __hip_atomic_store((CO_G u64 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
DFN uint32_t co_add(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_add((CO_G uint32_t *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DFN u64 co_add64(u64 *p, u64 v) {
    return __hip_atomic_fetch_add((CO_G u64 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DFN void co_max64(u64 *p, u64 v) {
    (void)__hip_atomic_fetch_max((CO_G u64 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DFN void co_min64(u64 *p, u64 v) {
    (void)__hip_atomic_fetch_min((CO_G u64 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DFN void co_or(uint32_t *p, uint32_t v) {
    (void)__hip_atomic_fetch_or((CO_G uint32_t *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DFN bool co_cas(uint32_t *p, uint32_t expect, uint32_t want) {
    return __hip_atomic_compare_exchange_strong((CO_G uint32_t *)p, &expect, want, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DFN void co_sleep() { __builtin_amdgcn_s_sleep(2); }

/* Path order key: 2-bit digits, most significant first (digit i at bit
   126 - 2i of hi:lo). */
struct DKey { u64 hi, lo; };
DFN DKey key_digit(DKey k, uint32_t i, uint32_t dg) {
    if (i < 32u) k.hi |= (u64)dg << (62u - 2u * i);
    else k.lo |= (u64)dg << (62u - 2u * (i - 32u));
    return k;
}
DFN DKey key_trunc(DKey k, uint32_t n) {
    if (n == 0u) return DKey{0, 0};
    if (n < 32u) {
        k.hi &= ~0ull << (64u - 2u * n);
        k.lo = 0;
    } else if (n == 32u) {
        k.lo = 0;
    } else if (n < 64u) {
        k.lo &= ~0ull << (64u - 2u * (n - 32u));
    }
    return k;
}
DFN bool key_lt(DKey a, DKey b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }

/* One job: a mesh traversal posted by the leader (26 words). */
struct DCoopJob {
    uint32_t ro[3], rd[3], rr[3], rs[3];   /* local ray: origin, direction, reciprocal, scaled origin (bits) */
    uint32_t neg, c0, nodes_lo, nodes_hi, tris_lo, tris_hi;
    uint32_t T, pad0, pad1, root_first, any_hit, max_steps, stack_size, seq;
};
/* A group's record (CO_HEADER bytes), then its queue of tasks
   (4 words of 64 bits each). */
struct DCoopGroup {
    u64 claim;      /* job word: seq << 32, | CO_CLOSED once the leader has ended the job */
    u64 first_hi;   /* any hit: least key (high word) of a task with a hit so far */
    uint32_t inflight;           /* helper blocks taking part in the job */
    uint32_t status;             /* CO_ST_* */
    uint32_t head, tail;         /* queue: next item to take, next free */
    uint32_t busy;               /* waves holding tasks */
    uint32_t quit;               /* the leader's launch base once it has no more pixels */
    uint32_t pad[6];
    DCoopJob job;
};
#define CO_HEADER 256u
static_assert(sizeof(DCoopGroup) <= CO_HEADER, "group record");
/* a block's result for a job (16 words): seq tag, hit, last plane hit */
struct DCoopSlot {
    uint32_t seq, found, t, tri, pm, k0, k1, k2, k3, lfound, l0, l1, l2, l3, ldist, pad;
};

/* a job as one thread holds it */
struct DCoopJobR {
    const uint32_t *nodes;
    const StDTriHot *tris;
    float c0;
    uint32_t T, root_first, any_hit, max_steps, stack_size, seq, valid;
};

/* one thread's (or block's) result so far */
struct DCoopRes {
    uint32_t found, tri, lfound;
    float t, pm, ldist;
    DKey k, lk;
};
DFN void res_merge(DCoopRes &r, const DCoopRes &o, bool any_hit) {
    if (o.found) {
        bool better = !r.found;
        if (!better) {
            if (!any_hit && o.t != r.t) better = o.t < r.t;
            else better = key_lt(o.k, r.k);
        }
        if (better) {
            r.found = 1;
            r.t = o.t;
            r.tri = o.tri;
            r.pm = o.pm;
            r.k = o.k;
        }
    }
    if (o.lfound && (!r.lfound || key_lt(r.lk, o.lk))) {
        r.lfound = 1;
        r.lk = o.lk;
        r.ldist = o.ldist;
    }
}
DFN uint32_t shfl_u(uint32_t v, uint32_t m) { return (uint32_t)__shfl_xor((int)v, (int)m, 64); }
DFN u64 shfl_u64(u64 v, uint32_t m) {
    return (u64)shfl_u((uint32_t)v, m) | ((u64)shfl_u((uint32_t)(v >> 32), m) << 32);
}
DFN void res_wave_reduce(DCoopRes &r, bool any_hit) {
#pragma unroll
    for (uint32_t m = 32; m > 0; m >>= 1) {
        DCoopRes o;
        o.found = shfl_u(r.found, m);
        o.t = __uint_as_float(shfl_u(__float_as_uint(r.t), m));
        o.tri = shfl_u(r.tri, m);
        o.pm = __uint_as_float(shfl_u(__float_as_uint(r.pm), m));
        o.k.hi = shfl_u64(r.k.hi, m);
        o.k.lo = shfl_u64(r.k.lo, m);
        o.lfound = shfl_u(r.lfound, m);
        o.lk.hi = shfl_u64(r.lk.hi, m);
        o.lk.lo = shfl_u64(r.lk.lo, m);
        o.ldist = __uint_as_float(shfl_u(__float_as_uint(r.ldist), m));
        res_merge(r, o, any_hit);
    }
}

#define CO_MAX_WAVES 4u   /* waves per redo block (REDO_THREADS / 64) */
/* Per block state of the redo stage (LDS). */
struct DCoopBlk {
    DCoopGroup *group;
    u64 *queue;
    DCoopSlot *slots;      /* the group's blocks' slots, by rank */
    uint32_t gsize, rank, q_cap, base, jobs, seq, cmd, status;
    uint32_t jw[26];       /* a helper: the job it takes part in */
    DCoopRes red[CO_MAX_WAVES];
    uint32_t ex[CO_MAX_WAVES][64 * 6];   /* per wave: tasks handed between lanes */
};
static __shared__ DCoopBlk co_blk;

#define T_STAT(m) ((u64)(m).triangle_count)
#ifdef CO_STATS
/* measurement builds: one record per cooperative job, redone pixel and
   sequential retry (4 words: kind | group << 8 | aux << 32, start clock,
   end clock, detail) */
static __device__ unsigned long long co_stats[32768][4];
static __device__ unsigned int co_stats_n;
DFN void co_stat(u64 a, u64 b, u64 c, u64 d) {
    const unsigned i = atomicAdd(&co_stats_n, 1u);
    if (i < 32768u) {
        co_stats[i][0] = a;
        co_stats[i][1] = b;
        co_stats[i][2] = c;
        co_stats[i][3] = d;
    }
}
#endif

/* Merge r over the block: every thread returns the block's result (all
   threads call it). */
DFN DCoopRes res_block_reduce(DCoopRes r, bool any_hit) {
    res_wave_reduce(r, any_hit);
    const uint32_t wv = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63u) == 0) co_blk.red[wv] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        DCoopRes a = co_blk.red[0];
        for (uint32_t i = 1; i < nw; i++) res_merge(a, co_blk.red[i], any_hit);
        co_blk.red[0] = a;
    }
    __syncthreads();
    const DCoopRes out = co_blk.red[0];
    __syncthreads();
    return out;
}

DFN void co_put_slot(DCoopSlot *s, uint32_t seq, const DCoopRes &r) {
    uint32_t *p = (uint32_t *)s;
    co_st(p + 1, r.found);
    co_st(p + 2, __float_as_uint(r.t));
    co_st(p + 3, r.tri);
    co_st(p + 4, __float_as_uint(r.pm));
    co_st(p + 5, (uint32_t)(r.k.hi >> 32));
    co_st(p + 6, (uint32_t)r.k.hi);
    co_st(p + 7, (uint32_t)(r.k.lo >> 32));
    co_st(p + 8, (uint32_t)r.k.lo);
    co_st(p + 9, r.lfound);
    co_st(p + 10, (uint32_t)(r.lk.hi >> 32));
    co_st(p + 11, (uint32_t)r.lk.hi);
    co_st(p + 12, (uint32_t)(r.lk.lo >> 32));
    co_st(p + 13, (uint32_t)r.lk.lo);
    co_st(p + 14, __float_as_uint(r.ldist));
    co_st(p + 0, seq);
}
DFN bool co_get_slot(const DCoopSlot *s, uint32_t seq, DCoopRes &r) {
    const uint32_t *p = (const uint32_t *)s;
    if (co_ld(p) != seq) return false;
    r.found = co_ld(p + 1);
    r.t = __uint_as_float(co_ld(p + 2));
    r.tri = co_ld(p + 3);
    r.pm = __uint_as_float(co_ld(p + 4));
    r.k.hi = ((u64)co_ld(p + 5) << 32) | co_ld(p + 6);
    r.k.lo = ((u64)co_ld(p + 7) << 32) | co_ld(p + 8);
    r.lfound = co_ld(p + 9);
    r.lk.hi = ((u64)co_ld(p + 10) << 32) | co_ld(p + 11);
    r.lk.lo = ((u64)co_ld(p + 12) << 32) | co_ld(p + 13);
    r.ldist = __uint_as_float(co_ld(p + 14));
    return true;
}

/* Queued tasks: node | depth << CO_NODE_BITS, path maximum, key; four
   64-bit words in ring slot (index % q_cap), each with the task's 16-bit tag
   (from the job and the task's index) in its top bits, so that a reader
   tells a whole task from a slot still being written or holding an older
   one. The reader of a task zeroes its first word once it has read it; a
   writer waits for that before it reuses the slot. */
DFN u64 co_tag(uint32_t seq, uint32_t idx) { return (u64)((seq * 4099u + idx) % 0xFFFFu + 1u) << 48; }
/* false: the slot never came free */
DFN bool co_push_item(u64 *slot, uint32_t seq, uint32_t idx, uint32_t nd, float pm, DKey k) {
    for (uint32_t spins = 0; co_ld64(slot) != 0ull; spins++) {
        if (spins > CO_SPIN_LIMIT) return false;
        co_sleep();
    }
    const u64 tg = co_tag(seq, idx);
    const uint32_t pb = __float_as_uint(pm);
    co_st64(slot + 1, tg | ((u64)(pb & 0xFFFFu) << 32) | (k.hi >> 32));
    co_st64(slot + 2, tg | ((k.hi & 0xFFFFFFFFull) << 16) | (k.lo >> 48));
    co_st64(slot + 3, tg | (k.lo & 0xFFFFFFFFFFFFull));
    co_st64(slot + 0, tg | ((u64)nd << 16) | (pb >> 16));
    return true;
}
/* Item idx of the queue (a slot this wave has claimed): waits for its words,
   frees the slot; 1, or -1 when its words never arrived or do not decode
   to a node pair of the mesh. */
DFN int co_pop_item(u64 *slot, uint32_t seq, uint32_t idx, uint32_t T, uint32_t &nd, float &pm, DKey &k) {
    const u64 tg = co_tag(seq, idx), m = 0xFFFFull << 48;
    u64 w0, w1, w2, w3;
    for (uint32_t spins = 0;; spins++) {
        w0 = co_ld64(slot + 0);
        w1 = co_ld64(slot + 1);
        w2 = co_ld64(slot + 2);
        w3 = co_ld64(slot + 3);
        if ((w0 & m) == tg && (w1 & m) == tg && (w2 & m) == tg && (w3 & m) == tg) break;
        if (spins > CO_SPIN_LIMIT) return -1;
        co_sleep();
    }
    co_st64(slot + 0, 0);
    nd = (uint32_t)(w0 >> 16);
    pm = __uint_as_float((uint32_t)(((w0 & 0xFFFFull) << 16) | ((w1 >> 32) & 0xFFFFull)));
    k.hi = ((w1 & 0xFFFFFFFFull) << 32) | ((w2 >> 16) & 0xFFFFFFFFull);
    k.lo = ((w2 & 0xFFFFull) << 48) | (w3 & 0xFFFFFFFFFFFFull);
    const uint32_t node = nd & ((1u << CO_NODE_BITS) - 1u), depth = nd >> CO_NODE_BITS;
    if (node + 1u >= 2u * T || depth == 0u || depth > CO_MAX_DIGITS) return -1;
    return 1;
}

/* The job in LDS as 26 words (a helper's thread 0 copies it from the
   record) -> one thread's copy. */
DFN void co_job_from_words(const uint32_t *v, uint32_t seq, DCoopJobR &J, DTrace &tr) {
    const DCoopJob &j = *(const DCoopJob *)v;
    tr.local.origin = V3(__uint_as_float(j.ro[0]), __uint_as_float(j.ro[1]), __uint_as_float(j.ro[2]));
    tr.local.direction = V3(__uint_as_float(j.rd[0]), __uint_as_float(j.rd[1]), __uint_as_float(j.rd[2]));
    tr.local.direction_reciprocal = V3(__uint_as_float(j.rr[0]), __uint_as_float(j.rr[1]), __uint_as_float(j.rr[2]));
    tr.local.scaled_origin = V3(__uint_as_float(j.rs[0]), __uint_as_float(j.rs[1]), __uint_as_float(j.rs[2]));
    tr.local.neg_x = j.neg & 1u;
    tr.local.neg_y = (j.neg >> 1) & 1u;
    tr.local.neg_z = (j.neg >> 2) & 1u;
    J.nodes = (const uint32_t *)(uintptr_t)(((u64)j.nodes_hi << 32) | j.nodes_lo);
    J.tris = (const StDTriHot *)(uintptr_t)(((u64)j.tris_hi << 32) | j.tris_lo);
    J.c0 = __uint_as_float(j.c0);
    J.T = j.T;
    J.root_first = j.root_first;
    J.any_hit = j.any_hit;
    J.max_steps = j.max_steps;
    J.stack_size = j.stack_size;
    J.seq = seq;
    J.valid = 1;
    tr.mesh_stack_size = j.stack_size;
}

/* Work of a job, per wave. Each lane holds at most one task: a subtree
   (node pair, path key, path maximum, depth) that it traverses with the
   reference's loop, one step per iteration. Every CO_BAL iterations the
   wave rebalances: lanes without a task take the oldest stack entry (the
   largest subtree left, visited last) of lanes that have one, through LDS;
   a wave that still has many idle lanes takes a batch of tasks from the
   group's queue, and a wave with no idle lane hands a batch of oldest
   entries to the queue while it is short. The queue's shared words change
   once per batch. The leader's first wave starts with the root. */
#define CO_BAL 4u                 /* iterations between rebalances */
#define CO_POP_MIN 16u            /* idle lanes that make a wave take tasks from the queue */

/* a task: node | depth << CO_NODE_BITS, path maximum, key */
struct DTask {
    uint32_t nd;
    float pm;
    DKey k;
};
/* a wave's task exchange area in LDS (6 words per lane) */
DFN volatile uint32_t *co_ex_slot(uint32_t i) { return co_blk.ex[threadIdx.x >> 6] + 6u * i; }

/* One wave's share of a job (every lane of the wave calls it; the wave
   leaves together). Results are merged into `res`; returns true when the
   wave took part in the job. */
DFN bool co_work(DTrace &tr, DCoopGroup *G, u64 *queue, uint32_t q_cap, const DCoopJobR &J, DCoopRes &res,
                 bool root) {
    DHit &closest = tr.closest_mesh_hit;
    uint32_t *stack = tr.mesh_stack, *pstack = tr.aux_stack;
    const uint32_t lane = threadIdx.x & 63u;
    bool has = false, it_found = false, err = false;
    bool wave_busy = root;            /* counted in the record's busy waves */
    bool took = root;
    uint32_t left = 0, sp = 0, steps = 0, dep = 0, idle_rounds = 0;
    float pm = 0.0f, it_pm = 0.0f;
    DKey key = {0, 0}, ks = {0, 0};
    DNodePair np = {};
    auto start = [&](uint32_t nd, float p, DKey k) {
        has = true;
        left = nd & ((1u << CO_NODE_BITS) - 1u);
        dep = nd >> CO_NODE_BITS;
        pm = p;
        key = ks = k;
        sp = steps = 0;
        it_found = false;
        closest.distance = J.c0;
        tr.current_hit.distance = __uint_as_float(ST_DIST_UNSET);
        np = load_pair(J.nodes, left);
    };
    auto finish = [&]() {
#ifdef CO_STATS
        {
            uint32_t *sw = G->pad;
            co_add(sw + 0, steps);
            co_add(sw + 1, 1u);
            (void)__hip_atomic_fetch_max((CO_G uint32_t *)(sw + 2), steps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
#endif
        DCoopRes o = {};
        if (it_found) {
            o.found = 1;
            o.t = closest.distance;
            o.tri = closest.object_id;
            o.pm = it_pm;
            o.k = ks;
            if (J.any_hit) co_min64(&G->first_hi, ks.hi);
        } else {
            const float cd = tr.current_hit.distance;
            if (__float_as_uint(cd) != ST_DIST_UNSET) {
                o.lfound = 1;
                o.lk = ks;
                o.ldist = cd;
            }
        }
        res_merge(res, o, J.any_hit != 0);
        has = false;
    };
    /* the oldest stack entry as a task of its own (removed from the stack) */
    auto take_oldest = [&](DTask &t) {
        const uint32_t e = stack[0];
        const uint32_t de = e >> CO_NODE_BITS;
        t.nd = e;
        t.pm = __uint_as_float(pstack[0]);
        t.k = key_digit(key_trunc(key, de - 1u), de - 1u, 3u);
        for (uint32_t i = 1; i < sp; i++) {
            stack[ST_STACK_STRIDE * (i - 1u)] = stack[ST_STACK_STRIDE * i];
            pstack[ST_STACK_STRIDE * (i - 1u)] = pstack[ST_STACK_STRIDE * i];
        }
        sp--;
    };
    if (root && lane == 0) start(J.root_first, 0.0f, DKey{0, 0});
    for (uint32_t it = 1;; it++) {
        if (has) {
            /* one step of the reference's loop (mesh.h:38-111): both
               children's boxes, the leaf children's triangles, then descent */
            float ld, rd;
            const bool hl = hit_aabb(np.lb, tr.local, J.c0, ld);
            const bool hr = hit_aabb(np.rb, tr.local, J.c0, rd);
            const float pml = ld > pm ? ld : pm, pmr = rd > pm ? rd : pm;
            uint32_t f0 = 0, n0 = 0, f1 = 0, n1 = 0;
            bool next = false, end = false;
            if (++steps > J.max_steps) {
                co_or(&G->status, CO_ST_STEPS);
                err = true;
            }
            if (hl && np.lcount) { f0 = np.lfirst; n0 = np.lcount; }
            if (hr && np.rcount) { f1 = np.rfirst; n1 = np.rcount; }
            const bool il = hl && !np.lcount, ir = hr && !np.rcount;
            if (il || ir) {
                if (dep + 1u >= CO_MAX_DIGITS) {
                    co_or(&G->status, CO_ST_DEEP);
                    err = true;
                }
                if (il && ir) {
                    uint32_t a = np.lfirst, b = np.rfirst;
                    float pa = pml, pb = pmr;
                    if (!J.any_hit && ld > rd) {
                        a = np.rfirst; b = np.lfirst;
                        pa = pmr; pb = pml;
                    }
                    if (sp + 1u >= J.stack_size || b >= (1u << CO_NODE_BITS)) {
                        co_or(&G->status, CO_ST_DEEP);
                        err = true;
                    } else {
                        stack[ST_STACK_STRIDE * sp] = b | ((dep + 1u) << CO_NODE_BITS);
                        pstack[ST_STACK_STRIDE * sp] = __float_as_uint(pb);
                        sp++;
                    }
                    left = a;
                    pm = pa;
                } else {
                    left = il ? np.lfirst : np.rfirst;
                    pm = il ? pml : pmr;
                }
                key = key_digit(key, dep, 2u);
                dep++;
                next = true;
            } else if (sp) {
                sp--;
                const uint32_t e = stack[ST_STACK_STRIDE * sp];
                const uint32_t de = e >> CO_NODE_BITS;
                left = e & ((1u << CO_NODE_BITS) - 1u);
                pm = __uint_as_float(pstack[ST_STACK_STRIDE * sp]);
                key = key_digit(key_trunc(key, de - 1u), de - 1u, 3u);
                dep = de;
                next = true;
            } else {
                end = true;
            }
            /* the successor's pair is in flight while this step's triangles are tested */
            if (next && !err) np = load_pair(J.nodes, left);
            if (!err && n0 + n1 && hit_triangles_batched(tr, J.tris, f0, n0, f1, n1, J.any_hit)) {
                it_found = true;
                it_pm = closest.object_id - f0 < n0 ? pml : pmr;
                if (J.any_hit) end = true;
            }
            if (err) has = false;
            else if (end) finish();
        }
        if (__ballot(err)) break;
        if (it % CO_BAL) continue;
        /* any hit: a task after a task that has a hit is moot */
        if (J.any_hit && (it & 31u) == 0 && has && co_ld64(&G->first_hi) < ks.hi) has = false;
        /* within the wave: idle lanes take busy lanes' oldest entries */
        const u64 idle = __ballot(!has);
        if (idle) {
            const u64 donors = __ballot(has && sp > 0);
            const uint32_t ni = (uint32_t)__builtin_popcountll(idle), nd = (uint32_t)__builtin_popcountll(donors);
            const uint32_t n = ni < nd ? ni : nd;
            if (n) {
                const uint32_t rd = __builtin_amdgcn_mbcnt_hi((uint32_t)(donors >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)donors, 0u));
                const uint32_t ri = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                if (has && sp > 0 && rd < n) {
                    DTask t;
                    take_oldest(t);
                    volatile uint32_t *x = co_ex_slot(rd);
                    x[0] = t.nd;
                    x[1] = __float_as_uint(t.pm);
                    x[2] = (uint32_t)(t.k.hi >> 32);
                    x[3] = (uint32_t)t.k.hi;
                    x[4] = (uint32_t)(t.k.lo >> 32);
                    x[5] = (uint32_t)t.k.lo;
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                __builtin_amdgcn_wave_barrier();
                if (!has && ri < n) {
                    volatile uint32_t *x = co_ex_slot(ri);
                    const uint32_t a0 = x[0], a1 = x[1], a2 = x[2], a3 = x[3], a4 = x[4], a5 = x[5];
                    start(a0, __uint_as_float(a1), DKey{((u64)a2 << 32) | a3, ((u64)a4 << 32) | a5});
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                __builtin_amdgcn_wave_barrier();
            }
        }
        const u64 idle2 = __ballot(!has);
        const uint32_t ni2 = (uint32_t)__builtin_popcountll(idle2);
        if (ni2 >= CO_POP_MIN) {
            /* take a batch from the group's queue */
            uint32_t h = 0, n = 0;
            if (lane == 0) {
                const uint32_t hh = co_ld(&G->head), tt = co_ld(&G->tail);
                if (tt > hh) {
                    n = tt - hh < ni2 ? tt - hh : ni2;
                    if (co_cas(&G->head, hh, hh + n)) {
                        h = hh;
                        /* counted once it holds tasks (a wave that sees no
                           busy wave and an empty queue in between leaves
                           early: this wave does their share) */
                        if (!wave_busy) co_add(&G->busy, 1u);
                    } else {
                        n = 0;
                    }
                }
            }
            h = (uint32_t)__shfl((int)h, 0, 64);
            n = (uint32_t)__shfl((int)n, 0, 64);
            if (n) {
                wave_busy = took = true;
                idle_rounds = 0;
                const uint32_t ri = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle2 >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)idle2, 0u));
                if (!has && ri < n) {
                    uint32_t tnd;
                    float tpm;
                    DKey tk;
                    if (co_pop_item(queue + 4 * (size_t)((h + ri) % q_cap), J.seq, h + ri, J.T, tnd, tpm, tk) > 0) {
                        start(tnd, tpm, tk);
                    } else {
                        co_or(&G->status, CO_ST_LOST);
                        err = true;
                    }
                }
                if (__ballot(err)) break;
                continue;
            }
        } else if (ni2 == 0 && (it % (4u * CO_BAL)) == 0) {
            /* no idle lane: hand oldest entries to the queue while it is short */
            const u64 donors = __ballot(sp > 0);
            const uint32_t nd = (uint32_t)__builtin_popcountll(donors);
            if (nd) {
                uint32_t t = 0, n = 0;
                if (lane == 0) {
                    const uint32_t q = co_ld(&G->tail) - co_ld(&G->head);
                    if (q < CO_EXPORT_LOW) {
                        n = CO_EXPORT_LOW - q < nd ? CO_EXPORT_LOW - q : nd;
                        t = co_add(&G->tail, n);
                    }
                }
                t = (uint32_t)__shfl((int)t, 0, 64);
                n = (uint32_t)__shfl((int)n, 0, 64);
                const uint32_t rd = __builtin_amdgcn_mbcnt_hi((uint32_t)(donors >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)donors, 0u));
                if (sp > 0 && rd < n) {
                    DTask tk;
                    take_oldest(tk);
                    if (!co_push_item(queue + 4 * (size_t)((t + rd) % q_cap), J.seq, t + rd, tk.nd, tk.pm, tk.k)) {
                        co_or(&G->status, CO_ST_LOST);
                        err = true;
                    }
                }
                if (__ballot(err)) break;
            }
        }
        if (ni2 == 64u) {
            /* the whole wave is idle: leave once no wave is busy and nothing
               is queued (the busy count is read before the queue: a wave
               queues its entries before it stops being busy) */
            uint32_t stop = 0;
            if (lane == 0) {
                if (wave_busy) co_add(&G->busy, 0xFFFFFFFFu);
                stop = co_ld(&G->busy) == 0u && co_ld(&G->head) >= co_ld(&G->tail);
            }
            wave_busy = false;
            if (__shfl((int)stop, 0, 64)) break;
            if (++idle_rounds > CO_SPIN_LIMIT / 16u) {
                co_or(&G->status, CO_ST_LOST);
                break;
            }
            for (uint32_t k = idle_rounds < 8u ? idle_rounds : 8u; k; k--) __builtin_amdgcn_s_sleep(8);
        }
    }
    if (wave_busy && lane == 0) co_add(&G->busy, 0xFFFFFFFFu);
    return took;
}

/* The leader's side of one cooperative mesh traversal (every thread of the
   leader block, all on the same ray): posts the job, takes part in it,
   waits for the helpers that joined, and merges the blocks' results.
   Returns 1 (hit in tr.closest_mesh_hit), 0 (none), or -1 when the result
   could not be shown equal to the sequential one (a path deeper than the
   order key, a stack limit, a lost wait, or a winner whose path is not
   admitted at t*): the caller then traces sequentially (start state
   restored). */
DFN int coop_trace_mesh(DTrace &tr, const StDScene &sc, const StDMesh &mesh, const uint32_t *nodes,
                        const StDTriHot *tris, uint32_t root_first, bool any_hit) {
    DHit &closest = tr.closest_mesh_hit;
    DCoopGroup *G = co_blk.group;
    const float c0 = closest.distance, cur0 = tr.current_hit.distance;
    const uint32_t T = mesh.triangle_count;
    uint32_t max_steps = 2u * T + 2u;
    if (sc.step_cap && sc.step_cap < max_steps) max_steps = sc.step_cap;
#ifdef CO_STATS
    const u64 st_t0 = __builtin_amdgcn_s_memrealtime();
#endif
    __syncthreads();  /* the previous job's readers of co_blk are done */
    if (co_blk.jobs >= CO_MAX_JOBS) return -1;
    if (threadIdx.x == 0) {
        const uint32_t seq = co_blk.base + ++co_blk.jobs;
        uint32_t v[26];
        DCoopJob &j = *(DCoopJob *)v;
        const DRay &r = tr.local;
        j.ro[0] = __float_as_uint(r.origin.x); j.ro[1] = __float_as_uint(r.origin.y); j.ro[2] = __float_as_uint(r.origin.z);
        j.rd[0] = __float_as_uint(r.direction.x); j.rd[1] = __float_as_uint(r.direction.y);
        j.rd[2] = __float_as_uint(r.direction.z);
        j.rr[0] = __float_as_uint(r.direction_reciprocal.x); j.rr[1] = __float_as_uint(r.direction_reciprocal.y);
        j.rr[2] = __float_as_uint(r.direction_reciprocal.z);
        j.rs[0] = __float_as_uint(r.scaled_origin.x); j.rs[1] = __float_as_uint(r.scaled_origin.y);
        j.rs[2] = __float_as_uint(r.scaled_origin.z);
        j.neg = r.neg_x | (r.neg_y << 1) | (r.neg_z << 2);
        j.c0 = __float_as_uint(c0);
        j.nodes_lo = (uint32_t)(uintptr_t)nodes;
        j.nodes_hi = (uint32_t)((u64)(uintptr_t)nodes >> 32);
        j.tris_lo = (uint32_t)(uintptr_t)tris;
        j.tris_hi = (uint32_t)((u64)(uintptr_t)tris >> 32);
        j.T = T;
        j.pad0 = j.pad1 = 0;
        j.root_first = root_first;
        j.any_hit = any_hit ? 1u : 0u;
        j.max_steps = max_steps;
        j.stack_size = tr.mesh_stack_size;
        j.seq = seq;
        uint32_t *p = (uint32_t *)&G->job;
#pragma unroll
        for (int i = 0; i < 26; i++) co_st(p + i, v[i]);
        co_st(&G->status, 0);
        co_st(&G->head, 0);
        co_st(&G->tail, 0);
        co_st(&G->busy, 1u);   /* the leader's first wave starts with the root */
        co_st64(&G->first_hi, ~0ull);
#ifdef CO_STATS
        for (int i = 0; i < 6; i++) co_st(G->pad + i, 0);
#endif
        /* the job and the reset words are in place before the claim word opens it */
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        co_st64(&G->claim, (u64)seq << 32);
        /* ... and the claim word is written before the block's threads claim from it */
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        co_blk.seq = seq;
    }
    __syncthreads();
    DCoopJobR J;
    J.nodes = nodes;
    J.tris = tris;
    J.c0 = c0;
    J.T = T;
    J.root_first = root_first;
    J.any_hit = any_hit ? 1u : 0u;
    J.max_steps = max_steps;
    J.stack_size = tr.mesh_stack_size;
    J.seq = co_blk.seq;
    J.valid = 1;
    DCoopRes res = {};
    (void)co_work(tr, G, co_blk.queue, co_blk.q_cap, J, res, threadIdx.x < 64u);
    const DCoopRes mine = res_block_reduce(res, any_hit);
    if (threadIdx.x == 0) {
        /* no unit can be claimed from here on; wait for the helpers that
           joined (each of them is running) */
        co_max64(&G->claim, ((u64)J.seq << 32) | CO_CLOSED);
        uint32_t st = 0;
        for (uint32_t spins = 0; co_ld(&G->inflight) != 0u; spins++) {
            if (spins > CO_SPIN_LIMIT) {
                st |= CO_ST_LOST;
                break;
            }
            co_sleep();
        }
        if (co_ld(&G->head) < co_ld(&G->tail)) st |= CO_ST_LOST;   /* a task nobody took */
        co_blk.status = st | co_ld(&G->status);
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    /* the helpers' results */
    DCoopRes all = {};
    if (threadIdx.x == 0) all = mine;
    for (uint32_t b = 1u + threadIdx.x; b < co_blk.gsize; b += blockDim.x) {
        DCoopRes o;
        if (co_get_slot(co_blk.slots + b, J.seq, o)) res_merge(all, o, any_hit);
    }
    const DCoopRes best = res_block_reduce(all, any_hit);
    const uint32_t status = co_blk.status;
#ifdef CO_STATS
    if (threadIdx.x == 0) {
        uint32_t joined = 0;
        for (uint32_t b = 1u; b < co_blk.gsize; b++) joined += co_ld((const uint32_t *)(co_blk.slots + b)) == J.seq;
        co_stat(1ull | ((u64)(blockIdx.x & 0xFFFFFFu) << 8) | ((u64)co_blk.jobs << 32), st_t0,
                __builtin_amdgcn_s_memrealtime(),
                (u64)T | ((u64)status << 32) | ((u64)best.found << 36) | ((u64)any_hit << 37) |
                    ((u64)(co_ld(&G->tail) & 0xFFFFu) << 40) | ((u64)(joined & 0xFFu) << 56));
        co_stat(4ull | ((u64)(blockIdx.x & 0xFFFFFFu) << 8) | ((u64)co_blk.jobs << 32),
                (u64)co_ld(G->pad + 0) | ((u64)co_ld(G->pad + 1) << 32), (u64)co_ld(G->pad + 2) | ((u64)co_ld(G->pad + 3) << 32),
                (u64)best.pm | 0);
    }
#endif
    closest.distance = c0;
    tr.current_hit.distance = cur0;
    if (status & CO_ST_STEPS) {
        trav_fault(sc, ST_FAULT_MESH_STEPS);
        return 0;
    }
    if (status) return -1;
    if (best.found) {
        if (!any_hit && !(best.pm <= best.t)) return -1;
        if (hit_triangles(tr, tris, best.tri, 1u, 0, 0, any_hit)) return 1;
        closest.distance = c0;
        tr.current_hit.distance = cur0;
        return -1;
    }
    if (best.lfound) tr.current_hit.distance = best.ldist;
    return 0;
}

/* A helper block of a redo group (every thread): waits for the leader's
   jobs, takes part in each, and leaves its result in its slot; returns
   once the leader has no more pixels (or no job came for a long time). */
DFN void coop_helper(uint32_t *mesh_stack, uint32_t *aux_stack) {
    DCoopGroup *G = co_blk.group;
    uint32_t last = co_blk.base;   /* thread 0: the latest job seen */
    for (;;) {
        if (threadIdx.x == 0) {
            uint32_t cmd = 2;
            /* a helper that sees no job for CO_HELPER_IDLE_TICKS leaves, so
               that its CU goes to other work (the leader needs no helper) */
            const u64 t_wait = __builtin_amdgcn_s_memrealtime();
            for (uint32_t spins = 0; spins <= CO_SPIN_LIMIT; spins++) {
                if (co_ld(&G->quit) == co_blk.base) break;
                if (__builtin_amdgcn_s_memrealtime() - t_wait > CO_HELPER_IDLE_TICKS) break;
                const u64 c = co_ld64(&G->claim);
                const uint32_t s = (uint32_t)(c >> 32);
                if (s - co_blk.base - 1u < CO_MAX_JOBS && s > last && (uint32_t)c < CO_CLOSED) {
                    /* join, then check that the job is still open: if so, the
                       leader's wait at its end counts this block and the
                       record holds this job until then; if not, the block
                       leaves the job's words alone */
                    co_add(&G->inflight, 1u);
                    const u64 c2 = co_ld64(&G->claim);
                    last = s;
                    if ((uint32_t)(c2 >> 32) == s && (uint32_t)c2 < CO_CLOSED) {
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                        const uint32_t *p = (const uint32_t *)&G->job;
                        for (int i = 0; i < 26; i++) co_blk.jw[i] = co_ld(p + i);
                        if (co_blk.jw[25] == s) {
                            co_blk.seq = s;
                            cmd = 1;
                            break;
                        }
                        co_or(&G->status, CO_ST_LOST);
                    }
                    co_add(&G->inflight, 0xFFFFFFFFu);
                }
                co_sleep();
            }
            co_blk.cmd = cmd;
        }
        __syncthreads();
        if (co_blk.cmd != 1u) break;
        DTrace tr = DTrace{};
        tr.closest_mesh_hit.object_type = ST_PRIM_MESH;
        tr.mesh_stack = mesh_stack;
        tr.aux_stack = aux_stack;
        DCoopJobR J;
        co_job_from_words(co_blk.jw, co_blk.seq, J, tr);
        DCoopRes res = {};
        const bool worked = co_work(tr, G, co_blk.queue, co_blk.q_cap, J, res, false);
        const uint32_t any_worked = __syncthreads_or(worked);
        const DCoopRes r = res_block_reduce(res, J.any_hit != 0u);
        if (threadIdx.x == 0) {
            if (any_worked) co_put_slot(co_blk.slots + co_blk.rank, J.seq, r);
            /* the slot is written before the leader can see this block leave */
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            co_add(&G->inflight, 0xFFFFFFFFu);
        }
        __syncthreads();
    }
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
#ifdef CO_STATS
        if (threadIdx.x == 0) co_stat(3ull | ((u64)blockIdx.x << 8), __builtin_amdgcn_s_memrealtime(), 0, T_STAT(mesh));
#endif
        /* r < 0: the cooperative result could not be shown equal to the
           sequential one; trace sequentially */
    }
    uint32_t left = root_first;
    uint32_t stack_size = 0;
    DStk stk = {stack, sc.stack_mask_m, 0};
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
                stk_push(stk, sc, sc.spill_s, stack_size++, rnext);
                if (stack_size == tr.mesh_stack_size) return false;
            }
            left = lnext;
        } else if (has_right) {
            left = rnext;
        } else {
            if (stack_size == 0) break;
            left = stk_pop(stk, sc, sc.spill_s, --stack_size);
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
    uint32_t stack_size = 0;
    DStk stk = {tr.scene_stack, sc.stack_mask_s, 0};
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
                stk_push(stk, sc, 0u, stack_size++, rnext);
            }
            left = lnext;
        } else if (has_right) {
            left = rnext;
        } else {
            if (stack_size == 0) break;
            left = stk_pop(stk, sc, 0u, --stack_size);
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
