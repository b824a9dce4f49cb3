/*
 * st_wave.h - the per-pixel path of the reference renderer split into
 * wavefront stages (gfx950, wave64).
 *
 * The reference traces one pixel start to finish (render/raytracer.h:81-134,
 * closest_hit/surface.h:146-255): primary ray, then a depth loop that shades
 * direct lights with one shadow ray each, emissive quads, light glow, and a
 * reflection/refraction bounce ray. Run as one GPU thread per pixel start to
 * finish, the lanes of a wave whose pixels stop early would idle while the
 * others bounce. Here the loop is cut at the first bounce: a level of it
 * (shading entry, the shadow rays, the shading from their results) runs in
 * the lane that found the level's surface, with the pixel's state in
 * registers, and the pixels that bounce after the first level are
 * compacted into a queue:
 *
 *   primary   one wave per tile: primary ray, Depth/Normals/UVs output,
 *             and for Beauty the first level (the shadow rays start at the
 *             hit points just found, so the tile's nodes are still in cache)
 *   bounce    per queued pixel: its bounce ray, the next level, and so on
 *             down the depth loop, in one lane
 *
 * Across the queue a pixel's state lives in HBM as planes of float4 (one
 * plane per four state values, slot = tile * 64 + lane), so that every state
 * read and write of a wave is one contiguous 1 KiB access.
 *
 * Results equal the sequential reference bit for bit, including its shared
 * hit state: the shading values are recomputed with the reference's
 * operation order, the closest-hit fields that later code reads after a
 * shadow ray (the last light's squared distance or its occluder's) are
 * carried explicitly, and so is the one field of the reference's
 * `current_hit` that flows from one traversal into a later one (its
 * `distance`, which a box hit leaves unwritten and copies into the closest
 * hit). A shadow ray's traversal does not depend on that value, so each
 * shadow ray of a pixel starts from a fresh hit state and the shading
 * replays their writes to it in light order.
 */
#ifndef ST_WAVE_H
#define ST_WAVE_H

#include "st_trace.h"

/* Per-launch wavefront buffers (device memory, passed by value). */
struct KWave {
    float *st;             /* ST_PLANES planes of n float4 */
    uint32_t n;            /* state slots: 64 per tile of the launch */
    uint32_t lmax;         /* planes of shadow_res (largest light count) */
    uint32_t *shadow_res;  /* [light][slot] 2 words: occluded, current-hit distance bits */
    float *occ_d2;         /* [slot] occluder's squared distance, last light's shadow ray */
    uint32_t *bounce_q;    /* [n] slots of the pixels with a first-level bounce ray: ctr[0] from the
                              front (refracting surfaces), ctr[1] from the back (the others) */
    uint32_t *ctr;         /* [2] lengths of bounce_q's front and back parts */
    uint32_t *counts;      /* [slot] 2 words: shadow rays, bounce rays of the pixel */
    uint32_t *sorted_q;    /* [n] or NULL: the bounce queue binned by request and direction octant
                              (k_bin), front part then back part */
    uint32_t *redo_q;      /* [slots of the launch] pixels to redo in the redo stage (every launch) */
    uint32_t *redo_ctr;    /* [1] length of redo_q */
};


/* state planes */
enum { P_COLOR, P_THRU, P_VO, P_VD, P_POS, P_NRM, P_REF, P_ALB, P_BDIR, P_MISC, ST_PLANES };

#define PF_LIGHTS_SHADED 1u
#define PF_REDO 2u         /* a traversal of this pixel gave up: the redo stage traces it again */


/* Shading state of one pixel between stages (surface.h:146-255 locals). */
struct DPix {
    sv3 color, throughput;
    sv3 vo, vd;            /* s.viewing_origin, s.viewing_direction */
    sv3 pos, nrm;          /* s.position, s.normal */
    sv3 ref, alb;          /* s.reflected_direction, s.albedo */
    sv3 bdir;              /* bounce ray direction */
    float z, closest_distance, cur_dist, max_distance, NdotRd, ior, hit_d2, duv_y;
    uint32_t prim_index, material_id, depth_left, flags;
    uint32_t n_shadow, n_bounce;   /* rays this pixel has traced so far */
};

DFN float4 pl_ld(const KWave &w, uint32_t plane, uint32_t slot) { return ldg4(w.st, (size_t)plane * w.n + slot); }
DFN void pl_st(const KWave &w, uint32_t plane, uint32_t slot, float a, float b, float c, float d) {
    st_vf4 v = {a, b, c, d};
    ((ST_GLOBAL st_vf4 *)w.st)[(size_t)plane * w.n + slot] = v;
}
DFN void pl_st3(const KWave &w, uint32_t plane, uint32_t slot, sv3 a, float d) { pl_st(w, plane, slot, a.x, a.y, a.z, d); }
DFN sv3 f4v3(float4 f) { return V3(f.x, f.y, f.z); }

DFN DPix pix_load(const KWave &w, uint32_t slot) {
    DPix p;
    float4 f;
    f = pl_ld(w, P_COLOR, slot); p.color = f4v3(f); p.throughput.x = f.w;
    f = pl_ld(w, P_THRU, slot); p.throughput.y = f.x; p.throughput.z = f.y; p.z = f.z; p.closest_distance = f.w;
    f = pl_ld(w, P_VO, slot); p.vo = f4v3(f); p.cur_dist = f.w;
    f = pl_ld(w, P_VD, slot); p.vd = f4v3(f); p.max_distance = f.w;
    f = pl_ld(w, P_POS, slot); p.pos = f4v3(f); p.NdotRd = f.w;
    f = pl_ld(w, P_NRM, slot); p.nrm = f4v3(f); p.ior = f.w;
    f = pl_ld(w, P_REF, slot); p.ref = f4v3(f); p.hit_d2 = f.w;
    f = pl_ld(w, P_ALB, slot); p.alb = f4v3(f); p.duv_y = f.w;
    f = pl_ld(w, P_BDIR, slot); p.bdir = f4v3(f); p.prim_index = __float_as_uint(f.w);
    int4 m = ldgi4(w.st, (size_t)P_MISC * w.n + slot);
    p.material_id = (uint32_t)m.x; p.depth_left = (uint32_t)m.y; p.flags = (uint32_t)m.z & 0xFFu;
    p.n_bounce = (uint32_t)m.z >> 8; p.n_shadow = (uint32_t)m.w;
    return p;
}
DFN void pix_store_all(const KWave &w, uint32_t slot, const DPix &p) {
    pl_st3(w, P_COLOR, slot, p.color, p.throughput.x);
    pl_st(w, P_THRU, slot, p.throughput.y, p.throughput.z, p.z, p.closest_distance);
    pl_st3(w, P_VO, slot, p.vo, p.cur_dist);
    pl_st3(w, P_VD, slot, p.vd, p.max_distance);
    pl_st3(w, P_POS, slot, p.pos, p.NdotRd);
    pl_st3(w, P_NRM, slot, p.nrm, p.ior);
    pl_st3(w, P_REF, slot, p.ref, p.hit_d2);
    pl_st3(w, P_ALB, slot, p.alb, p.duv_y);
    pl_st3(w, P_BDIR, slot, p.bdir, __uint_as_float(p.prim_index));
    stgi4(w.st, (size_t)P_MISC * w.n + slot, (int)p.material_id, (int)p.depth_left, (int)(p.flags | (p.n_bounce << 8)),
          (int)p.n_shadow);
}

/* ---- wave-level queue helpers ------------------------------------------ */

DFN uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

/* Append one entry per active lane with `want` to a queue whose length is
   *ctr: one atomic per wave. Returns this lane's index (valid if want). */
DFN uint32_t wave_append(uint32_t *ctr, bool want) {
    const unsigned long long m = __ballot(want);
    uint32_t base = 0;
    if (m) {
        const uint32_t leader = (uint32_t)__builtin_ctzll(m);
        if (lane_id() == leader) base = atomicAdd(ctr, (uint32_t)__builtin_popcountll(m));
        base = __builtin_amdgcn_readlane(base, leader);
    }
    return base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

/* ---- shared shading steps ------------------------------------------------ */

/* Per-light set-up of shade_lights_direct (surface.h:9-35): the vector to the
   light, its squared length and length, and the normalised direction. */
struct DLightDir {
    sv3 pos, dir;
    float d2, dist;
    float4 la, lb;
};
DFN DLightDir light_dir(const StDScene &sc, uint32_t i, sv3 P) {
    DLightDir L;
    const void *lp = sc.lights + i;
    L.la = ldg4(lp, 0);
    L.lb = ldg4(lp, 1);
    L.pos = V3(L.la.x, L.la.y, L.la.z);
    sv3 d = v_sub(L.pos, P);
    L.d2 = v_sqlen(d);
    L.dist = sqrtf(L.d2);
    L.dir = v_scale(d, 1.0f / L.dist);
    return L;
}

DFN DShaded shaded_of(const DPix &p) {
    DShaded s;
    s.albedo = p.alb;
    s.position = p.pos;
    s.normal = p.nrm;
    s.viewing_direction = p.vd;
    s.viewing_origin = p.vo;
    s.reflected_direction = p.ref;
    s.light_direction = V3s(0);
    s.uv = V2(0, 0);
    s.prim_index = p.prim_index;
    for (int j = 0; j < 4; j++) s.quad_vertices[j] = V3s(0);
    return s;
}

/* Fresh traversal state around one ray; `cur_dist` is the reference's
   current_hit.distance carried over from the pixel's earlier traversals. */
DFN void trace_init(DTrace &tr, const StDScene &sc, uint32_t *scene_stack, uint32_t *mesh_stack, float cur_dist,
                    uint32_t *aux_stack = nullptr) {
    tr = DTrace{};
    tr.closest_mesh_hit.object_type = ST_PRIM_MESH;
    tr.current_hit.distance = cur_dist;
    tr.mesh_stack_size = sc.mesh_stack_size;
    tr.scene_stack = scene_stack;
    tr.mesh_stack = mesh_stack;
    tr.aux_stack = aux_stack;
}

/* Per-lane traversal stacks of a stage kernel: scene and mesh columns, and
   for the redo stage the cooperative traversal's path maxima. */
struct DStacks {
    uint32_t *scene, *mesh, *aux;
};

/* The unnormalised primary direction `current` of traced pixel (x, y): the
   row's sequential sum (render/raytracer.h:156-170) continued from its
   checkpoint at the previous multiple of ST_DIR_MARK. */
DFN sv3 primary_dir(const StDView &v, uint32_t x, uint32_t y) {
    const float4 m = ldg4(v.dir_marks, (size_t)y * ST_DIR_MARKS(v.width1) + x / ST_DIR_MARK);
    sv3 current = V3(m.x, m.y, m.z);
    const sv3 right = V3(v.right[0], v.right[1], v.right[2]);
    for (uint32_t k = x % ST_DIR_MARK; k; k--) current = v_add(current, right);
    return current;
}

/* The window pixel's final value (raytracer.h:105-134): light glow along
   the camera ray when no shading pass hit a light, tone mapping, and the
   squared 0-255 colour with the camera-space depth. */
DFN void finish_pixel(const StDScene &sc, const StDView &v, uint32_t pix, sv3 color, bool hit_found,
                      bool lights_shaded, float z, float closest_distance) {
    if (v.mode == ST_MODE_BEAUTY && sc.light_count && !lights_shaded) {
        const sv3 Rd = v_norm(primary_dir(v, pix % v.width1, pix / v.width1));
        const sv3 Ro = V3(v.cam_pos[0], v.cam_pos[1], v.cam_pos[2]);
        DSphereHit sh = {};
        if (shade_light_glow(sc, Ro, Rd, closest_distance, sh, color)) {
            hit_found = true;
            z = sh.closest_hit_distance;
        }
    }
    float4 rgba;
    double depth;
    if (hit_found) {
        if (v.mode == ST_MODE_BEAUTY) {
            color.x = tone_map(color.x);
            color.y = tone_map(color.y);
            color.z = tone_map(color.z);
        }
        color = v_scale(color, 255.0f);
        color = v_mul(color, color);
        rgba = make_float4(color.x, color.y, color.z, 1.0f);
        depth = (double)z;
    } else {
        rgba = make_float4(0, 0, 0, 0);
        depth = (double)INFINITY;
    }
    /* FloatPixel is 24 bytes: three 8-byte words */
    void *px = static_cast<char *>(v.pixels) + 24 * (size_t)pix;
    stg2(px, 0, __float_as_uint(rgba.x), __float_as_uint(rgba.y));
    stg2(px, 1, __float_as_uint(rgba.z), __float_as_uint(rgba.w));
    const unsigned long long db = (unsigned long long)__double_as_longlong(depth);
    stg2(px, 2, (uint32_t)db, (uint32_t)(db >> 32));
}

/* Top of the reference's depth loop (surface.h:190-200) for the hit in
   tr.closest_hit: the surface frame and reflection into the pixel state and
   the number of shadow rays the light loop (surface.h:9-35) will cast (one
   per light facing the surface; trace_level_shadows casts them). The state
   stays in registers: the same lane traces the shadow rays and shades. */
DFN void begin_level(const StDScene &sc, const DMat &M, const DHit &hit, DPix &p) {
    p.pos = hit.position;
    p.nrm = hit.from_behind ? v_neg(hit.normal) : hit.normal;
    p.ior = hit.from_behind ? M.n2_over_n1 : M.n1_over_n2;
    const bool is_ref = (M.is & ST_REFLECTIVE) || (M.is & ST_REFRACTIVE);
    if (is_ref || M.brdf == ST_BRDF_PHONG) {
        p.NdotRd = -inv_dot(p.nrm, p.vd);
        p.ref = reflect_with_dot(p.vd, p.nrm, p.NdotRd);
    }
    for (uint32_t i = 0; i < sc.light_count; i++)
        if (v_dot(p.nrm, light_dir(sc, i, p.pos).dir) > 0) p.n_shadow++;
}

DFN void record_counts(const KWave &w, uint32_t slot, uint32_t n_shadow, uint32_t n_bounce) {
    stg2(w.counts, slot, n_shadow, n_bounce);
}

/* One shadow ray (surface.h:13-25 and trace.h:8-76 any-hit): records
   whether the light is occluded and the last current-hit distance the
   traversal wrote. */
template <bool COOP = false>
DFN void shadow_ray(const KWave &w, const StDScene &sc, DTrace &tr, uint32_t slot, uint32_t light, sv3 pos,
                    const DLightDir &L) {
    tr.closest_hit.position = L.pos;
    tr.closest_hit.distance_squared = L.d2;
    tr.closest_hit.distance = L.dist;
    DRay ray;
    ray.origin = pos;
    ray.direction = L.dir;
    const bool occluded = trace_scene<COOP>(ray, tr, sc, true);
    stg2(w.shadow_res, (size_t)light * w.n + slot, occluded ? 1u : 0u, __float_as_uint(tr.current_hit.distance));
    if (occluded && light + 1 == sc.light_count)
        ((ST_GLOBAL float *)w.occ_d2)[slot] = tr.closest_hit.distance_squared;
}

/* The shadow rays of a pixel whose level just began (surface point pos,
   normal nrm), traced by the lane that found the surface: one per light
   that faces it, each from a fresh hit state. The traversal reads none of
   the pixel's shading state, so the rays need no order among themselves;
   the shading replays their current-hit writes in light order. */
template <bool COOP = false>
DFN bool trace_level_shadows(const KWave &w, const StDScene &sc, const DStacks &stk, uint32_t slot, sv3 pos,
                             sv3 nrm) {
    uint32_t abandoned = 0;
    for (uint32_t i = 0; i < sc.light_count; i++) {
        const DLightDir L = light_dir(sc, i, pos);
        if (!(v_dot(nrm, L.dir) > 0)) continue;
        DTrace tr;
        trace_init(tr, sc, stk.scene, stk.mesh, __uint_as_float(ST_DIST_UNSET), stk.aux);
        shadow_ray<COOP>(w, sc, tr, slot, i, pos, L);
        abandoned |= tr.abandoned;
    }
    return abandoned != 0;
}

/* The shading of one pixel's level (surface.h:200-243): direct light from
   the level's shadow results, emissive quads and glow. Returns true when the pixel
   continues with a bounce ray (direction in p.bdir), false when it is done. */
DFN bool shade_level(const KWave &w, const StDScene &sc, DTrace &tr, uint32_t slot, DPix &p) {
    const DMat M = load_mat(sc, p.material_id);
    const bool is_ref = (M.is & ST_REFLECTIVE) || (M.is & ST_REFRACTIVE);
    DShaded s = shaded_of(p);
    sv3 current_color = is_ref ? V3s(0) : V3(sc.ambient[0], sc.ambient[1], sc.ambient[2]);
    float cur = p.cur_dist;
    float closest_d2 = p.hit_d2;
    for (uint32_t i = 0; i < sc.light_count; i++) {
        const DLightDir L = light_dir(sc, i, s.position);
        closest_d2 = L.d2;
        s.light_direction = L.dir;
        const float NdotL = v_dot(s.normal, s.light_direction);
        if (NdotL <= 0) continue;
        const float intensity = L.la.w / L.d2;
        const uint32_t *res = w.shadow_res + 2 * ((size_t)i * w.n + slot);
        const uint32_t occluded = ldgu(res, 0), dist_bits = ldgu(res, 1);
        if (dist_bits != ST_DIST_UNSET) cur = __uint_as_float(dist_bits);
        if (occluded) {
            if (i + 1 == sc.light_count) closest_d2 = ((const ST_GLOBAL float *)w.occ_d2)[slot];
            continue;
        }
        current_color = v_mad(shade_point(s, M, NdotL), v_scale(V3(L.lb.x, L.lb.y, L.lb.z), intensity), current_color);
    }
    if (sc.has_emissive_quads) {
        tr.current_hit.distance = cur;
        tr.closest_hit.distance_squared = closest_d2;
        DRay ray;
        ray.origin = s.position;
        ray.direction = V3s(0);
        if (shade_emissive_quads(s, M, ray, tr, sc, current_color)) p.max_distance = tr.closest_hit.distance;
        cur = tr.current_hit.distance;
    }
    p.cur_dist = cur;
    p.color = v_mad(current_color, p.throughput, p.color);
    DSphereHit sh = {};
    if (sc.light_count && shade_light_glow(sc, s.viewing_origin, s.viewing_direction, p.max_distance, sh, p.color))
        p.flags |= PF_LIGHTS_SHADED;
    if (!(is_ref && --p.depth_left)) return false;
    if (M.is & ST_REFLECTIVE) {
        p.bdir = s.reflected_direction;
    } else {
        p.bdir = refract_dir(s.viewing_direction, s.normal, p.ior, p.NdotRd);
        if (p.bdir.x == 0 && p.bdir.y == 0 && p.bdir.z == 0) p.bdir = s.reflected_direction;
    }
    return true;
}

#endif
