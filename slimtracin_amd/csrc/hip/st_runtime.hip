/*
 * st_runtime.hip - device residency, kernels and the extern "C" API of the
 * HIP tracing library (gfx950, wave64).
 *
 * Kernels
 *   k_primary_dirs  one lane per traced row and vector component: walks
 *                   the row with the reference's sequential
 *                   `current += right` fp32 sum (render/raytracer.h:156-170)
 *                   and stores the unnormalised `current` vector at every
 *                   ST_DIR_MARK-th pixel (the primary stage finishes the sum
 *                   and normalises it). Rows of every request of a launch
 *                   are flattened into one grid.
 *   k_tile_order    the primary stage's tile claim order: tile order,
 *                   super-tile order, or tiles slower than the mean in the
 *                   previous launch first.
 *   k_primary       one wavefront per pixel tile: primary ray, and for
 *   k_primary4      Beauty the first level of the depth loop (shading
 *                   entry, shadow rays, shading; st_wave.h); the same body
 *                   held to 3 or 4 waves per SIMD. Tiles of all requests
 *                   of a launch form one list; in persistent mode each
 *                   wavefront keeps pulling the next tile from a global
 *                   counter, or first from its XCD's eighth of the claim
 *                   order, until the list is exhausted (the loop's exit
 *                   condition is the same for every wave), otherwise tile =
 *                   wave id with an XCD-aware block order. A single request
 *                   is just a list with one entry, so plain and batched
 *                   launches run the same code. Persistent launches keep
 *                   short traversal stacks in LDS and spill older entries
 *                   to a per-launch buffer.
 *   k_bounce        the Beauty depth loop after its first level, one launch:
 *                   per queued pixel, in one lane, its bounce ray, then the
 *                   new surface's level as k_primary runs the first, and so
 *                   on while the pixel bounces; the part of the queue that
 *                   holds the longest chains first.
 *   k_redo          the pixels whose traversal gave up in the two stages
 *                   above (a mesh traversal past ST_ABANDON_STEPS steps):
 *                   each pixel's whole path again by a group of blocks,
 *                   with large mesh traversals split across the group.
 *   k_counts        per-request shadow/bounce ray totals.
 */
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <vector>

#include "st_wave.h"
#include "../st_ssb.h"
#include "st_lbvh.h"
#include "st_mesh_lbvh.h"
#include "../st_mesh.h"
#include "../host/st_host.h"
#include "slimtracin.h"
#include "st_internal.h"
#include "st_sort.h"

#define WAVES_PER_BLOCK 4
#define BLOCK_THREADS (64 * WAVES_PER_BLOCK)
/* Launch attributes of the primary stage (occupancy experiments): 3 waves
   per SIMD, a register budget of 168 */
#ifndef K_PRIMARY_ATTR
#define K_PRIMARY_ATTR __attribute__((amdgpu_waves_per_eu(3)))
#endif
/* Launch attributes of the bounce stage (occupancy experiments). */
#ifndef K_BOUNCE_ATTR
/* 2 waves per SIMD: the register budget of 256 (a few values spill) */
#define K_BOUNCE_ATTR __attribute__((amdgpu_waves_per_eu(2)))
#endif

/* Every use of the shared stream, arena and staging buffer holds this lock
   (the server's worker and direct callers may run on different threads). */
static std::recursive_mutex g_mu;
#define RT_LOCK std::lock_guard<std::recursive_mutex> rt_lock_(g_mu)

/* ---- errors ------------------------------------------------------------ */

static __thread char g_err[512];

static int set_err(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_err(ST_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));        \
    } while (0)

extern "C" const char *st_last_error(void) { return g_err; }

/* ---- kernels ------------------------------------------------------------ */

/* Per-request launch record shared by both kernels (device copy). */
struct KReq {
    StDView v;
    uint32_t row_base;     /* first flattened row of this request */
    uint32_t tile_count;
    uint32_t out_base;     /* first flattened window pixel (display resolve) */
    uint32_t count_block;  /* first block of k_counts for this request */
    uint32_t *bgra;        /* window image (width1-1)*(height1-1), or NULL */
    uint64_t pad2;
    /* device SSB pass (k_ssb) when device_ssb: camera rotation_matrix_inverted
       (rows X, Y, Z) and focal length; ssb_raw / ssb_flags optionally receive
       the rects as computeSSB leaves them and the flags */
    float ssb_rmi[9], focal_length;
    uint32_t device_ssb, pad3;
    int32_t *ssb_raw;
    uint8_t *ssb_flags;
};

__device__ static inline uint32_t find_req(const KReq *reqs, uint32_t n, uint32_t key, bool by_row) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        uint32_t base = by_row ? reqs[mid].row_base : reqs[mid].v.tile_base;
        if (base <= key) lo = mid; else hi = mid - 1;
    }
    return lo;
}

/* One lane per traced row and component of the direction: walks that
   component of the row's sequential `current += right` sum (a vector sum's
   components are independent fp32 chains, so three lanes run a row's
   chain a third as long each) and stores the value at every
   ST_DIR_MARK-th pixel, into the x, y or z word of the pixel's float4
   checkpoint. A pixel's direction is then its checkpoint plus at most
   ST_DIR_MARK - 1 further additions, the same fp32 operations in the same
   order (primary_dir). */
#define DIRS_BLOCK 192u
__global__ void __launch_bounds__(DIRS_BLOCK) k_primary_dirs(const KReq *reqs, uint32_t nreq, uint32_t total_rows) {
    const uint32_t t = blockIdx.x * DIRS_BLOCK + threadIdx.x;
    const uint32_t row = t / 3u, c = t - 3u * row;
    if (row >= total_rows) return;
    const uint32_t r = find_req(reqs, nreq, row, true);
    const StDView &v = reqs[r].v;
    const uint32_t y = row - reqs[r].row_base;
    if (y >= v.height1) return;
    const float4 rs = ldg4(v.row_starts, y);
    float current = c == 0 ? rs.x : (c == 1 ? rs.y : rs.z);
    const float right = v.right[c];
    const uint32_t marks = ST_DIR_MARKS(v.width1);
    ST_GLOBAL float *out = (ST_GLOBAL float *)v.dir_marks + (size_t)y * marks * 4 + c;
    for (uint32_t m = 0; m < marks; m++) {
        out[(size_t)m * 4] = current;
        if (m + 1 < marks) {
#pragma unroll
            for (uint32_t k = 0; k < ST_DIR_MARK; k++) current = current + right;
        }
    }
}

/* Tile coordinates of launch tile `tile` within its request's tile grid. */
__device__ static inline void tile_coords(const StDView &v, uint32_t tile, uint32_t &tx, uint32_t &ty) {
    const uint32_t local = tile - v.tile_base;
    tx = local % v.tiles_x;
    ty = local / v.tiles_x;
}

/* Launch tile `tile`'s pixel for lane `lane`; false outside the traced grid. */
__device__ static inline bool tile_pixel(const StDView &v, uint32_t tile, uint32_t lane, uint32_t &x, uint32_t &y) {
    uint32_t tx, ty;
    tile_coords(v, tile, tx, ty);
    const uint32_t tw = 1u << v.tile_shift, th = 64u >> v.tile_shift;
    x = tx * tw + (lane & (tw - 1u));
    y = ty * th + (lane >> v.tile_shift);
    return x < v.width1 && y < v.height1;
}

/* One depth-loop level of a pixel (surface.h:190-243) from its surface hit
   `hit`: shading entry, the light loop's shadow rays, then the shading from
   their results (direct light, emissive quads, glow). Returns true when the
   pixel continues with a bounce ray (direction in p.bdir). In the main
   stages a pixel whose shadow ray gave up is left for the redo stage. */
template <bool COOP = false>
__device__ static inline __attribute__((always_inline)) bool run_level(const KWave &w, const StDScene &sc,
                                                                      uint32_t slot, const DMat &M, const DHit &hit,
                                                                      DPix &p, const DStacks &stk) {
    begin_level(sc, M, hit, p);
    if (trace_level_shadows<COOP>(w, sc, stk, slot, p.pos, p.nrm)) {
        p.flags |= PF_REDO;
        return false;
    }
    DTrace tr;
    trace_init(tr, sc, nullptr, nullptr, p.cur_dist);
    const bool bounce = shade_level(w, sc, tr, slot, p);
    if (bounce) p.n_bounce++;
    return bounce;
}

/* A pixel whose depth loop has ended: its frame value and ray counts. */
__device__ static inline void end_pixel(const KWave &w, const StDScene &sc, const StDView &v, uint32_t slot,
                                        const DPix &p) {
    uint32_t x, y;
    tile_pixel(v, slot >> 6, slot & 63u, x, y);
    finish_pixel(sc, v, y * v.width1 + x, p.color, true, (p.flags & PF_LIGHTS_SHADED) != 0, p.z,
                 p.closest_distance);
    record_counts(w, slot, p.n_shadow, p.n_bounce);
}

/* The depth loop after a level that ends in a bounce ray (surface.h:221-240):
   closest-hit traversal of the bounce ray; a hit that is not an emissive
   quad starts the pixel's next level (shading entry, shadow rays, shading)
   with the state in registers, and so on while the pixel bounces. In the
   main stages a traversal that gave up ends the loop with PF_REDO set. */
template <bool COOP = false>
__device__ static inline __attribute__((always_inline)) void bounce_chain(const KWave &w, const StDScene &sc,
                                                                         uint32_t slot, DPix &p,
                                                                         const DStacks &stk) {
    bool again;
    do {
        DTrace tr;
        trace_init(tr, sc, stk.scene, stk.mesh, p.cur_dist, stk.aux);
        DRay ray;
        ray.origin = p.pos;
        ray.direction = p.bdir;
        DHit &hit = tr.closest_hit;
        hit.distance = hit.distance_squared = INFINITY;
        const bool found = trace_scene<COOP>(ray, tr, sc, false);
        if (tr.abandoned) {
            p.flags |= PF_REDO;
            break;
        }
        p.cur_dist = tr.current_hit.distance;
        again = false;
        if (found) {
            p.prim_index = hit.object_id;
            const DMat M = load_mat(sc, hit.material_id);
            p.alb = M.albedo;
            if (M.use && M.texture_count) {
                DShaded s = shaded_of(p);
                sv2 dUV = V2(0, p.duv_y);
                apply_textures(tr, sc, s, M, dUV, false);
                p.alb = s.albedo;
            }
            if (hit.object_type == ST_PRIM_QUAD && (M.is & ST_EMISSIVE)) {
                p.color = hit.from_behind ? V3s(0) : M.emission;
            } else {
                if (M.brdf != ST_BRDF_COOK_TORRANCE) p.throughput = v_mul(p.throughput, M.reflectivity);
                p.vo = ray.origin;
                p.vd = ray.direction;
                p.material_id = hit.material_id;
                p.hit_d2 = hit.distance_squared;
                again = run_level<COOP>(w, sc, slot, M, hit, p, stk);
            }
        }
    } while (again);
}

/* Primary stage of one pixel (raytracer.h:81-134 up to the Beauty depth
   loop): the primary ray over all primitives with the screen-space gate, the
   hit record, and then either the finished pixel (Depth/Normals/UVs,
   misses, emissive hits, depth 0) or the Beauty shading entry
   (surface.h:146-170) and the first level of its depth loop. The main stage
   queues a pixel that bounces for the bounce stage and one whose traversal
   gave up for the redo stage; the redo stage (COOP, a whole block on one
   pixel) runs the pixel's whole depth loop here. `slot` = tile * 64 + lane
   indexes the pixel's depth-loop state. */
template <bool COOP = false>
__device__ static inline __attribute__((always_inline)) void primary_pixel(const StDScene &sc, const StDView &v,
                                                                          uint32_t slot, uint32_t x, uint32_t y,
                                                                          bool valid, const DTileBox &box,
                                                                          const DStacks &stk, const KWave &w) {
    /* Every lane traces, pixels outside the traced grid included (they read
       pixel 0's direction and no rect admits them), so that the whole wave
       takes part in culling the primitives against the tile. */
    const uint32_t pix = valid ? y * v.width1 + x : 0u;
    const sv3 cur = primary_dir(v, valid ? x : 0u, valid ? y : 0u);
    DTrace tr;
    trace_init(tr, sc, stk.scene, stk.mesh, 0.0f, stk.aux);
    DRay ray;
    ray.origin = V3(v.cam_pos[0], v.cam_pos[1], v.cam_pos[2]);
    /* dir = normVec3(current) (render/raytracer.h:159) */
    ray.direction = v_norm(cur);
    ray.direction_reciprocal = v_rcp(ray.direction);
    DHit &hit = tr.closest_hit;
    hit.distance = hit.distance_squared = INFINITY;
    hit.cone_angle = v.cone_angle;
    hit.cone_width = 0;
    tr.primary_triangle = -1;
    const bool found = hit_primitives<COOP>(ray, tr, sc, sc.scene_leaf_ids, sc.prim_count, false, v.rects,
                                            valid ? x : ~0u, valid ? y : ~0u, &box);
    bool redo = tr.abandoned != 0;
    bool entered = false;   /* entered the Beauty depth loop */
    if (valid) {
        if (v.hits)
            stgi4(v.hits, pix, found ? (int)hit.object_id : -1, found ? (int)hit.object_type : 0,
                  found ? tr.primary_triangle : -1, __float_as_int(found ? hit.distance : INFINITY));
        const float closest_distance = found ? hit.distance : INFINITY;
        float z = INFINITY;
        sv3 color = V3s(0);
        if (found) {
            sq4 cr;
            cr.v = V3(v.cam_rot_inv[0], v.cam_rot_inv[1], v.cam_rot_inv[2]);
            cr.w = v.cam_rot_inv[3];
            z = q_rot(v_sub(hit.position, ray.origin), cr).z;
            if (v.mode == ST_MODE_BEAUTY) {
                const DMat M = load_mat(sc, hit.material_id);
                if (M.is & ST_EMISSIVE) {
                    color = hit.from_behind ? V3s(0) : M.emission;
                } else if (v.depth) {
                    DPix p;
                    p.color = V3s(0);
                    p.throughput = V3s(1);
                    p.vo = ray.origin;
                    p.vd = ray.direction;
                    p.ref = V3s(0);
                    p.bdir = V3s(0);
                    p.prim_index = hit.object_id;
                    p.alb = M.albedo;
                    sv2 dUV = V2(0, 0);
                    if (M.use && M.texture_count) {
                        DShaded s = shaded_of(p);
                        apply_textures(tr, sc, s, M, dUV, true);
                        p.alb = s.albedo;
                    }
                    p.duv_y = dUV.y;
                    p.NdotRd = 0;
                    p.ior = 0;
                    p.max_distance = hit.distance;
                    p.hit_d2 = hit.distance_squared;
                    p.z = z;
                    p.closest_distance = closest_distance;
                    p.cur_dist = tr.current_hit.distance;
                    p.material_id = hit.material_id;
                    p.depth_left = v.depth;
                    p.flags = 0;
                    p.n_shadow = p.n_bounce = 0;
                    bool bounce = false;
                    if (!redo) {
                        bounce = run_level<COOP>(w, sc, slot, M, hit, p, stk);
                        redo = (p.flags & PF_REDO) != 0;
                    }
                    if (COOP) {
                        if (bounce) bounce_chain<true>(w, sc, slot, p, stk);
                        end_pixel(w, sc, v, slot, p);
                    } else {
                        /* refracting pixels, which go on for the most levels,
                           fill the queue from the front, the others from the
                           back; the bounce stage starts with the front */
                        const bool go = bounce && !redo;
                        const bool deep = (M.is & ST_REFRACTIVE) != 0;
                        const uint32_t qf = wave_append(w.ctr, go && deep);
                        const uint32_t qb = wave_append(w.ctr + 1, go && !deep);
                        if (go) {
                            ((ST_GLOBAL uint32_t *)w.bounce_q)[deep ? qf : w.n - 1u - qb] = slot;
                            pix_store_all(w, slot, p);
                        } else {
                            end_pixel(w, sc, v, slot, p);
                        }
                    }
                    entered = true;
                }
            } else if (v.mode == ST_MODE_DEPTH) {
                color = V3s(4 / hit.distance);
            } else if (v.mode == ST_MODE_NORMALS) {
                const DMat M = load_mat(sc, hit.material_id);
                if (M.texture_count > 1 && (M.use & ST_NORMAL_MAP)) {
                    sv2 uv = V2(hit.uv.x * M.uv_repeat.x, hit.uv.y * M.uv_repeat.y);
                    hit.uv_area /= M.uv_repeat.x / M.uv_repeat.y;
                    sv2 dUV = V2(0, 0);
                    dUV.x = duv_ray_cone(hit.NdotV, hit.cone_width, hit.area, hit.uv_area);
                    sq4 rot = normal_rotation(sample_normal(sc, M.tex1, uv, dUV));
                    hit.normal = q_rot(hit.normal, rot);
                }
                color = V3(0.5f * (hit.normal.x + 1), 0.5f * (hit.normal.y + 1), 0.5f * (hit.normal.z + 1));
            } else {
                color = V3(0.5f * hit.uv.x, 0.5f * hit.uv.y, 0.5f);
            }
        }
        if (!entered) finish_pixel(sc, v, pix, color, found, false, z, closest_distance);
    }
    if (w.n && !entered) record_counts(w, slot, 0, 0);
    if (!COOP) {
        const bool r = valid && redo;
        const uint32_t q = wave_append(w.redo_ctr, r);
        if (r) ((ST_GLOBAL uint32_t *)w.redo_q)[q] = slot;
    }
}

/* Primary stage of one tile: each lane its pixel (primary_pixel). */
__device__ static inline __attribute__((always_inline)) void primary_tile(const StDScene *scenes, const KReq *reqs,
                                                                         uint32_t nreq, uint32_t tile,
                                                                         const DStacks &stk, const KWave &w) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t r = find_req(reqs, nreq, tile, false);
    const StDView v = reqs[r].v;
    const StDScene sc = scenes[v.scene];
    uint32_t x, y;
    const bool valid = tile_pixel(v, tile, lane, x, y);
    DTileBox box;
    {
        uint32_t tx, ty;
        tile_coords(v, tile, tx, ty);
        const uint32_t tw = 1u << v.tile_shift, th = 64u >> v.tile_shift;
        box.x0 = (int32_t)(tx * tw);
        box.y0 = (int32_t)(ty * th);
        box.x1 = box.x0 + (int32_t)tw;
        box.y1 = box.y0 + (int32_t)th;
    }
    primary_pixel<false>(sc, v, tile * 64u + lane, x, y, valid, box, stk, w);
}

/* Traversal stacks live in LDS, one column per lane (no bank conflicts):
   [wave][entry][lane]. */
#define ST_LDS_STACKS(scene_entries, mesh_entries)                                                              \
    extern __shared__ uint32_t st_lds[];                                                                        \
    uint32_t *wave_base_ = st_lds + (size_t)(threadIdx.x >> 6) * 64u * ((scene_entries) + (mesh_entries)) +     \
                           (threadIdx.x & 63u);                                                                 \
    uint32_t *scene_stack = wave_base_;                                                                         \
    uint32_t *mesh_stack = wave_base_ + 64u * (scene_entries)

/* words of the per-launch counters (zeroed at launch): the XCD-aware
   claim's 8 chunk counters from XCD_CLAIM on, each on a 128-byte line of
   its own (atomics to one line serialise) */
#define CLAIM_WORDS 1024u
#define XCD_MIN_TILES 32u   /* tiles per wave below which a launch claims from one counter */
#define XCD_CLAIM 64u
#define XCD_CLAIM_STRIDE 32u

/* XCD-aware tile claim (StTuning.persistent 2): the claim order in 8
   contiguous chunks. The blocks of XCD k (a launch's workgroups are dealt
   to the 8 XCDs round-robin) take chunk k's tiles first, so each XCD's L2
   serves a compact part of the frames instead of every 8th tile of a
   region 8 times as large, then chunk k+1's, k+2's, ... as each runs out.
   chunk and tried are the wave's state (first chunk blockIdx.x mod 8, none
   tried); every failed claim moves to the next chunk, so a wave makes at
   most 8 of them and the counters stay below tiles + 8 x waves. Returns
   total once all 8 chunks are done. */
__device__ static inline uint32_t claim_xcd(uint32_t *ctr, uint32_t total, uint32_t &chunk, uint32_t &tried) {
    while (tried < 8u) {
        const uint32_t lo = (uint32_t)((uint64_t)total * chunk / 8u);
        const uint32_t hi = (uint32_t)((uint64_t)total * (chunk + 1u) / 8u);
        const uint32_t p = atomicAdd(ctr + XCD_CLAIM_STRIDE * chunk, 1u);
        if (p < hi - lo) return lo + p;
        chunk = (chunk + 1u) & 7u;
        tried++;
    }
    return total;
}

__device__ static inline __attribute__((always_inline)) void
primary_stage(const StDScene *__restrict__ scenes, const KReq *__restrict__ reqs, uint32_t nreq, uint32_t total_tiles,
              uint32_t *tile_counter, uint32_t persistent, uint32_t xcd_blocks, uint32_t scene_entries,
              uint32_t mesh_entries, const KWave &w, const uint32_t *__restrict__ tile_order,
              uint32_t *__restrict__ tile_cost) {
    ST_LDS_STACKS(scene_entries, mesh_entries);
    const DStacks stk = {scene_stack, mesh_stack, nullptr};
    /* One call site, so primary_tile is inlined once. Persistent (1): waves
       claim tiles from the global counter until it runs past the end; 2:
       from their XCD's chunk of the claim order first (claim_xcd).
       Otherwise each wave renders one tile: blocks are dealt to the 8 XCDs
       round-robin, so each XCD gets a contiguous run of tiles and
       neighbouring tiles share its L2. */
    const uint32_t b = blockIdx.x;
    uint32_t t = ((b % 8u) * (xcd_blocks / 8u) + b / 8u) * WAVES_PER_BLOCK + (threadIdx.x >> 6);
    uint32_t chunk = b & 7u, tried = 0;
    for (uint32_t round = 0;; round++) {
        if (persistent) {
            uint32_t c = 0;
            if ((threadIdx.x & 63u) == 0)
                c = persistent == 2u ? claim_xcd(tile_counter + XCD_CLAIM, total_tiles, chunk, tried)
                                     : atomicAdd(tile_counter, 1u);
            t = c;
        } else if (round > 0) {
            break;
        }
        /* the tile is wave-uniform: say so, so that the request, view and
           scene records load into scalar registers */
        t = __builtin_amdgcn_readfirstlane(t);
        if (t >= total_tiles) break;
        /* claim order -> tile (k_tile_order) */
        t = __builtin_amdgcn_readfirstlane(ldgu(tile_order, t));
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        primary_tile(scenes, reqs, nreq, t, stk, w);
        /* the tile's time (100 MHz ticks), which orders the next launch */
        if ((threadIdx.x & 63u) == 0)
            ((ST_GLOBAL uint32_t *)tile_cost)[t] = (uint32_t)(__builtin_amdgcn_s_memrealtime() - t0);
    }
}

#define K_PRIMARY_ARGS                                                                                                \
    const StDScene *__restrict__ scenes, const KReq *__restrict__ reqs, uint32_t nreq, uint32_t total_tiles,         \
        uint32_t *tile_counter, uint32_t persistent, uint32_t xcd_blocks, uint32_t scene_entries,                     \
        uint32_t mesh_entries, KWave w, const uint32_t *__restrict__ tile_order, uint32_t *__restrict__ tile_cost
/* The primary stage at the occupancy its registers allow (3 waves per
   SIMD), and held to 4 waves per SIMD (128 registers, a few values in
   scratch) for launches whose traversal stacks leave LDS for 4 waves
   (StTuning.primary_waves): with 16-slot stack windows the headline batch's
   primary stage took 30.0 ms at 4 waves against 34.2 ms at 3. */
__global__ void __launch_bounds__(BLOCK_THREADS) K_PRIMARY_ATTR k_primary(K_PRIMARY_ARGS) {
    primary_stage(scenes, reqs, nreq, total_tiles, tile_counter, persistent, xcd_blocks, scene_entries, mesh_entries,
                  w, tile_order, tile_cost);
}
__global__ void __launch_bounds__(BLOCK_THREADS) __attribute__((amdgpu_waves_per_eu(4))) k_primary4(K_PRIMARY_ARGS) {
    primary_stage(scenes, reqs, nreq, total_tiles, tile_counter, persistent, xcd_blocks, scene_entries, mesh_entries,
                  w, tile_order, tile_cost);
}

/* Redo stage: the pixels whose traversal gave up in the main stages, each
   traced again by a group of blocks: its leader block runs the pixel's
   primary ray and whole depth loop again (every thread computing the same
   values, so the block's barriers are reached by all of its threads), with
   every traversal of a large mesh split across the group's blocks
   (coop_trace_mesh; the other blocks of the group run coop_helper). With
   fewer pixels than blocks each pixel gets a group of gridDim / pixels
   blocks; with more, each block is a group of one and takes every
   gridDim-th pixel. The pixel's frame value, hit record and ray counts
   replace what the main stages wrote for it. Each thread has a scene, a
   mesh and a path-maximum stack column in LDS. */
#define REDO_THREADS 256u
struct KCoop {
    char *base;            /* group records: group_stride bytes each (CO_HEADER, then the task queue) */
    DCoopSlot *slots;      /* one per block of the launch */
    uint32_t group_stride, q_cap;
    const uint32_t *epoch; /* launch number (k_tile_order) */
};
__global__ void __launch_bounds__(REDO_THREADS) k_redo(const StDScene *__restrict__ scenes,
                                                       const KReq *__restrict__ reqs, uint32_t nreq, KWave w,
                                                       uint32_t scene_entries, uint32_t mesh_entries, KCoop co) {
    ST_LDS_STACKS(scene_entries, 2u * mesh_entries);
    const DStacks stk = {scene_stack, mesh_stack, mesh_stack + 64u * mesh_entries};
    const uint32_t count = __builtin_amdgcn_readfirstlane(ldgu(w.redo_ctr, 0));
    if (count == 0) return;
    const uint32_t blocks = gridDim.x;
    const uint32_t ngroups = count < blocks ? count : blocks;
    const uint32_t gsize = blocks / ngroups;
    const uint32_t grp = blockIdx.x / gsize, rank = blockIdx.x - grp * gsize;
    if (grp >= ngroups) return;
    DCoopGroup *rec = (DCoopGroup *)(co.base + (size_t)grp * co.group_stride);
    if (threadIdx.x == 0) {
        co_blk.group = rec;
        co_blk.queue = (u64 *)((char *)rec + CO_HEADER);
        co_blk.slots = co.slots + (size_t)grp * gsize;
        co_blk.gsize = gsize;
        co_blk.rank = rank;
        co_blk.q_cap = co.q_cap;
        co_blk.base = co_ld(co.epoch) << 12;
        co_blk.jobs = 0;
    }
    __syncthreads();
    if (rank) {
        coop_helper(stk.mesh, stk.aux);
        return;
    }
    /* the leader: a clean queue and record before its first job */
    for (uint32_t i = threadIdx.x; i < 4u * co.q_cap; i += blockDim.x) co_st64(co_blk.queue + i, 0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        co_st(&rec->inflight, 0);
        co_st(&rec->status, 0);
        co_st(&rec->head, 0);
        co_st(&rec->tail, 0);
        co_st(&rec->busy, 0);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        co_st64(&rec->claim, ((u64)co_blk.base << 32) | CO_CLOSED);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    }
    __syncthreads();
    for (uint32_t i = grp; i < count; i += ngroups) {
        const uint32_t slot = __builtin_amdgcn_readfirstlane(ldgu(w.redo_q, i));
        const uint32_t tile = slot >> 6;
        const uint32_t r = find_req(reqs, nreq, tile, false);
        const StDView v = reqs[r].v;
        const StDScene sc = scenes[v.scene];
        uint32_t x, y;
        if (!tile_pixel(v, tile, slot & 63u, x, y)) continue;
        const DTileBox box = {(int32_t)x, (int32_t)y, (int32_t)x + 1, (int32_t)y + 1};
#ifdef CO_STATS
        const u64 st_t0 = __builtin_amdgcn_s_memrealtime();
#endif
        primary_pixel<true>(sc, v, slot, x, y, true, box, stk, w);
#ifdef CO_STATS
        if (threadIdx.x == 0)
            co_stat(2ull | ((u64)blockIdx.x << 8) | ((u64)gsize << 32), st_t0, __builtin_amdgcn_s_memrealtime(),
                    (u64)slot | ((u64)v.mode << 32) | ((u64)v.depth << 40));
#endif
    }
    __syncthreads();
    if (threadIdx.x == 0) co_st(&rec->quit, co_blk.base);
}

/* Redo stage launch shape: blocks of up to REDO_THREADS threads, as many
   waves as the LDS stack columns (scene, mesh and path maxima per thread)
   leave room for beside the cooperative traversal's block state. */
static void redo_shape(uint32_t s_entries, uint32_t m_entries, uint32_t &threads, size_t &lds) {
    const size_t per_wave = (size_t)64 * (s_entries + 2 * m_entries) * sizeof(uint32_t);
    uint32_t waves = REDO_THREADS / 64u;
    while (waves > 1 && per_wave * waves > 160 * 1024 - sizeof(DCoopBlk) - 256) waves >>= 1;
    threads = 64u * waves;
    lds = per_wave * waves;
}

/* queued tasks per group (main launches; st_render_noalloc keeps fewer) */
#define REDO_QUEUE 2048u
#define REDO_QUEUE_NOALLOC 256u
static size_t redo_group_stride(uint32_t q_cap) { return CO_HEADER + (size_t)q_cap * 32u; }
/* device bytes of the redo stage's group records and slots */
static size_t redo_slots_offset(uint32_t blocks, uint32_t q_cap) {
    return ((size_t)blocks * redo_group_stride(q_cap) + 255) & ~(size_t)255;
}
static size_t redo_coop_bytes(uint32_t blocks, uint32_t q_cap) {
    return redo_slots_offset(blocks, q_cap) + (((size_t)blocks * sizeof(DCoopSlot) + 255) & ~(size_t)255);
}

/* Sum of the previous launch's tile times (cost_sum zeroed before): one
   atomic per block. */
#define COST_THREADS 256u
__global__ void __launch_bounds__(COST_THREADS) k_cost_total(const uint32_t *__restrict__ tile_cost, uint32_t tiles,
                                                             unsigned long long *cost_sum) {
    __shared__ unsigned long long part[COST_THREADS / 64];
    unsigned long long c = 0;
    for (uint32_t i = blockIdx.x * COST_THREADS + threadIdx.x; i < tiles; i += gridDim.x * COST_THREADS)
        c += ldgu(tile_cost, i);
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (uint32_t k = 0; k < COST_THREADS / 64; k++) t += part[k];
        if (t) atomicAdd(cost_sum, t);
    }
}

/* Claim order of the primary stage's tiles, from each tile's time in the
   previous launch of the same shape (written by k_primary): tiles that took
   longer than the mean (cost_sum / tiles, k_cost_total) first, then the
   others. A launch ends when its slowest tiles end, so those should start
   first and the cheap ones fill the end (on the 200k-triangle frame the
   tiles over the mesh take 200-600 us, the floor's a few tens). Without a
   previous launch of this shape the order is the tile order (mode 0), or
   (mode 2) the super-tile order of super_tile. One thread per tile; each
   block keeps its tiles in order within each group and claims its ranges
   with one atomic per group (ctr[0] from the front, ctr[1] from the back,
   both zero at launch). Any cost values give a permutation of the tiles,
   so the order never changes a result. Thread 0 also advances the launch
   number `epoch` the redo stage tags its cooperative jobs with. */
#define ORDER_THREADS 1024u
#define ORDER_TILE 0u      /* tile order: each request's tiles row by row */
#define ORDER_HISTORY 1u   /* slowest first, from the previous launch's tile times */
#define ORDER_SUPER 2u     /* super-tile order */
#define ST_SUPER 8u        /* super-tile: ST_SUPER x ST_SUPER tiles */

/* Launch tile at claim position i in super-tile order: each request's tile
   rows in bands of ST_SUPER, each band in blocks of ST_SUPER tile columns
   (narrower at the right edge, lower at the bottom), the tiles of a block
   row by row. A bijection on each request's tiles, so that a run of claims
   covers a compact part of the frame rather than a few full-width rows. */
__device__ static inline uint32_t super_tile(const KReq *reqs, uint32_t nreq, uint32_t i) {
    const StDView &v = reqs[find_req(reqs, nreq, i, false)].v;
    const uint32_t w = v.tiles_x, th = 64u >> v.tile_shift, h = (v.height1 + th - 1u) / th;
    const uint32_t local = i - v.tile_base;
    const uint32_t band = local / (ST_SUPER * w), q = local - band * ST_SUPER * w;
    const uint32_t hb = min(ST_SUPER, h - band * ST_SUPER);
    const uint32_t col = q / (ST_SUPER * hb), q2 = q - col * ST_SUPER * hb;
    const uint32_t wc = min(ST_SUPER, w - col * ST_SUPER);
    return v.tile_base + (band * ST_SUPER + q2 / wc) * w + col * ST_SUPER + q2 % wc;
}

__global__ void __launch_bounds__(ORDER_THREADS) k_tile_order(const uint32_t *__restrict__ tile_cost,
                                                              uint32_t *__restrict__ tile_order, uint32_t tiles,
                                                              const unsigned long long *cost_sum, uint32_t *ctr,
                                                              uint32_t mode, uint32_t *epoch,
                                                              const KReq *__restrict__ reqs, uint32_t nreq) {
    __shared__ uint32_t s_h[ORDER_THREADS / 64], s_l[ORDER_THREADS / 64], s_base[2];
    const uint32_t i = blockIdx.x * ORDER_THREADS + threadIdx.x;
    /* a new launch number for the redo stage's group records */
    if (i == 0) ((ST_GLOBAL uint32_t *)epoch)[0] += 1u;
    if (mode != ORDER_HISTORY) {
        if (i < tiles) ((ST_GLOBAL uint32_t *)tile_order)[i] = mode == ORDER_SUPER ? super_tile(reqs, nreq, i) : i;
        return;
    }
    const uint32_t wv = threadIdx.x >> 6;
    const unsigned long long total = *(const volatile unsigned long long *)cost_sum;
    const bool live = i < tiles;
    const bool heavy = live && (unsigned long long)ldgu(tile_cost, i) * tiles > total;
    const unsigned long long mh = __ballot(heavy), ml = __ballot(live && !heavy);
    const uint32_t rh = __builtin_amdgcn_mbcnt_hi((uint32_t)(mh >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mh, 0u));
    const uint32_t rl = __builtin_amdgcn_mbcnt_hi((uint32_t)(ml >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ml, 0u));
    if ((threadIdx.x & 63u) == 0) {
        s_h[wv] = (uint32_t)__builtin_popcountll(mh);
        s_l[wv] = (uint32_t)__builtin_popcountll(ml);
    }
    __syncthreads();
    uint32_t bh = 0, bl = 0, nh = 0, nl = 0;
    for (uint32_t k = 0; k < ORDER_THREADS / 64; k++) {
        if (k < wv) { bh += s_h[k]; bl += s_l[k]; }
        nh += s_h[k];
        nl += s_l[k];
    }
    if (threadIdx.x == 0) {
        s_base[0] = atomicAdd(ctr, nh);
        s_base[1] = atomicAdd(ctr + 1, nl);
    }
    __syncthreads();
    if (heavy) ((ST_GLOBAL uint32_t *)tile_order)[s_base[0] + bh + rh] = i;
    else if (live) ((ST_GLOBAL uint32_t *)tile_order)[tiles - 1u - (s_base[1] + bl + rl)] = i;
}

/* word of the per-launch counters holding the bounce stage's chunk claims */
#define BOUNCE_CLAIM 32u
/* words of the per-launch counters holding k_tile_order's two claims */
#define ORDER_CLAIM 40u

/* Wave loop over `chunks` chunks of 64 queue entries. entry(chunk, lane,
   slot, aux) decodes one entry into its state slot (and an extra word) and
   says whether it is live. Each wave's first chunk is its wave id; after that a wave claims
   the next chunk from *claim (zero at launch; one atomic per chunk, so
   few: same-address atomics serialise). A wave whose chunk runs long thus
   takes no further chunks, which shortens the launch's tail, and every
   wave leaves after its first claim past the end. Lanes are grouped by
   request and f(request, slot, aux) runs once per group with a
   wave-uniform request, so the view and scene records load into scalar
   registers. */
template <class E, class F>
__device__ static inline __attribute__((always_inline)) void drain_queue(const KReq *reqs, uint32_t nreq,
                                                                        uint32_t chunks, uint32_t *claim,
                                                                        E &&entry, F &&f) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = gridDim.x * WAVES_PER_BLOCK;
    uint32_t chunk = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES_PER_BLOCK + (threadIdx.x >> 6));
    while (chunk < chunks) {
        uint32_t slot = 0, aux = 0;
        bool pending = entry(chunk, lane, slot, aux);
        const uint32_t r = pending ? find_req(reqs, nreq, slot >> 6, false) : 0u;
        while (__ballot(pending)) {
            if (pending) {
                const uint32_t r0 = __builtin_amdgcn_readfirstlane(r);
                if (r == r0) {
                    pending = false;
                    f(r0, slot, aux);
                }
            }
        }
        uint32_t c = 0;
        if (lane == 0) c = atomicAdd(claim, 1u);
        chunk = waves + __builtin_amdgcn_readfirstlane(c);
    }
}

/* Bounce stage (surface.h:221-240), one launch for every further level:
   closest-hit traversal of each queued bounce ray; a hit that is not an
   emissive quad starts the pixel's next level (shading entry, shadow rays,
   shading) in the same lane, and a pixel that bounces again goes straight
   on to its next bounce ray with its state in registers. No level waits for
   the previous level's slowest wave, which at these depths, with a few
   rays per SIMD, sets a level's time. The stage ends with its longest
   chains (a chunk whose 64 pixels all bounce to the full depth over large
   meshes), so the part of the queue that holds them should start first:
   back_first 0 takes the refracting pixels (the queue's front part) first,
   1 the others (its back part; on the headline batch the reflecting
   pixels' chains are the long ones: 4.9 against 5.3 ms). */
__global__ void __launch_bounds__(BLOCK_THREADS) K_BOUNCE_ATTR k_bounce(const StDScene *__restrict__ scenes,
                                                          const KReq *__restrict__ reqs, uint32_t nreq, KWave w,
                                                          uint32_t *claim, uint32_t scene_entries,
                                                          uint32_t mesh_entries, uint32_t back_first) {
    ST_LDS_STACKS(scene_entries, mesh_entries);
    const DStacks stk = {scene_stack, mesh_stack, nullptr};
    /* the queue's front part (ctr[0] entries) and its back part (ctr[1]
       entries, stored from the end of the buffer down), in back_first's order */
    const uint32_t n_front = __builtin_amdgcn_readfirstlane(ldgu(w.ctr, 0));
    const uint32_t n_back = __builtin_amdgcn_readfirstlane(ldgu(w.ctr, 1));
    const uint32_t c_front = (n_front + 63u) / 64u, c_back = (n_back + 63u) / 64u;
    const uint32_t c_first = back_first ? c_back : c_front;
    const uint32_t chunks = w.sorted_q ? (n_front + n_back + 63u) / 64u : c_front + c_back;
    drain_queue(reqs, nreq, chunks, claim,
                [&](uint32_t chunk, uint32_t lane, uint32_t &slot, uint32_t &) {
                    if (w.sorted_q) {
                        /* binned (k_bin): one list */
                        const uint32_t qi = chunk * 64u + lane;
                        if (qi >= n_front + n_back) return false;
                        slot = ldgu(w.sorted_q, qi);
                        return true;
                    }
                    const bool back = (chunk < c_first) == (back_first != 0u);
                    const uint32_t qi = (chunk < c_first ? chunk : chunk - c_first) * 64u + lane;
                    if (qi >= (back ? n_back : n_front)) return false;
                    slot = ldgu(w.bounce_q, back ? w.n - 1u - qi : qi);
                    return true;
                },
                [&](uint32_t r, uint32_t slot, uint32_t) {
                    const StDView v = reqs[r].v;
                    const StDScene sc = scenes[v.scene];
                    DPix p = pix_load(w, slot);
                    bounce_chain<false>(w, sc, slot, p, stk);
                    end_pixel(w, sc, v, slot, p);
                    /* a traversal gave up: the redo stage traces the pixel again */
                    if (p.flags & PF_REDO)
                        ((ST_GLOBAL uint32_t *)w.redo_q)[atomicAdd(w.redo_ctr, 1u)] = slot;
                });
}

/* Ray sort of the bounce queue (autotuned, StTuning.ray_sort): a counting
   sort of its entries into bins by part (the refracting front, then the
   back), request and bounce direction octant (sign bits). A wave of the
   bounce stage then takes rays of one request leaving in one octant, which
   keeps its traversals together in the tree and its view and scene records
   scalar. pass 0 counts the bins, pass 1 (after an exclusive scan of the
   counts) places each entry; entries of one bin land in any order (a
   pixel's result does not depend on its place in the queue). */
#define BIN_THREADS 256u
__global__ void __launch_bounds__(BIN_THREADS) k_bin(const KReq *__restrict__ reqs, uint32_t nreq, KWave w,
                                                     uint32_t *__restrict__ bins, uint32_t pass) {
    const uint32_t n_front = ldgu(w.ctr, 0), total = n_front + ldgu(w.ctr, 1);
    for (uint32_t i = blockIdx.x * BIN_THREADS + threadIdx.x; i < total; i += gridDim.x * BIN_THREADS) {
        const uint32_t part = i < n_front ? 0u : 1u;
        const uint32_t slot = ldgu(w.bounce_q, part ? w.n - 1u - (i - n_front) : i);
        const uint32_t r = find_req(reqs, nreq, slot >> 6, false);
        const float4 bd = pl_ld(w, P_BDIR, slot);
        const uint32_t oct = (__builtin_signbit(bd.x) ? 1u : 0u) | (__builtin_signbit(bd.y) ? 2u : 0u) |
                             (__builtin_signbit(bd.z) ? 4u : 0u);
        const uint32_t b = (part * nreq + r) * 8u + oct;
        if (pass == 0) atomicAdd(&bins[b], 1u);
        else ((ST_GLOBAL uint32_t *)w.sorted_q)[atomicAdd(&bins[b], 1u)] = slot;
    }
}

/* Per-request shadow and bounce ray totals from the per-slot counts: block b
   sums COUNT_TILES tiles of one request (reqs[r].count_block = its first) and
   adds them with one atomic per counter. */
#define COUNT_TILES 64u
__global__ void __launch_bounds__(256) k_counts(const KReq *__restrict__ reqs, uint32_t nreq,
                                                const uint32_t *__restrict__ counts) {
    __shared__ unsigned long long part[2][4];
    uint32_t lo = 0, hi = nreq - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (reqs[mid].count_block <= blockIdx.x) lo = mid; else hi = mid - 1;
    }
    const KReq &q = reqs[lo];
    const uint32_t t0 = q.v.tile_base + (blockIdx.x - q.count_block) * COUNT_TILES;
    const uint32_t t1 = min(t0 + COUNT_TILES, q.v.tile_base + q.tile_count);
    unsigned long long sh = 0, bo = 0;
    /* one 8-byte load per slot, unrolled so that a lane's loads are all in
       flight together */
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
    const ST_GLOBAL u32x2 *c2 = (const ST_GLOBAL u32x2 *)counts + (size_t)t0 * 64;
    const uint32_t n_slots = (t1 - t0) * 64u;
#pragma unroll 16
    for (uint32_t i = threadIdx.x; i < n_slots; i += 256) {
        const u32x2 c = c2[i];
        sh += c.x;
        bo += c.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sh += __shfl_xor(sh, off, 64);
        bo += __shfl_xor(bo, off, 64);
    }
    if ((threadIdx.x & 63u) == 0) {
        part[0][threadIdx.x >> 6] = sh;
        part[1][threadIdx.x >> 6] = bo;
    }
    __syncthreads();
    if (threadIdx.x == 0 && q.v.counters) {
        atomicAdd(q.v.counters + 1, part[0][0] + part[0][1] + part[0][2] + part[0][3]);
        atomicAdd(q.v.counters + 2, part[1][0] + part[1][1] + part[1][2] + part[1][3]);
    }
}

/* Screen-space bounds on the device (updateSceneSSB, render/SSB.h:137-173),
   one block per request, the same operations as the host pass (st_ssb.h):
   the gate rect of every primitive into the request's rect array. */
__global__ void __launch_bounds__(256) k_ssb(const StDScene *__restrict__ scenes, const KReq *__restrict__ reqs) {
    const KReq &q = reqs[blockIdx.x];
    if (!q.device_ssb) return;
    const StDView &v = q.v;
    const StDScene sc = scenes[v.scene];
    const StSsbDims d = st_ssb_dims(v.width1, v.height1);
    sm3 rmi;
    rmi.X = V3(q.ssb_rmi[0], q.ssb_rmi[1], q.ssb_rmi[2]);
    rmi.Y = V3(q.ssb_rmi[3], q.ssb_rmi[4], q.ssb_rmi[5]);
    rmi.Z = V3(q.ssb_rmi[6], q.ssb_rmi[7], q.ssb_rmi[8]);
    const sv3 cam = V3(v.cam_pos[0], v.cam_pos[1], v.cam_pos[2]);
    for (uint32_t i = threadIdx.x; i < sc.prim_count; i += blockDim.x) {
        const DPrim p = load_prim(sc, i);
        uint8_t flags = (uint8_t)p.flags;
        float mb[6] = {0, 0, 0, 0, 0, 0};
        if (p.type == ST_PRIM_MESH) {
            const float *m = sc.mesh_aabbs + 6 * (size_t)p.mesh_id;
            for (int k = 0; k < 6; k++) mb[k] = ((const ST_GLOBAL float *)m)[k];
        }
        int32_t gate[4], raw[4];
        st_ssb_prim(p.type, &flags, p.pos, p.scale, mb, cam, rmi, q.focal_length, &d, gate, raw);
        stgi4(const_cast<int32_t *>(v.rects), i, gate[0], gate[1], gate[2], gate[3]);
        if (q.ssb_raw) stgi4(q.ssb_raw, i, raw[0], raw[1], raw[2], raw[3]);
        if (q.ssb_flags) ((ST_GLOBAL uint8_t *)q.ssb_flags)[i] = flags;
    }
}

/* Picking (the first-press branch of manipulateSelection,
   viewport/manipulation.h:55-124), one lane per window point: the closed-form
   ray of setRayFromCoords (render/shaders/common.h:170-175), the closest
   primitive over all primitives with the SSB gate at the point, then each
   light's glow sphere (radius intensity/128) in front of it; the view-space
   z of the hit through the inverted camera rotation. Fresh hit state per
   point. */
struct KPick {
    float start[3], right[3], down[3];
    float rmi[9];                  /* camera rotation_matrix_inverted, rows X, Y, Z */
    uint32_t n;
    const uint32_t *xy;            /* n (x, y) pairs */
    StPick *out;
};

__global__ void __launch_bounds__(64) k_pick(const StDScene *__restrict__ scenes, const KReq *__restrict__ reqs,
                                             KPick pk, uint32_t mesh_entries) {
    ST_LDS_STACKS(1u, mesh_entries);
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= pk.n) return;
    const StDView v = reqs[0].v;
    const StDScene sc = scenes[0];
    const int32_t x = (int32_t)ldgu(pk.xy, 2 * (size_t)i), y = (int32_t)ldgu(pk.xy, 2 * (size_t)i + 1);
    DTrace tr;
    trace_init(tr, sc, scene_stack, mesh_stack, 0.0f);
    DRay ray;
    ray.origin = V3(v.cam_pos[0], v.cam_pos[1], v.cam_pos[2]);
    ray.direction = v_smad(V3(pk.right[0], pk.right[1], pk.right[2]), (float)x, V3(pk.start[0], pk.start[1], pk.start[2]));
    ray.direction = v_smad(V3(pk.down[0], pk.down[1], pk.down[2]), (float)y, ray.direction);
    ray.direction = v_norm(ray.direction);
    ray.direction_reciprocal = v_rcp(ray.direction);
    DHit &hit = tr.closest_hit;
    hit.distance = hit.distance_squared = INFINITY;
    uint32_t type = 0, id = 0;
    bool found = hit_primitives(ray, tr, sc, sc.scene_leaf_ids, sc.prim_count, false, v.rects, (uint16_t)x,
                                (uint16_t)y);
    if (tr.abandoned) trav_fault(sc, ST_FAULT_MESH_STEPS);
    if (found) {
        type = hit.object_type;
        id = hit.object_id;
    }
    for (uint32_t l = 0; l < sc.light_count; l++) {
        const float4 la = ldg4(sc.lights + l, 0);
        const float light_radius = la.w * (1.0f / (8.0f * 16.0f));
        tr.sphere_hit.furthest = hit.distance / light_radius;
        if (hit_sphere_simple(ray.origin, ray.direction, V3(la.x, la.y, la.z), 1.0f / light_radius, tr.sphere_hit)) {
            hit.distance = tr.sphere_hit.t_near * light_radius;
            hit.position = v_smad(ray.direction, hit.distance, ray.origin);
            type = ST_PRIM_LIGHT;
            id = l;
            found = true;
        }
    }
    float out[6] = {0, 0, 0, 0, 0, 0};
    if (found) {
        sm3 m;
        m.X = V3(pk.rmi[0], pk.rmi[1], pk.rmi[2]);
        m.Y = V3(pk.rmi[3], pk.rmi[4], pk.rmi[5]);
        m.Z = V3(pk.rmi[6], pk.rmi[7], pk.rmi[8]);
        out[0] = __uint_as_float(type);
        out[1] = __uint_as_float(id);
        out[2] = hit.position.x;
        out[3] = hit.position.y;
        out[4] = hit.position.z;
        out[5] = v_mat(v_sub(hit.position, ray.origin), m).z;
    }
    void *o = pk.out + i;
    stg2(o, 0, __float_as_uint(out[0]), __float_as_uint(out[1]));
    stg2(o, 1, __float_as_uint(out[2]), __float_as_uint(out[3]));
    stg2(o, 2, __float_as_uint(out[4]), __float_as_uint(out[5]));
}

/* Display resolve of one window pixel (drawViewportToWindowContent,
   viewport/viewport.h:12-66, QCAA branch): the 4 traced corners are averaged
   as (TL.c*TL.o + (TR.c*TR.o + (BL.c*BL.o + BR.c*BR.o))) * 0.25 and each
   channel is (u8)sqrt(c), 255 above 255^2. (u8) of the double square root is
   floor(sqrt(c)): computed exactly in integers from the correctly rounded
   fp32 root (squares up to 255^2 are exact in fp32). Background (all four
   depths infinite) is (0,0,0,255); alpha is 255 everywhere; NaN or negative
   channels give 0. Bytes B, G, R, A. */
__device__ static inline uint32_t resolve_channel(float c) {
    if (c > 65025.0f) return 255u;
    if (!(c >= 0.0f)) return 0u;
    uint32_t s = (uint32_t)sqrtf(c);
    if (s > 255u) s = 255u;
    while (s < 255u && (float)((s + 1u) * (s + 1u)) <= c) s++;
    while (s > 0u && (float)(s * s) > c) s--;
    return s;
}

__global__ void __launch_bounds__(256) k_resolve(const KReq *reqs, uint32_t nreq, uint32_t total_out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total_out) return;
    uint32_t lo = 0, hi = nreq - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (reqs[mid].out_base <= i) lo = mid; else hi = mid - 1;
    }
    const KReq &r = reqs[lo];
    if (!r.bgra) return;
    const uint32_t w1 = r.v.width1, w = w1 - 1u;
    const uint32_t local = i - r.out_base;
    if (local >= w * (r.v.height1 - 1u)) return;
    const uint32_t x = local % w, y = local / w;
    const char *base = static_cast<const char *>(r.v.pixels);
    const size_t tl = (size_t)y * w1 + x, corner[4] = {tl, tl + 1, tl + w1, tl + w1 + 1};
    float c[4][4];
    bool background = true;
    for (int k = 0; k < 4; k++) {
        /* 24-byte records are 8-byte aligned: three 8-byte loads */
        const float2 rg = *reinterpret_cast<const float2 *>(base + 24 * corner[k]);
        const float2 bo = *reinterpret_cast<const float2 *>(base + 24 * corner[k] + 8);
        const double d = *reinterpret_cast<const double *>(base + 24 * corner[k] + 16);
        c[k][0] = rg.x; c[k][1] = rg.y; c[k][2] = bo.x; c[k][3] = bo.y;
        background = background && d == (double)INFINITY;
    }
    uint32_t px = 0xFF000000u;
    if (!background) {
        uint32_t ch[3];
        for (int k = 0; k < 3; k++) {
            float acc = c[3][k] * c[3][3];
            acc = c[2][k] * c[2][3] + acc;
            acc = c[1][k] * c[1][3] + acc;
            acc = c[0][k] * c[0][3] + acc;
            ch[k] = resolve_channel(acc * 0.25f);
        }
        px |= ch[2] | (ch[1] << 8) | (ch[0] << 16);
    }
    r.bgra[local] = px;
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
__global__ void k_zero_u64(unsigned long long *p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

/* ---- context ------------------------------------------------------------ */

struct Arena {
    char *base;
    size_t cap, used;
};

#define ST_STAGES_MAX 64
/* Everything one launch in flight owns: its stream, the per-launch device
   buffers, the pinned staging of its uploads and readbacks, the primary
   stage's tile history and the events that time it. The direct API runs on
   the library's main context (under the library lock); a server keeps one
   context per launch it can have in flight, so launches on different
   contexts overlap on the device. */
struct LaunchCtx {
    hipStream_t stream;
    Arena work;            /* per-launch device buffers (grow-only) */
    char *staging;         /* pinned host staging: uploads, then the ray-count readback */
    size_t staging_cap;
    uint32_t *tile_cost;   /* [tile_cap] each tile's time in the last primary stage */
    uint32_t *tile_order;  /* [tile_cap] the primary stage's claim order (k_tile_order) */
    uint32_t tile_cap;
    uint64_t tile_sig;     /* shape of the launch whose times tile_cost holds (0: none) */
    unsigned long long *cost_sum; /* sum of tile_cost over the last launch (k_cost_total) */
    uint32_t *tile_counter; /* CLAIM_WORDS words zeroed per launch: [0] the persistent primary stage's
                               tile counter, [BOUNCE_CLAIM] the bounce stage's chunk claims (128 B
                               apart), [ORDER_CLAIM] k_tile_order's, [XCD_CLAIM + 32k] chunk k's */
    hipEvent_t ev[4];
    int stage_timing;      /* record an event after every stage launch */
    int graph_mode;        /* capture launches into a graph, replay while the launch shape repeats */
    hipGraphExec_t graph_exec;
    std::vector<uint64_t> graph_sig;
    hipEvent_t stage_ev[ST_STAGES_MAX + 1];
    uint32_t stage_n;
    float stage_ms[ST_STAGES_MAX];
    uint32_t last_redo;    /* pixels the redo stage traced again in the last launch */
};

static struct {
    int ready;
    int device;
    LaunchCtx main;        /* the direct API's launches, scene builds and copies */
    int cus;               /* compute units of the device */
    StTuning tune;         /* launch shape chosen by st_set_tuning / st_autotune */
    uint32_t step_cap;     /* st_debug_step_cap: lower traversal step bound, 0 = none */
    uint32_t abandon_steps, coop_min_tris;   /* st_debug_redo (0: ST_ABANDON_STEPS, ST_COOP_MIN_TRIS) */
    std::atomic<uint64_t> graph_captures, graph_launches;
} g;

/* The per-launch traversal limits of a scene table entry. */
static void launch_limits(StDScene &d, uint32_t step_cap, uint32_t abandon_steps, uint32_t coop_min_tris,
                          uint32_t mask_s = ~0u, uint32_t mask_m = ~0u) {
    d.stack_mask_s = mask_s;
    d.stack_mask_m = mask_m;
    d.step_cap = step_cap;
    uint32_t lim = abandon_steps ? abandon_steps : ST_ABANDON_STEPS;
    if (step_cap && step_cap < lim) lim = step_cap;
    d.main_step_limit = lim;
    d.coop_min_tris = coop_min_tris ? coop_min_tris : ST_COOP_MIN_TRIS;
}

/* LDS slots and ring mask of a traversal stack of `need` entries kept in a
   window of `win` slots (0, or a window at least the stack: all of it) */
static void stack_window(uint32_t need, uint32_t win, uint32_t &slots, uint32_t &mask) {
    if (win && win < need) {
        slots = win;
        mask = win - 1u;
    } else {
        slots = need;
        mask = ~0u;
    }
}

/* Device buffers of st_render_noalloc: one array per GPU, part of the code
   object, held until the process ends. In host code `scratch` names a host
   stand-in, not device memory: host code uses only the address
   hipGetSymbolAddress returns for the current device (g.noalloc_scratch). */
static __device__ __align__(256) char scratch[ST_NOALLOC_SCRATCH_BYTES];
#define ST_MAX_DEVICES 64
static char *g_noalloc_scratch[ST_MAX_DEVICES];
/* host side of a st_render_noalloc launch: the upload block (request,
   scene record, row starts) and the ray counts read back */
#define ST_NOALLOC_HOST_BYTES (1u << 21)
static char g_noalloc_host[ST_NOALLOC_HOST_BYTES] __attribute__((aligned(256)));
static unsigned long long g_noalloc_counts[5];

static int ctx_init(LaunchCtx *c) {
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_TRY(hipMalloc(&c->tile_counter, CLAIM_WORDS * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&c->cost_sum, 256));
    HIP_TRY(hipMemset(c->cost_sum, 0, 256));
    for (int i = 0; i < 4; i++) HIP_TRY(hipEventCreate(&c->ev[i]));
    return ST_OK;
}

static void ctx_release(LaunchCtx *c) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->work.base) (void)hipFree(c->work.base);
    if (c->staging) (void)hipHostFree(c->staging);
    if (c->tile_cost) (void)hipFree(c->tile_cost);
    if (c->cost_sum) (void)hipFree(c->cost_sum);
    if (c->tile_counter) (void)hipFree(c->tile_counter);
    for (int i = 0; i < 4; i++)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    for (int i = 0; i <= ST_STAGES_MAX; i++)
        if (c->stage_ev[i]) (void)hipEventDestroy(c->stage_ev[i]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

static int ensure_init(void) {
    if (g.ready) return ST_OK;
    return st_init(0);
}

extern "C" int st_device_info(StDeviceInfo *info) {
    RT_LOCK;
    if (!info) return set_err(ST_ERR_ARG, "null info");
    int rc = ensure_init();
    if (rc) return rc;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, g.device));
    memset(info, 0, sizeof *info);
    snprintf(info->arch, sizeof info->arch, "%s", prop.gcnArchName);
    /* some runtimes report no marketing name: the architecture stands in */
    if (prop.name[0])
        snprintf(info->name, sizeof info->name, "%s", prop.name);
    else
        snprintf(info->name, sizeof info->name, "AMD GPU %.*s", (int)strcspn(info->arch, ":"), info->arch);
    info->device = g.device;
    info->compute_units = (uint32_t)g.cus;
    info->hbm_bytes = (uint64_t)prop.totalGlobalMem;
    return ST_OK;
}

extern "C" int st_init(int device) {
    RT_LOCK;
    if (g.ready) {
        if (device == g.device) return ST_OK;
        return set_err(ST_ERR_ARG, "library already initialised on device %d", g.device);
    }
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return set_err(ST_ERR_ARG, "device %d not present (%d devices)", device, count);
    HIP_TRY(hipSetDevice(device));
    int rc = ctx_init(&g.main);
    if (rc) return rc;
    g.tune.tile_shift = 3;          /* 8 x 8 tiles */
    g.tune.blocks_per_cu = 4;
    g.tune.persistent = 2;          /* persistent, each XCD claiming from its own chunk first ... */
    g.tune.cost_order = 0;          /* ... of the claim order in tile order */
    g.tune.ray_sort = 0;
    g.tune.stack_entries = 16;      /* 16-slot LDS stack windows ... */
    g.tune.primary_waves = 4;       /* ... leave room for 4 waves per SIMD in the primary stage */
    g.tune.bounce_order = 1;        /* the bounce stage starts with the pixels that do not refract */
    hipDeviceProp_t prop;
    g.cus = 256;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        g.cus = prop.multiProcessorCount;
    /* st_render_noalloc's buffers: this device's copy of the scratch array */
    if (device >= ST_MAX_DEVICES) return set_err(ST_ERR_ARG, "device %d above %d", device, ST_MAX_DEVICES);
    void *sp = NULL;
    HIP_TRY(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
    g_noalloc_scratch[device] = (char *)sp;
    g.device = device;
    g.ready = 1;
    return ST_OK;
}

void st_rt_lock(void) { g_mu.lock(); }
void st_rt_unlock(void) { g_mu.unlock(); }

int st_rt_ensure_init(void) { return ensure_init(); }
hipStream_t st_rt_stream(void) { return g.main.stream; }

int st_rt_ctx_create(LaunchCtx **out) {
    *out = NULL;
    int rc = ensure_init();
    if (rc) return rc;
    LaunchCtx *c = new LaunchCtx();
    if ((rc = ctx_init(c))) {
        ctx_release(c);
        delete c;
        return rc;
    }
    *out = c;
    return ST_OK;
}

void st_rt_ctx_destroy(LaunchCtx *c) {
    if (!c) return;
    ctx_release(c);
    delete c;
}
int st_rt_error(int code, const char *msg) { return set_err(code, "%s", msg); }
extern "C" int st_host_error(int code, const char *msg) { return set_err(code, "%s", msg); }

static int arena_reserve(Arena *a, size_t bytes) {
    if (a->cap >= bytes) { a->used = 0; return ST_OK; }
    if (a->base) HIP_TRY(hipFree(a->base));
    a->base = NULL;
    a->cap = 0;
    size_t cap = bytes + bytes / 4 + (1u << 20);
    HIP_TRY(hipMalloc(&a->base, cap));
    a->cap = cap;
    a->used = 0;
    return ST_OK;
}
static void *arena_take(Arena *a, size_t bytes) {
    size_t off = (a->used + 255) & ~(size_t)255;
    a->used = off + bytes;
    return a->base + off;
}
static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

static int staging_reserve(LaunchCtx *c, size_t bytes) {
    if (c->staging_cap >= bytes) return ST_OK;
    if (c->staging) HIP_TRY(hipHostFree(c->staging));
    c->staging = NULL;
    c->staging_cap = 0;
    size_t cap = bytes + bytes / 4 + 65536;
    HIP_TRY(hipHostMalloc((void **)&c->staging, cap, hipHostMallocDefault));
    c->staging_cap = cap;
    return ST_OK;
}

/* ---- resident scenes ---------------------------------------------------- */

static int build_device_meshes(StSceneHandle *h);

extern "C" int st_scene_create_ex(const void *blob, size_t size, uint32_t flags, StSceneHandle **out) {
    if (!out || !blob) return set_err(ST_ERR_ARG, "null argument");
    *out = NULL;
    RT_LOCK;
    int rc = ensure_init();
    if (rc) return rc;
    if ((flags & ST_SCENE_MESH_LBVH) && (flags & ST_SCENE_MESH_BINNED))
        return set_err(ST_ERR_ARG, "ST_SCENE_MESH_LBVH and ST_SCENE_MESH_BINNED together");
    const char *why = NULL;
    uint32_t builder = (flags & ST_SCENE_MESH_LBVH)     ? ST_MESH_BUILD_DEVICE
                       : (flags & ST_SCENE_MESH_BINNED) ? ST_MESH_BUILD_BINNED
                                                        : ST_MESH_BUILD_SAH;
    StScene *sc = st_scene_from_blob_ex(blob, size, builder, &why);
    if (!sc) return set_err(ST_ERR_SCENE, "invalid scene: %s", why ? why : "unknown");
    StSceneHandle *h = (StSceneHandle *)calloc(1, sizeof(StSceneHandle));
    if (!h) { st_scene_free(sc); return set_err(ST_ERR_MEMORY, "out of host memory"); }
    h->host = sc;
    rc = st_scene_upload(h);
    if (!rc) rc = build_device_meshes(h);
    if (rc) { st_scene_destroy(h); return rc; }
    *out = h;
    return ST_OK;
}

extern "C" int st_scene_create(const void *blob, size_t size, StSceneHandle **out) {
    return st_scene_create_ex(blob, size, 0, out);
}

/* Device LBVH builds of the meshes the host left unbuilt: geometry goes to a
   scratch buffer, nodes and triangle records straight into the scene's
   allocation, then the heights come back for the traversal stack size. */
static int build_device_meshes(StSceneHandle *h) {
    StScene *sc = h->host;
    size_t geom = 0, work = 0, sort_bytes = 0;
    int any = 0;
    for (uint32_t m = 0; m < sc->mesh_count; m++) {
        if (!sc->mesh_on_device[m]) continue;
        const StMeshPart *s = sc->deferred + m;
        const StMeshRec *q = &s->rec;
        size_t gb = round256(12ull * q->vertex_count) + round256(12ull * q->triangle_count) +
                    (q->normals_count ? round256(12ull * q->normals_count) + round256(12ull * q->triangle_count) : 0) +
                    (q->uvs_count ? round256(8ull * q->uvs_count) + round256(12ull * q->triangle_count) : 0);
        size_t sb = 0;
        size_t wb = st_mesh_lbvh_work_bytes(q->triangle_count, &sb);
        if (gb > geom) geom = gb;
        if (wb > work) { work = wb; sort_bytes = sb; }
        any = 1;
    }
    if (!any) return ST_OK;
    char *buf = NULL;
    hipError_t e = hipMalloc(&buf, geom + work + 256);
    if (e != hipSuccess) return set_err(ST_ERR_MEMORY, "hipMalloc(%zu) for mesh builds failed", geom + work);
    uint32_t *height_d = (uint32_t *)(buf + geom + work);
    int rc = ST_OK;
    uint32_t max_height = 0;
    for (uint32_t m = 0; m < sc->mesh_count && !rc; m++) {
        if (!sc->mesh_on_device[m]) { if (sc->meshes[m].height > max_height) max_height = sc->meshes[m].height; continue; }
        const StMeshPart *s = sc->deferred + m;
        const StMeshRec *q = &s->rec;
        char *p = buf;
        StMeshGeomDev in;
        memset(&in, 0, sizeof in);
        in.T = q->triangle_count;
        struct { const void *src; size_t bytes; const void **dst; } up[6] = {
            {s->pos, 12ull * q->vertex_count, (const void **)&in.pos},
            {s->pidx, 12ull * q->triangle_count, (const void **)&in.pidx},
            {q->normals_count ? s->nrm : NULL, 12ull * q->normals_count, (const void **)&in.nrm},
            {q->normals_count ? s->nidx : NULL, 12ull * q->triangle_count, (const void **)&in.nidx},
            {q->uvs_count ? s->uvs : NULL, 8ull * q->uvs_count, (const void **)&in.uvs},
            {q->uvs_count ? s->uidx : NULL, 12ull * q->triangle_count, (const void **)&in.uidx},
        };
        for (int k = 0; k < 6 && !rc; k++) {
            if (!up[k].src || !up[k].bytes) continue;
            if (hipMemcpyAsync(p, up[k].src, up[k].bytes, hipMemcpyHostToDevice, g.main.stream) != hipSuccess)
                rc = set_err(ST_ERR_HIP, "mesh geometry upload failed");
            *up[k].dst = p;
            p += round256(up[k].bytes);
        }
        if (rc) break;
        const StDMesh &dm = sc->meshes[m];
        uint32_t *nodes = (uint32_t *)h->dev.mesh_nodes + 8ull * dm.node_offset;
        StDTriHot *hot = (StDTriHot *)h->dev.tri_hot + dm.tri_offset;
        StDTriCold *cold = (StDTriCold *)h->dev.tri_cold + dm.tri_offset;
        if (st_mesh_lbvh_build(g.main.stream, in, buf + geom, sort_bytes, nodes, hot, cold, height_d)) {
            rc = set_err(ST_ERR_HIP, "mesh %u: device BVH build failed: %s", m, hipGetErrorString(hipGetLastError()));
            break;
        }
        uint32_t height = 0;
        if (hipMemcpyAsync(&height, height_d, 4, hipMemcpyDeviceToHost, g.main.stream) != hipSuccess ||
            hipStreamSynchronize(g.main.stream) != hipSuccess) {
            rc = set_err(ST_ERR_HIP, "mesh %u: device BVH build failed", m);
            break;
        }
        if (height == 0 || height + 2 > ST_MESH_STACK_MAX) {
            rc = set_err(ST_ERR_SCENE, "mesh %u: BVH height %u exceeds the device traversal stack", m, height);
            break;
        }
        sc->meshes[m].height = height;
        if (height > max_height) max_height = height;
        memset(sc->deferred + m, 0, sizeof(StMeshPart));
    }
    (void)hipStreamSynchronize(g.main.stream);
    (void)hipFree(buf);
    if (rc) return rc;
    sc->max_mesh_height = max_height;
    sc->mesh_stack_size = (uint8_t)(max_height + 2); /* core/init.h:523-527 */
    h->dev.mesh_stack_size = sc->mesh_stack_size;
    HIP_TRY(hipMemcpy((void *)h->dev.meshes, sc->meshes, sc->mesh_count * sizeof(StDMesh), hipMemcpyHostToDevice));
    return ST_OK;
}

/* Mesh m's BVH node slots and Triangle records (leaf order) as resident on
   the device. A device-built mesh has 2T-1 slots (1 for one triangle);
   slots below a multi-triangle leaf are unused. */
extern "C" int st_scene_mesh_bvh(const StSceneHandle *h, uint32_t m, StBVHNodeRec *nodes, uint32_t node_cap,
                                 StTriangleRec *tris, uint32_t *node_count, uint32_t *height) {
    RT_LOCK;
    if (!h || !h->host || !h->dbuf || !nodes || !tris || !node_count || !height)
        return set_err(ST_ERR_ARG, "null argument");
    const StScene *sc = h->host;
    if (m >= sc->mesh_count) return set_err(ST_ERR_ARG, "no mesh %u", m);
    const StDMesh &dm = sc->meshes[m];
    uint32_t next = (m + 1 < sc->mesh_count) ? sc->meshes[m + 1].node_offset : sc->mesh_node_total;
    uint32_t count = sc->mesh_on_device[m] ? (dm.triangle_count > 1 ? 2 * dm.triangle_count - 1 : 1)
                                           : next - dm.node_offset;
    if (!sc->mesh_on_device[m]) {
        /* host-built: trailing unused slots (odd-offset padding) are not part of the tree */
        while (count > 1 && sc->mesh_nodes[dm.node_offset + count - 1].first_child_id == 0 &&
               sc->mesh_nodes[dm.node_offset + count - 1].child_count == 0 &&
               sc->mesh_nodes[dm.node_offset + count - 1].depth == 0)
            count--;
    }
    if (count > node_cap) return set_err(ST_ERR_ARG, "node buffer holds %u of %u slots", node_cap, count);
    HIP_TRY(hipStreamSynchronize(g.main.stream));
    HIP_TRY(hipMemcpy(nodes, (const uint32_t *)h->dev.mesh_nodes + 8ull * dm.node_offset, 32ull * count,
                      hipMemcpyDeviceToHost));
    std::vector<StDTriHot> hot(dm.triangle_count);
    std::vector<StDTriCold> cold(dm.triangle_count);
    HIP_TRY(hipMemcpy(hot.data(), h->dev.tri_hot + dm.tri_offset, sizeof(StDTriHot) * dm.triangle_count,
                      hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cold.data(), h->dev.tri_cold + dm.tri_offset, sizeof(StDTriCold) * dm.triangle_count,
                      hipMemcpyDeviceToHost));
    for (uint32_t t = 0; t < dm.triangle_count; t++) st_tri_join(&hot[t], &cold[t], tris + t);
    *node_count = count;
    *height = dm.height;
    return ST_OK;
}

/* One allocation per scene. BVH node arrays start 32 bytes past a 64-byte
   boundary: node 0 (the root) sits alone and every child pair (2k+1, 2k+2)
   covers exactly one 64-byte segment. */
extern "C" int st_scene_upload(StSceneHandle *h) {
    RT_LOCK;
    if (!h || !h->host) return set_err(ST_ERR_ARG, "scene has nothing to upload");
    StScene *sc = h->host;
    size_t o_prims = 0;
    size_t o_mats = o_prims + round256((sc->prim_count + 1) * sizeof(StDPrim));
    size_t o_lights = o_mats + round256((sc->material_count + 1) * sizeof(StDMat));
    size_t o_snodes = o_lights + round256((sc->light_count + 1) * sizeof(StDLight));
    size_t o_sleaf = o_snodes + round256(32 + (2 * (size_t)sc->prim_count + 2) * sizeof(StBVHNodeRec));
    size_t o_meshes = o_sleaf + round256(((size_t)sc->prim_count + 1) * 4);
    size_t o_mnodes = o_meshes + round256((sc->mesh_count + 1) * sizeof(StDMesh));
    size_t o_hot = o_mnodes + round256(((size_t)sc->mesh_node_total + 1) * sizeof(StBVHNodeRec));
    size_t o_cold = o_hot + round256(((size_t)sc->tri_total + 1) * sizeof(StDTriHot));
    size_t o_tex = o_cold + round256(((size_t)sc->tri_total + 1) * sizeof(StDTriCold));
    size_t o_mips = o_tex + round256((sc->texture_count + 1) * sizeof(StDTex));
    size_t o_texels = o_mips + round256((sc->mip_total + 1) * sizeof(StDMip));
    size_t o_maabb = o_texels + round256(sc->texel_bytes + 16);
    size_t total = o_maabb + round256((sc->mesh_count + 1) * 24ull);
    if (h->dbuf) HIP_TRY(hipFree(h->dbuf));
    h->dbuf = NULL;
    hipError_t e = hipMalloc(&h->dbuf, total);
    if (e != hipSuccess) return set_err(ST_ERR_MEMORY, "hipMalloc(%zu) for scene failed: %s", total, hipGetErrorString(e));
    h->dbytes = total;
    char *b = h->dbuf;
    /* mesh nodes: the host array already has odd per-mesh offsets; index 0
       is a dummy node, so the device array starts on a 64-byte boundary */
    struct { size_t off; const void *src; size_t bytes; } parts[] = {
        {o_prims, sc->prims, sc->prim_count * sizeof(StDPrim)},
        {o_mats, sc->mats, sc->material_count * sizeof(StDMat)},
        {o_lights, sc->lights, sc->light_count * sizeof(StDLight)},
        {o_snodes + 32, sc->scene_nodes, (size_t)sc->scene_node_count * sizeof(StBVHNodeRec)},
        {o_sleaf, sc->scene_leaf_ids, (size_t)sc->prim_count * 4},
        {o_meshes, sc->meshes, sc->mesh_count * sizeof(StDMesh)},
        {o_mnodes, sc->mesh_nodes, sc->mesh_host_data ? (size_t)sc->mesh_node_total * sizeof(StBVHNodeRec) : 0},
        {o_hot, sc->tri_hot, sc->mesh_host_data ? (size_t)sc->tri_total * sizeof(StDTriHot) : 0},
        {o_cold, sc->tri_cold, sc->mesh_host_data ? (size_t)sc->tri_total * sizeof(StDTriCold) : 0},
        {o_tex, sc->textures, sc->texture_count * sizeof(StDTex)},
        {o_mips, sc->mips, sc->mip_total * sizeof(StDMip)},
        {o_texels, sc->texels, sc->texel_bytes},
    };
    for (size_t i = 0; i < sizeof parts / sizeof parts[0]; i++)
        if (parts[i].bytes) HIP_TRY(hipMemcpy(b + parts[i].off, parts[i].src, parts[i].bytes, hipMemcpyHostToDevice));
    if (sc->mesh_count) {
        float *ma = (float *)malloc(24ull * sc->mesh_count);
        for (uint32_t m = 0; m < sc->mesh_count; m++) {
            const sbox *bb = sc->mesh_aabbs + m;
            float v[6] = {bb->min.x, bb->min.y, bb->min.z, bb->max.x, bb->max.y, bb->max.z};
            memcpy(ma + 6 * m, v, 24);
        }
        hipError_t ce = hipMemcpy(b + o_maabb, ma, 24ull * sc->mesh_count, hipMemcpyHostToDevice);
        free(ma);
        if (ce != hipSuccess) return set_err(ST_ERR_HIP, "mesh AABB upload failed: %s", hipGetErrorString(ce));
    }
    h->d_mesh_aabbs = (float *)(b + o_maabb);
    h->bvh_on_device_only = 0;
    StDScene &d = h->dev;
    memset(&d, 0, sizeof d);
    d.mesh_aabbs = h->d_mesh_aabbs;
    d.prims = (const StDPrim *)(b + o_prims);
    d.mats = (const StDMat *)(b + o_mats);
    d.lights = (const StDLight *)(b + o_lights);
    d.scene_nodes = (const uint32_t *)(b + o_snodes + 32);
    d.scene_leaf_ids = (const uint32_t *)(b + o_sleaf);
    d.meshes = (const StDMesh *)(b + o_meshes);
    d.mesh_nodes = (const uint32_t *)(b + o_mnodes);
    d.tri_hot = (const StDTriHot *)(b + o_hot);
    d.tri_cold = (const StDTriCold *)(b + o_cold);
    d.textures = (const StDTex *)(b + o_tex);
    d.mips = (const StDMip *)(b + o_mips);
    d.texels = (const uint8_t *)(b + o_texels);
    d.prim_count = sc->prim_count;
    d.light_count = sc->light_count;
    d.mesh_count = sc->mesh_count;
    d.texture_count = sc->texture_count;
    memcpy(d.ambient, sc->ambient, 12);
    d.mesh_stack_size = sc->mesh_stack_size;
    d.has_emissive_quads = sc->has_emissive_quads;
    d.scene_stack_need = sc->scene_real_height + 1;
    d.stack_mask_s = d.stack_mask_m = ~0u;   /* whole stacks in LDS unless a launch sets a window */
    d.spill = NULL;
    d.spill_s = d.spill_m = 0;
    return ST_OK;
}

extern "C" void st_scene_destroy(StSceneHandle *h) {
    RT_LOCK;
    if (!h) return;
    if (h->dbuf) (void)hipFree(h->dbuf);
    if (h->lbvh_work) (void)hipFree(h->lbvh_work);
    if (h->d_height) (void)hipFree(h->d_height);
    st_seam_state_free(h->seam);
    st_scene_free(h->host);
    free(h);
}

extern "C" int st_scene_info(const StSceneHandle *h, StSceneInfo *info) {
    if (!h || !info) return set_err(ST_ERR_ARG, "null argument");
    memset(info, 0, sizeof *info);
    if (!h->host) return set_err(ST_ERR_ARG, "scene is not resident yet");
    const StScene *sc = h->host;
    info->prim_count = sc->prim_count;
    info->mesh_count = sc->mesh_count;
    info->triangle_count = sc->tri_total;
    info->texture_count = sc->texture_count;
    info->scene_node_count = sc->scene_node_count;
    info->scene_height = sc->scene_height;
    info->mesh_node_count = sc->mesh_node_total;
    info->max_mesh_height = sc->max_mesh_height;
    info->device_bytes = h->dbytes;
    return ST_OK;
}

extern "C" int st_scene_bvh(const StSceneHandle *h, StBVHNodeRec *nodes, uint32_t *leaf_ids, uint32_t *node_count,
                            uint32_t *height) {
    RT_LOCK;
    if (!h || !nodes || !leaf_ids || !node_count || !height) return set_err(ST_ERR_ARG, "null argument");
    if (!h->host) return set_err(ST_ERR_ARG, "scene is not resident yet");
    const StScene *sc = h->host;
    if (h->bvh_on_device_only) {
        HIP_TRY(hipStreamSynchronize(g.main.stream));
        HIP_TRY(hipMemcpy(sc->scene_nodes, h->dev.scene_nodes, sizeof(StBVHNodeRec) * sc->scene_node_count,
                          hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(sc->scene_leaf_ids, h->dev.scene_leaf_ids, 4u * sc->prim_count, hipMemcpyDeviceToHost));
        ((StSceneHandle *)h)->bvh_on_device_only = 0;
    }
    memcpy(nodes, sc->scene_nodes, sizeof(StBVHNodeRec) * sc->scene_node_count);
    memcpy(leaf_ids, sc->scene_leaf_ids, 4u * sc->prim_count);
    *node_count = sc->scene_node_count;
    *height = sc->scene_height;
    return ST_OK;
}

/* ---- moving primitives ------------------------------------------------- */

extern "C" int st_scene_set_prims(StSceneHandle *h, const StPrimRec *recs, uint32_t first, uint32_t count) {
    RT_LOCK;
    if (!h || !recs) return set_err(ST_ERR_ARG, "null argument");
    if (!h->host || !h->dbuf) return set_err(ST_ERR_ARG, "scene is not resident yet");
    StScene *sc = h->host;
    if ((uint64_t)first + count > sc->prim_count) return set_err(ST_ERR_ARG, "primitive range out of bounds");
    for (uint32_t i = 0; i < count; i++) {
        const StPrimRec *r = recs + i;
        if (r->material_id >= sc->material_count || r->material_id > 255)
            return set_err(ST_ERR_ARG, "primitive %u: material_id out of range", first + i);
        if (r->type == ST_PRIM_MESH && r->mesh_id >= sc->mesh_count)
            return set_err(ST_ERR_ARG, "primitive %u: mesh_id out of range", first + i);
    }
    for (uint32_t i = 0; i < count; i++) sc->prims[first + i] = st_prim_from_rec(recs + i);
    sc->has_emissive_quads = 0;
    for (uint32_t i = 0; i < sc->prim_count; i++) {
        uint32_t mid = (sc->prims[i].flags_mat >> 8) & 0xFFu;
        if (sc->prims[i].type == ST_PRIM_QUAD && (sc->mats[mid].bits & ST_EMISSIVE)) sc->has_emissive_quads = 1;
    }
    h->dev.has_emissive_quads = sc->has_emissive_quads;
    HIP_TRY(hipMemcpyAsync((void *)(h->dev.prims + first), sc->prims + first, count * sizeof(StDPrim),
                           hipMemcpyHostToDevice, g.main.stream));
    HIP_TRY(hipStreamSynchronize(g.main.stream));
    return ST_OK;
}

/* Rebuild the scene BVH over the current primitive transforms: the
   reference's SAH builder on the host, or the Morton LBVH on the device. */
extern "C" int st_scene_rebuild(StSceneHandle *h, uint32_t mode, float *ms) {
    RT_LOCK;
    if (!h) return set_err(ST_ERR_ARG, "null scene");
    int rc = ensure_init();
    if (rc) return rc;
    if (!h->host || !h->dbuf) return set_err(ST_ERR_ARG, "scene is not resident yet");
    StScene *sc = h->host;
    uint32_t P = sc->prim_count;
    if (!P) return ST_OK;
    if (mode == ST_REBUILD_SAH) {
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        if (st_scene_rebuild_sah(sc)) return set_err(ST_ERR_MEMORY, "SAH rebuild failed");
        int height = st_bvh_validate(sc->scene_nodes, sc->scene_node_count, P);
        if (height < 0 || height + 1 > ST_SCENE_STACK_MAX) return set_err(ST_ERR_SCENE, "rebuilt BVH too deep");
        sc->scene_real_height = (uint32_t)height;
        HIP_TRY(hipMemcpy((void *)h->dev.scene_nodes, sc->scene_nodes, sizeof(StBVHNodeRec) * sc->scene_node_count,
                          hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy((void *)h->dev.scene_leaf_ids, sc->scene_leaf_ids, 4u * P, hipMemcpyHostToDevice));
        h->dev.scene_stack_need = sc->scene_real_height + 1;
        h->bvh_on_device_only = 0;
        clock_gettime(CLOCK_MONOTONIC, &t1);
        if (ms) *ms = (float)((t1.tv_sec - t0.tv_sec) * 1e3 + (t1.tv_nsec - t0.tv_nsec) * 1e-6);
        return ST_OK;
    }
    if (mode != ST_REBUILD_LBVH) return set_err(ST_ERR_ARG, "unknown rebuild mode %u", mode);
    const size_t wb = st_lbvh_work_bytes(P);
    if (h->lbvh_work_bytes < wb) {
        if (h->lbvh_work) HIP_TRY(hipFree(h->lbvh_work));
        h->lbvh_work = NULL;
        h->lbvh_work_bytes = 0;
        HIP_TRY(hipMalloc(&h->lbvh_work, wb));
        h->lbvh_work_bytes = wb;
    }
    if (!h->d_height) HIP_TRY(hipMalloc(&h->d_height, 256));
    HIP_TRY(hipEventRecord(g.main.ev[2], g.main.stream));
    if (st_lbvh_launch(g.main.stream, h->dev.prims, h->d_mesh_aabbs, sc->mesh_count, P, h->lbvh_work,
                       (uint32_t *)h->dev.scene_nodes,
                       (uint32_t *)h->dev.scene_leaf_ids, h->d_height))
        return set_err(ST_ERR_HIP, "LBVH launch failed: %s", hipGetErrorString(hipGetLastError()));
    HIP_TRY(hipEventRecord(g.main.ev[3], g.main.stream));
    uint32_t height = 0;
    HIP_TRY(hipMemcpyAsync(&height, h->d_height, 4, hipMemcpyDeviceToHost, g.main.stream));
    HIP_TRY(hipStreamSynchronize(g.main.stream));
    if (ms) HIP_TRY(hipEventElapsedTime(ms, g.main.ev[2], g.main.ev[3]));
    if (height == 0 || height + 1 > ST_SCENE_STACK_MAX) return set_err(ST_ERR_SCENE, "LBVH height %u too deep", height);
    sc->scene_node_count = 2 * P - 1;
    sc->scene_height = height;
    sc->scene_real_height = height;
    h->dev.scene_stack_need = height + 1;
    h->bvh_on_device_only = 1;
    return ST_OK;
}

/* ---- launches ----------------------------------------------------------- */

/* blocks of BLOCK_THREADS that are resident on the whole device at once */
static uint32_t resident_blocks(const void *kernel, size_t lds) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, BLOCK_THREADS, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return (uint32_t)per_cu * (uint32_t)g.cus;
}

/* LDS per CU on gfx950 */
#define CU_LDS_BYTES (160u * 1024u)
/* the redo stage's grid: as many blocks as are resident at once */
static uint32_t redo_blocks_for(uint32_t threads, size_t lds) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k_redo), (int)threads,
                                                     lds) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return (uint32_t)per_cu * (uint32_t)g.cus;
}

/* resident_blocks with at most `cap` blocks per CU (0: no cap): the block's
   LDS request is raised until no more than `cap` fit; *lds_launch is the
   request to launch with */
static uint32_t capped_blocks(const void *kernel, size_t lds, uint32_t cap, size_t *lds_launch) {
    size_t want = lds;
    if (cap) {
        /* just over a (cap+1)-th of the CU's LDS: cap+1 blocks do not fit,
           and cap blocks fit with room for the allocation granularity */
        const size_t pad = (((size_t)CU_LDS_BYTES / (cap + 1)) & ~(size_t)1023) + 2048;
        if (pad > want) want = pad;
    }
    *lds_launch = want;
    return resident_blocks(kernel, want);
}

struct LaunchPlan {
    uint32_t nreq, total_rows, total_tiles, nscenes;
    size_t work_bytes, stage_bytes;
};

/* 64-bit FNV-1a over words: the shape signature that keys a context's tile
   history (a tile time only orders the same tile of a launch of the same
   shape). */
static uint64_t sig_hash(const std::vector<uint64_t> &w) {
    uint64_t h = 1469598103934665603ull;
    for (uint64_t x : w)
        for (int k = 0; k < 8; k++) {
            h ^= (x >> (8 * k)) & 0xFFu;
            h *= 1099511628211ull;
        }
    return h ? h : 1;
}

/* Upload and run one launch on context c over `n` requests whose host views
   are already prepared (camera basis, row starts, SSB rects). Each
   request's frame (and hit records) is copied to pixels_host[i] /
   hits_host[i] when given. Returns when the launch is done. Runs without
   the library lock on a server context: only c's own buffers are written,
   the scenes are read. */
int st_rt_launch_ctx(LaunchCtx *c, StSceneHandle *const *scenes, StHostView *hv, uint32_t n, int want_hits,
                     int persistent, StRayCounts *counts, StFloatPixel *const *pixels_host,
                     StHitRec *const *hits_host, uint32_t *const *bgra_host, float *kernel_ms) {
    int rc = ensure_init();
    if (rc) return rc;
    if (n == 0) return set_err(ST_ERR_ARG, "no requests");
    StTuning tune;
    uint32_t step_cap, abandon_steps, coop_min_tris;
    {
        RT_LOCK;
        tune = g.tune;
        step_cap = g.step_cap;
        abandon_steps = g.abandon_steps;
        coop_min_tris = g.coop_min_tris;
    }
    std::vector<uint32_t> scene_idx(n);
    std::vector<StSceneHandle *> uniq(n);
    uint32_t nuniq = 0;
    /* one upload block (requests, scene table, row starts, rects), staged in
       pinned memory and copied with a single transfer */
    std::vector<size_t> rs_off(n), rect_off(n);
    size_t up = round256(n * sizeof(KReq));
    size_t per_req_out = 0;
    uint32_t rows = 0, tiles = 0;
    uint64_t out_px = 0;
    std::vector<uint64_t> shape = {n, tune.tile_shift};
    for (uint32_t i = 0; i < n; i++) {
        StSceneHandle *s = scenes[i];
        if (!s || !s->host || !s->dbuf) return set_err(ST_ERR_ARG, "request %u has no resident scene", i);
        uint32_t k = 0;
        while (k < nuniq && uniq[k] != s) k++;
        if (k == nuniq) uniq[nuniq++] = s;
        scene_idx[i] = k;
        size_t px = (size_t)hv[i].dv.width1 * hv[i].dv.height1;
        per_req_out += round256((size_t)hv[i].dv.height1 * ST_DIR_MARKS(hv[i].dv.width1) * 16) + round256(px * 24) +
                       (want_hits ? round256(px * 16) : 0);
        if (bgra_host && bgra_host[i]) {
            if (hv[i].dv.width1 < 2 || hv[i].dv.height1 < 2) return set_err(ST_ERR_ARG, "request %u: empty window", i);
            per_req_out += round256((size_t)(hv[i].dv.width1 - 1) * (hv[i].dv.height1 - 1) * 4);
        }
        const uint32_t tw = 1u << tune.tile_shift, th = 64u >> tune.tile_shift;
        hv[i].dv.tile_shift = tune.tile_shift;
        hv[i].dv.tiles_x = (hv[i].dv.width1 + tw - 1) / tw;
        uint32_t ty = (hv[i].dv.height1 + th - 1) / th;
        hv[i].dv.tile_base = tiles;
        tiles += hv[i].dv.tiles_x * ty;
        rows += hv[i].dv.height1;
        out_px += (uint64_t)(hv[i].dv.width1 - 1) * (hv[i].dv.height1 - 1);
        shape.insert(shape.end(), {(uint64_t)(uintptr_t)s, hv[i].dv.width1, hv[i].dv.height1, hv[i].dv.mode,
                                   hv[i].dv.depth});
    }
    if (out_px > 0xFFFFFFF0ull || tiles > 0x7FFFFFF0u || (uint64_t)rows * 3 > 0xFFFFFFF0ull)
        return set_err(ST_ERR_ARG, "batch too large");
    const uint64_t tile_sig = sig_hash(shape);
    size_t o_scenes = up;
    up += round256(nuniq * sizeof(StDScene));
    for (uint32_t i = 0; i < n; i++) {
        rs_off[i] = up;
        up += round256(hv[i].dv.height1 * 16ull);
        rect_off[i] = up;
        up += round256((scenes[i]->host->prim_count + 1) * 16ull);
    }
    /* wavefront buffers of the Beauty depth loop: levels = largest depth,
       shadow planes = largest light count among Beauty requests */
    uint32_t levels = 0, lmax = 1;
    for (uint32_t i = 0; i < n; i++)
        if (hv[i].dv.mode == ST_MODE_BEAUTY && hv[i].dv.depth) {
            if (hv[i].dv.depth > levels) levels = hv[i].dv.depth;
            if (scenes[i]->host->light_count > lmax) lmax = scenes[i]->host->light_count;
        }
    const size_t slots = (size_t)tiles * 64;
    size_t wave_bytes = 0;
    /* the bounce-ray sort: its bins and the binned queue */
    const bool ray_sort = levels > 1 && tune.ray_sort;
    const uint32_t nbins = 2u * n * 8u;
    if (levels) {
        if (slots > 0xFFFFFFC0ull || slots * lmax > 0xFFFFFFC0ull) return set_err(ST_ERR_ARG, "batch too large");
        wave_bytes = round256(slots * 16 * ST_PLANES) + round256(slots * lmax * 8) + round256(slots * 4) +
                     round256(slots * 4) + round256(8) + round256(slots * 8) + 6 * 256;
        if (ray_sort) wave_bytes += round256(slots * 4) + 2 * round256(nbins * 4ull) + 512;
    }
    /* per request: primary, shadow, bounce totals; then the fault word and
       the redo stage's queue length */
    const size_t cnt_bytes = n * 24ull + 16;
    /* the redo stage's queue: every launch (a primary ray can give up too) */
    uint32_t s_entries = 1, m_entries = 1;
    for (uint32_t i = 0; i < n; i++) {
        const StDScene &d = scenes[i]->dev;
        if (d.scene_stack_need > s_entries) s_entries = d.scene_stack_need;
        if (d.mesh_stack_size > m_entries) m_entries = d.mesh_stack_size;
    }
    /* a persistent launch keeps each traversal stack's newest entries in an
       LDS window of the primary and bounce stages (StTuning.stack_entries)
       and their older entries in a spill buffer with a region per wave of a
       stage kernel (persistent grids and resident-only grids bound the wave
       ids); one wavefront per tile keeps whole stacks */
    const uint32_t win = persistent ? tune.stack_entries : 0u;
    uint32_t ks, km, mask_s, mask_m;
    stack_window(s_entries, win, ks, mask_s);
    stack_window(m_entries, win, km, mask_m);
    if (persistent && (size_t)BLOCK_THREADS * (ks + km) * sizeof(uint32_t) > 160 * 1024) {
        stack_window(s_entries, 32u, ks, mask_s);
        stack_window(m_entries, 32u, km, mask_m);
    }
    const uint32_t spill_s = mask_s != ~0u ? s_entries : 0u, spill_m = mask_m != ~0u ? m_entries : 0u;
    uint32_t spill_waves = 0;
    if (spill_s + spill_m) {
        const uint32_t pk = std::min<uint32_t>((tiles + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK,
                                               (uint32_t)g.cus * tune.blocks_per_cu);
        spill_waves = std::max<uint32_t>(pk * WAVES_PER_BLOCK, (uint32_t)g.cus * 32u);
    }
    const size_t spill_bytes = (size_t)spill_waves * 64u * (spill_s + spill_m) * 4u;
    /* redo stage: as many blocks as fit at once, grouped per redone pixel */
    uint32_t redo_threads;
    size_t lds_redo;
    redo_shape(s_entries, m_entries, redo_threads, lds_redo);
    const uint32_t redo_blocks = redo_blocks_for(redo_threads, lds_redo);
    const size_t redo_bytes = round256(slots * 4) + 256 + redo_coop_bytes(redo_blocks, REDO_QUEUE) + 256;
    size_t work = up + round256(cnt_bytes) + per_req_out + wave_bytes + redo_bytes + round256(spill_bytes) + 512;
    if ((rc = arena_reserve(&c->work, work))) return rc;
    /* staging: the upload block, then the ray counts and fault word read back */
    const size_t o_readback = round256(up);
    if ((rc = staging_reserve(c, o_readback + cnt_bytes))) return rc;
    if (tiles > c->tile_cap) {   /* per-tile times and claim order, kept across launches */
        if (c->tile_cost) HIP_TRY(hipFree(c->tile_cost));
        c->tile_cost = c->tile_order = NULL;
        c->tile_cap = 0;
        c->tile_sig = 0;
        HIP_TRY(hipMalloc(&c->tile_cost, (size_t)tiles * 8));
        HIP_TRY(hipMemset(c->tile_cost, 0, (size_t)tiles * 8));
        c->tile_order = c->tile_cost + tiles;
        c->tile_cap = tiles;
    }
    /* the previous launch on this context had this shape: order the tiles
       by its tile times (when the tuning asks for it) */
    const uint32_t use_history = (c->tile_sig == tile_sig && g.tune.cost_order == 1u) ? 1u : 0u;
    const uint32_t order_mode = use_history ? ORDER_HISTORY : g.tune.cost_order == 2u ? ORDER_SUPER : ORDER_TILE;
    char *st = c->staging;
    char *blk = (char *)arena_take(&c->work, up);
    KReq *reqs_h = (KReq *)st;
    KReq *reqs_d = (KReq *)blk;
    StDScene *sc_h = (StDScene *)(st + o_scenes);
    StDScene *sc_d = (StDScene *)(blk + o_scenes);
    /* per request: primary, shadow, bounce ray totals; then the launch's
       traversal fault word (StDScene.fault) */
    unsigned long long *cnt_d = (unsigned long long *)arena_take(&c->work, cnt_bytes);
    unsigned long long *cnt_h = (unsigned long long *)(st + o_readback);
    for (uint32_t k = 0; k < nuniq; k++) {
        sc_h[k] = uniq[k]->dev;
        sc_h[k].fault = (uint32_t *)(cnt_d + 3 * (size_t)n);
        launch_limits(sc_h[k], step_cap, abandon_steps, coop_min_tris, mask_s, mask_m);
        sc_h[k].spill_s = spill_s;
        sc_h[k].spill_m = spill_m;
    }
    uint32_t row_base = 0, out_base = 0, count_blocks = 0, any_device_ssb = 0;
    for (uint32_t i = 0; i < n; i++) {
        size_t px = (size_t)hv[i].dv.width1 * hv[i].dv.height1;
        StDView &dv = hv[i].dv;
        dv.scene = scene_idx[i];
        uint32_t P = scenes[i]->host->prim_count;
        memcpy(st + rs_off[i], hv[i].row_starts, dv.height1 * 16ull);
        if (!hv[i].device_ssb) memcpy(st + rect_off[i], hv[i].rects, P * 16ull);
        dv.row_starts = (const float *)(blk + rs_off[i]);
        dv.rects = (const int32_t *)(blk + rect_off[i]);
        dv.dir_marks = (float *)arena_take(&c->work, (size_t)dv.height1 * ST_DIR_MARKS(dv.width1) * 16);
        dv.pixels = arena_take(&c->work, px * 24);
        dv.hits = want_hits ? arena_take(&c->work, px * 16) : NULL;
        dv.counters = cnt_d + 3 * i;
        memset(&reqs_h[i], 0, sizeof(KReq));
        reqs_h[i].v = dv;
        reqs_h[i].row_base = row_base;
        reqs_h[i].out_base = out_base;
        reqs_h[i].tile_count = dv.tiles_x * ((dv.height1 + (64u >> dv.tile_shift) - 1) / (64u >> dv.tile_shift));
        reqs_h[i].count_block = count_blocks;
        count_blocks += (reqs_h[i].tile_count + COUNT_TILES - 1) / COUNT_TILES;
        uint32_t wpx = (dv.width1 - 1) * (dv.height1 - 1);
        reqs_h[i].bgra = (bgra_host && bgra_host[i]) ? (uint32_t *)arena_take(&c->work, (size_t)wpx * 4) : NULL;
        if (hv[i].device_ssb) {
            const sm3 &m = hv[i].rotation_matrix_inverted;
            const float rmi[9] = {m.X.x, m.X.y, m.X.z, m.Y.x, m.Y.y, m.Y.z, m.Z.x, m.Z.y, m.Z.z};
            memcpy(reqs_h[i].ssb_rmi, rmi, sizeof rmi);
            reqs_h[i].focal_length = hv[i].focal_length;
            reqs_h[i].device_ssb = 1;
            any_device_ssb = 1;
        }
        row_base += dv.height1;
        out_base += wpx;
    }
    KWave w;
    memset(&w, 0, sizeof w);
    if (levels) {
        w.n = (uint32_t)slots;
        w.lmax = lmax;
        w.st = (float *)arena_take(&c->work, slots * 16 * ST_PLANES);
        w.shadow_res = (uint32_t *)arena_take(&c->work, slots * lmax * 8);
        w.occ_d2 = (float *)arena_take(&c->work, slots * 4);
        w.bounce_q = (uint32_t *)arena_take(&c->work, slots * 4);
        w.ctr = (uint32_t *)arena_take(&c->work, 8);
        w.counts = (uint32_t *)arena_take(&c->work, slots * 8);
    }
    uint32_t *bins = NULL, *bin_sums = NULL;
    if (ray_sort) {
        w.sorted_q = (uint32_t *)arena_take(&c->work, slots * 4);
        bins = (uint32_t *)arena_take(&c->work, nbins * 4ull);
        bin_sums = (uint32_t *)arena_take(&c->work, nbins * 4ull);
    }
    w.redo_q = (uint32_t *)arena_take(&c->work, slots * 4);
    w.redo_ctr = (uint32_t *)(cnt_d + 3 * (size_t)n + 1);
    KCoop coop;
    coop.base = (char *)arena_take(&c->work, redo_coop_bytes(redo_blocks, REDO_QUEUE));
    coop.slots = (DCoopSlot *)(coop.base + redo_slots_offset(redo_blocks, REDO_QUEUE));
    coop.group_stride = (uint32_t)redo_group_stride(REDO_QUEUE);
    coop.q_cap = REDO_QUEUE;
    coop.epoch = (const uint32_t *)(c->cost_sum + 1);
    uint32_t *spill = spill_bytes ? (uint32_t *)arena_take(&c->work, spill_bytes) : NULL;
    for (uint32_t k = 0; k < nuniq; k++) sc_h[k].spill = spill;
    /* launch shape, fixed before anything is enqueued so that a captured
       graph of this launch can be matched against the next one */
    uint32_t blocks;
    uint32_t xcd_blocks = 0;
    uint32_t claim_mode = 0;
    if (persistent) {
        uint32_t want = (uint32_t)g.cus * tune.blocks_per_cu;
        uint32_t need = (tiles + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
        blocks = need < want ? need : want;
        /* the primary stage's tile claim: the global counter (1), or XCD
           chunks first (2, when tuned so and the launch has at least
           XCD_MIN_TILES tiles per wave: with a few tiles per wave the
           chunks' claims, all made at once, only add contention) */
        claim_mode = (tune.persistent == 2u && tiles >= XCD_MIN_TILES * blocks * WAVES_PER_BLOCK) ? 2u : 1u;
    } else {
        blocks = (tiles + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
        blocks = (blocks + 7u) & ~7u;
        xcd_blocks = blocks;
    }
    size_t lds = (size_t)BLOCK_THREADS * (ks + km) * sizeof(uint32_t);
    if (lds > 160 * 1024) return set_err(ST_ERR_SCENE, "traversal stacks need %zu B of LDS", lds);
    /* the primary stage at 4 waves per SIMD when asked for and the stacks
       leave LDS for 4 blocks of 4 waves per CU */
    const bool waves4 = tune.primary_waves == 4 && 4 * lds <= 160 * 1024;
    /* the depth loop's queue kernels run exactly as many waves as fit on
       the device at once: they take queue chunks in a fixed interleaved
       order, so a wave that had to wait for a slot would run its whole
       share late */
    uint32_t b_bounce = 0;
    size_t lds_primary = lds;
    /* one wave per tile: at most blocks_per_cu of its blocks per CU (fewer
       tiles in flight share the caches better) */
    if (!persistent && tune.blocks_per_cu < 4)
        (void)capped_blocks(waves4 ? reinterpret_cast<const void *>(k_primary4) : reinterpret_cast<const void *>(k_primary),
                            lds, tune.blocks_per_cu, &lds_primary);
    if (levels) {
        b_bounce = resident_blocks(reinterpret_cast<const void *>(k_bounce), lds);
    }
    const uint32_t cost_blocks = std::min<uint32_t>((tiles + COST_THREADS - 1) / COST_THREADS, 1024u);

    /* per-stage timing (bench): an event before the first stage and after each */
    uint32_t marks = 0;
    const bool graph = c->graph_mode != 0;
    auto mark = [&]() -> hipError_t {
        if (graph || !c->stage_timing || marks > ST_STAGES_MAX) return hipSuccess;
        if (!c->stage_ev[marks]) {
            hipError_t e = hipEventCreate(&c->stage_ev[marks]);
            if (e != hipSuccess) return e;
        }
        return hipEventRecord(c->stage_ev[marks++], c->stream);
    };
    /* everything from the upload to the ray counts; the same stream work is
       what a graph captures. ev[0] .. ev[1] brackets every kernel of the
       launch (the render time st_render_timed reports). */
    auto enqueue = [&]() -> int {
        HIP_TRY(hipMemcpyAsync(blk, st, up, hipMemcpyHostToDevice, c->stream));
        if (!graph) HIP_TRY(hipEventRecord(c->ev[0], c->stream));
        hipLaunchKernelGGL(k_zero_u64, dim3((3 * n + 2 + 255) / 256), dim3(256), 0, c->stream, cnt_d, 3 * n + 2);
        hipLaunchKernelGGL(k_zero_u32, dim3(CLAIM_WORDS / 256u), dim3(256), 0, c->stream, c->tile_counter,
                           CLAIM_WORDS);
        static_assert(BOUNCE_CLAIM < ORDER_CLAIM && ORDER_CLAIM + 1 < XCD_CLAIM &&
                          XCD_CLAIM + 8 * XCD_CLAIM_STRIDE <= CLAIM_WORDS && CLAIM_WORDS % 256u == 0,
                      "claim counters share the zeroed words");
        if (any_device_ssb)
            hipLaunchKernelGGL(k_ssb, dim3(n), dim3(256), 0, c->stream, (const StDScene *)sc_d, (const KReq *)reqs_d);
        if (levels)
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c->stream, w.ctr, 2u);
        hipLaunchKernelGGL(k_primary_dirs, dim3((uint32_t)(((uint64_t)rows * 3 + DIRS_BLOCK - 1) / DIRS_BLOCK)),
                           dim3(DIRS_BLOCK), 0, c->stream, reqs_d, n, rows);
        HIP_TRY(hipGetLastError());
        if (use_history) {
            hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, c->cost_sum, 1u);
            hipLaunchKernelGGL(k_cost_total, dim3(cost_blocks), dim3(COST_THREADS), 0, c->stream,
                               (const uint32_t *)c->tile_cost, tiles, c->cost_sum);
        }
        hipLaunchKernelGGL(k_tile_order, dim3((tiles + ORDER_THREADS - 1) / ORDER_THREADS), dim3(ORDER_THREADS), 0,
                           c->stream, (const uint32_t *)c->tile_cost, c->tile_order, tiles,
                           (const unsigned long long *)c->cost_sum, c->tile_counter + ORDER_CLAIM, order_mode,
                           (uint32_t *)(c->cost_sum + 1), (const KReq *)reqs_d, n);
        HIP_TRY(hipGetLastError());
        HIP_TRY(mark());
        if (waves4)
            hipLaunchKernelGGL(k_primary4, dim3(blocks), dim3(BLOCK_THREADS), lds_primary, c->stream,
                               (const StDScene *)sc_d, (const KReq *)reqs_d, n, tiles, c->tile_counter,
                               claim_mode, xcd_blocks, ks, km, w, (const uint32_t *)c->tile_order,
                               c->tile_cost);
        else
            hipLaunchKernelGGL(k_primary, dim3(blocks), dim3(BLOCK_THREADS), lds_primary, c->stream,
                               (const StDScene *)sc_d, (const KReq *)reqs_d, n, tiles, c->tile_counter,
                               claim_mode, xcd_blocks, ks, km, w, (const uint32_t *)c->tile_order,
                               c->tile_cost);
        HIP_TRY(hipGetLastError());
        HIP_TRY(mark());
        /* the rest of the depth loop: the bounce stage reads its queue
           length on the device, so it needs no host round trip */
        if (levels > 1) {
            if (ray_sort) {
                const uint32_t bb = std::min<uint32_t>((uint32_t)((slots + BIN_THREADS - 1) / BIN_THREADS),
                                                       4u * (uint32_t)g.cus);
                hipLaunchKernelGGL(k_zero_u32, dim3((nbins + 255) / 256), dim3(256), 0, c->stream, bins, nbins);
                hipLaunchKernelGGL(k_bin, dim3(bb), dim3(BIN_THREADS), 0, c->stream, (const KReq *)reqs_d, n, w, bins,
                                   0u);
                if (st_scan_exclusive(c->stream, bins, nbins, bin_sums))
                    return set_err(ST_ERR_HIP, "bin scan failed: %s", hipGetErrorString(hipGetLastError()));
                hipLaunchKernelGGL(k_bin, dim3(bb), dim3(BIN_THREADS), 0, c->stream, (const KReq *)reqs_d, n, w, bins,
                                   1u);
                HIP_TRY(hipGetLastError());
            }
            hipLaunchKernelGGL(k_bounce, dim3(b_bounce), dim3(BLOCK_THREADS), lds, c->stream,
                               (const StDScene *)sc_d, (const KReq *)reqs_d, n, w, c->tile_counter + BOUNCE_CLAIM,
                               ks, km, tune.bounce_order);
            HIP_TRY(mark());
        }
        /* the pixels that gave up, traced again (it reads the queue's length
           on the device: no round trip; nothing to do is one quick launch) */
        hipLaunchKernelGGL(k_redo, dim3(redo_blocks), dim3(redo_threads), lds_redo, c->stream, (const StDScene *)sc_d,
                           (const KReq *)reqs_d, n, w, s_entries, m_entries, coop);
        HIP_TRY(mark());
        if (levels && count_blocks)
            hipLaunchKernelGGL(k_counts, dim3(count_blocks), dim3(256), 0, c->stream, (const KReq *)reqs_d, n,
                               (const uint32_t *)w.counts);
        HIP_TRY(hipGetLastError());
        if (!graph) HIP_TRY(hipEventRecord(c->ev[1], c->stream));
        return ST_OK;
    };

    if (!graph) {
        if ((rc = enqueue())) return rc;
    } else {
        /* everything a kernel argument or copy depends on: the buffers are
           at fixed arena offsets, so an equal signature means the captured
           graph does exactly this launch's work (the per-frame data is in
           the upload block, copied from the same pinned staging) */
        std::vector<uint64_t> sig = {n, up, rows, tiles, levels, lmax, count_blocks, blocks, xcd_blocks, any_device_ssb,
                                     (uint64_t)claim_mode, s_entries, m_entries, ks, km, (uint64_t)waves4, b_bounce, tune.bounce_order,
                                     lds_primary, (uint64_t)(uintptr_t)blk,
                                     (uint64_t)(uintptr_t)st,
                                     (uint64_t)(uintptr_t)cnt_d, (uint64_t)(uintptr_t)w.st, w.n,
                                     (uint64_t)(uintptr_t)c->tile_cost, order_mode, cost_blocks,
                                     (uint64_t)(uintptr_t)w.redo_q, redo_blocks, redo_threads, lds_redo,
                                     (uint64_t)(uintptr_t)coop.base, (uint64_t)ray_sort,
                                     (uint64_t)(uintptr_t)w.sorted_q, (uint64_t)(uintptr_t)spill, spill_s, spill_m};
        for (uint32_t i = 0; i < n; i++) {
            const StDView &dv = reqs_h[i].v;
            sig.insert(sig.end(), {dv.width1, dv.height1, dv.mode, dv.depth, dv.tile_shift,
                                   (uint64_t)(uintptr_t)dv.dir_marks, (uint64_t)(uintptr_t)dv.pixels,
                                   (uint64_t)(uintptr_t)dv.hits, (uint64_t)(uintptr_t)reqs_h[i].bgra});
        }
        if (!c->graph_exec || sig != c->graph_sig) {
            if (c->graph_exec) {
                (void)hipGraphExecDestroy(c->graph_exec);
                c->graph_exec = nullptr;
            }
            HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
            const int erc = enqueue();
            hipGraph_t gr = nullptr;
            const hipError_t ee = hipStreamEndCapture(c->stream, &gr);
            if (erc || ee != hipSuccess) {
                if (gr) (void)hipGraphDestroy(gr);
                return erc ? erc : set_err(ST_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ee));
            }
            const hipError_t ie = hipGraphInstantiate(&c->graph_exec, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (ie != hipSuccess) {
                c->graph_exec = nullptr;
                return set_err(ST_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ie));
            }
            c->graph_sig = sig;
            g.graph_captures++;
        }
        HIP_TRY(hipEventRecord(c->ev[0], c->stream));
        HIP_TRY(hipGraphLaunch(c->graph_exec, c->stream));
        HIP_TRY(hipEventRecord(c->ev[1], c->stream));
        g.graph_launches++;
    }
    /* this launch's tile times are now what tile_cost will hold */
    c->tile_sig = tile_sig;
    if (bgra_host && out_px) {
        hipLaunchKernelGGL(k_resolve, dim3((uint32_t)((out_px + 255) / 256)), dim3(256), 0, c->stream,
                           (const KReq *)reqs_d, n, (uint32_t)out_px);
        HIP_TRY(hipGetLastError());
    }
    for (uint32_t i = 0; i < n; i++) {
        size_t px = (size_t)hv[i].dv.width1 * hv[i].dv.height1;
        if (pixels_host && pixels_host[i])
            HIP_TRY(hipMemcpyAsync(pixels_host[i], reqs_h[i].v.pixels, px * 24, hipMemcpyDeviceToHost, c->stream));
        if (hits_host && hits_host[i] && want_hits)
            HIP_TRY(hipMemcpyAsync(hits_host[i], reqs_h[i].v.hits, px * 16, hipMemcpyDeviceToHost, c->stream));
        if (reqs_h[i].bgra)
            HIP_TRY(hipMemcpyAsync(bgra_host[i], reqs_h[i].bgra,
                                   (size_t)(hv[i].dv.width1 - 1) * (hv[i].dv.height1 - 1) * 4, hipMemcpyDeviceToHost,
                                   c->stream));
    }
    HIP_TRY(hipMemcpyAsync(cnt_h, cnt_d, cnt_bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->last_redo = (uint32_t)cnt_h[3 * (size_t)n + 1];
    if (cnt_h[3 * (size_t)n])
        return set_err(ST_ERR_SCENE, "a traversal ran past its bound (fault bits 0x%llx): malformed BVH",
                       cnt_h[3 * (size_t)n]);
    if (counts)
        for (uint32_t i = 0; i < n; i++) {
            counts[i].primary = (uint64_t)hv[i].dv.width1 * hv[i].dv.height1;
            counts[i].shadow = cnt_h[3 * i + 1];
            counts[i].bounce = cnt_h[3 * i + 2];
        }
    if (kernel_ms) HIP_TRY(hipEventElapsedTime(kernel_ms, c->ev[0], c->ev[1]));
    c->stage_n = marks ? marks - 1 : 0;
    for (uint32_t k = 0; k < c->stage_n; k++)
        HIP_TRY(hipEventElapsedTime(&c->stage_ms[k], c->stage_ev[k], c->stage_ev[k + 1]));
    return rc;
}

int st_rt_launch_prepared(StSceneHandle *const *scenes, StHostView *hv, uint32_t n, int want_hits, int persistent,
                          StRayCounts *counts, StFloatPixel *const *pixels_host, StHitRec *const *hits_host,
                          uint32_t *const *bgra_host, float *kernel_ms) {
    RT_LOCK;
    return st_rt_launch_ctx(&g.main, scenes, hv, n, want_hits, persistent, counts, pixels_host, hits_host, bgra_host,
                            kernel_ms);
}

static int launch(StSceneHandle *const *scenes, const StView *views, uint32_t n, int want_hits, int persistent,
                  StRayCounts *counts, StFloatPixel *const *pixels_host, StHitRec *const *hits_host,
                  uint32_t *const *bgra_host, float *kernel_ms) {
    int rc = ensure_init();
    if (rc) return rc;
    if (n == 0) return set_err(ST_ERR_ARG, "no requests");
    StHostView *hv = (StHostView *)calloc(n, sizeof(StHostView));
    if (!hv) return set_err(ST_ERR_MEMORY, "out of host memory");
    for (uint32_t i = 0; i < n && !rc; i++) {
        if (!scenes[i] || !scenes[i]->host) rc = set_err(ST_ERR_ARG, "request %u has no resident scene", i);
        else if (st_view_prepare_ex(scenes[i]->host, views + i, hv + i, 0))
            rc = set_err(ST_ERR_ARG, "request %u: invalid view (%ux%u, mode %u)", i, views[i].width, views[i].height,
                         views[i].render_mode);
    }
    if (!rc)
        rc = st_rt_launch_prepared(scenes, hv, n, want_hits, persistent, counts, pixels_host, hits_host, bgra_host,
                                   kernel_ms);
    for (uint32_t i = 0; i < n; i++) st_view_release(hv + i);
    free(hv);
    return rc;
}

extern "C" int st_render(StSceneHandle *scene, const StView *view, StFloatPixel *pixels, StHitRec *hits,
                         StRayCounts *counts) {
    if (!scene || !view) return set_err(ST_ERR_ARG, "null argument");
    int rc = ensure_init();
    if (rc) return rc;
    return launch(&scene, view, 1, hits != NULL, (int)g.tune.persistent, counts, &pixels, &hits, NULL, NULL);
}

/* st_render without allocating: every per-launch device buffer is carved
   from this device's scratch array, the host side uses static buffers, and
   the launch runs on the main stream. The same kernels, in the same order,
   as a single-request st_render (tiles in tile order: the array keeps no
   tile history). */
extern "C" int st_render_noalloc(StSceneHandle *scene, const StView *view, StFloatPixel *pixels, StHitRec *hits,
                                 StRayCounts *counts) {
    if (!scene || !view || !pixels) return set_err(ST_ERR_ARG, "null argument");
    RT_LOCK;
    if (!g.ready) return set_err(ST_ERR_ARG, "st_init has not run: the scratch array's address is taken there");
    char *const base = g_noalloc_scratch[g.device];
    if (!base) return set_err(ST_ERR_HIP, "no scratch array address for device %d", g.device);
    if (!scene->host || !scene->dbuf) return set_err(ST_ERR_ARG, "scene is not resident");
    const StScene *hs = scene->host;
    if (view->width == 0 || view->height == 0 || view->width >= 0xFFFF || view->height >= 0xFFFF ||
        view->render_mode > ST_MODE_UVS)
        return set_err(ST_ERR_ARG, "invalid view (%ux%u, mode %u)", view->width, view->height, view->render_mode);
    const StTuning tune = g.tune;
    const uint32_t width1 = view->width + 1u, height1 = view->height + 1u;
    const size_t px = (size_t)width1 * height1;
    const uint32_t tw = 1u << tune.tile_shift, th = 64u >> tune.tile_shift;
    const uint32_t tiles_x = (width1 + tw - 1) / tw, tiles = tiles_x * ((height1 + th - 1) / th);
    const uint32_t P = hs->prim_count;
    const bool beauty = view->render_mode == ST_MODE_BEAUTY && (uint8_t)view->depth;
    const uint32_t lmax = hs->light_count > 1 ? hs->light_count : 1;
    const size_t slots = (size_t)tiles * 64;
    /* layout: upload block (request, scene record, row starts, rects), ray
       counts + fault word, direction checkpoints, frame, hit records, tile
       claim order, tile times, counters, depth-loop state */
    const size_t o_scene = round256(sizeof(KReq)), o_rows = o_scene + round256(sizeof(StDScene));
    const size_t o_rects = o_rows + round256((size_t)height1 * 16), up_host = o_rects;
    size_t off = o_rects + round256(((size_t)P + 1) * 16);
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += round256(bytes);
        return o;
    };
    const size_t o_cnt = take(40), o_dirs = take((size_t)height1 * ST_DIR_MARKS(width1) * 16), o_pix = take(px * 24);
    const size_t o_hits = hits ? take(px * 16) : 0;
    const size_t o_order = take((size_t)tiles * 4), o_cost = take((size_t)tiles * 4), o_ctr = take(CLAIM_WORDS * 4);
    const size_t o_redo = take(slots * 4);
    const uint32_t s_entries = scene->dev.scene_stack_need > 1 ? scene->dev.scene_stack_need : 1;
    const uint32_t m_entries = scene->dev.mesh_stack_size > 1 ? scene->dev.mesh_stack_size : 1;
    uint32_t redo_threads;
    size_t lds_redo;
    redo_shape(s_entries, m_entries, redo_threads, lds_redo);
    const uint32_t redo_blocks = redo_blocks_for(redo_threads, lds_redo);
    const size_t o_coop = take(redo_coop_bytes(redo_blocks, REDO_QUEUE_NOALLOC));
    /* whole traversal stacks in LDS (a spill buffer would not fit the array) */
    const uint32_t ks = s_entries, km = m_entries, mask_s = ~0u, mask_m = ~0u;
    size_t o_st = 0, o_sres = 0, o_occ = 0, o_bq = 0, o_wctr = 0, o_wcnt = 0;
    if (beauty) {
        o_st = take(slots * 16 * ST_PLANES);
        o_sres = take(slots * lmax * 8);
        o_occ = take(slots * 4);
        o_bq = take(slots * 4);
        o_wctr = take(8);
        o_wcnt = take(slots * 8);
    }
    /* the last 256 bytes hold the redo stage's launch number (kept across calls) */
    if (off > ST_NOALLOC_SCRATCH_BYTES - 256)
        return set_err(ST_ERR_MEMORY, "request needs %zu bytes of scratch; st_render_noalloc has %u", off + 256,
                       (unsigned)ST_NOALLOC_SCRATCH_BYTES);
    uint32_t *const epoch = (uint32_t *)(base + ST_NOALLOC_SCRATCH_BYTES - 256);
    if (up_host > ST_NOALLOC_HOST_BYTES) return set_err(ST_ERR_ARG, "view too tall for st_render_noalloc");
    StHostView hv;
    if (st_view_prepare_rows(hs, view, &hv, (float *)(g_noalloc_host + o_rows)))
        return set_err(ST_ERR_ARG, "invalid view");
    KReq *rq = (KReq *)g_noalloc_host;
    memset(rq, 0, sizeof(KReq));
    StDView &dv = rq->v;
    dv = hv.dv;
    dv.scene = 0;
    dv.tile_base = 0;
    dv.tiles_x = tiles_x;
    dv.tile_shift = tune.tile_shift;
    dv.row_starts = (const float *)(base + o_rows);
    dv.rects = (const int32_t *)(base + o_rects);
    dv.dir_marks = (float *)(base + o_dirs);
    dv.pixels = base + o_pix;
    dv.hits = hits ? base + o_hits : NULL;
    unsigned long long *cnt_d = (unsigned long long *)(base + o_cnt);
    dv.counters = cnt_d;
    rq->row_base = 0;
    rq->out_base = 0;
    rq->tile_count = tiles;
    rq->count_block = 0;
    const sm3 &m = hv.rotation_matrix_inverted;
    const float rmi[9] = {m.X.x, m.X.y, m.X.z, m.Y.x, m.Y.y, m.Y.z, m.Z.x, m.Z.y, m.Z.z};
    memcpy(rq->ssb_rmi, rmi, sizeof rmi);
    rq->focal_length = hv.focal_length;
    rq->device_ssb = 1;
    StDScene dsc = scene->dev;
    dsc.fault = (uint32_t *)(cnt_d + 3);
    launch_limits(dsc, g.step_cap, g.abandon_steps, g.coop_min_tris, mask_s, mask_m);
    dsc.spill = NULL;
    dsc.spill_s = dsc.spill_m = 0;
    memcpy(g_noalloc_host + o_scene, &dsc, sizeof dsc);
    KWave w;
    memset(&w, 0, sizeof w);
    if (beauty) {
        w.n = (uint32_t)slots;
        w.lmax = lmax;
        w.st = (float *)(base + o_st);
        w.shadow_res = (uint32_t *)(base + o_sres);
        w.occ_d2 = (float *)(base + o_occ);
        w.bounce_q = (uint32_t *)(base + o_bq);
        w.ctr = (uint32_t *)(base + o_wctr);
        w.counts = (uint32_t *)(base + o_wcnt);
    }
    w.redo_q = (uint32_t *)(base + o_redo);
    w.redo_ctr = (uint32_t *)(cnt_d + 4);
    KCoop coop;
    coop.base = base + o_coop;
    coop.slots = (DCoopSlot *)(coop.base + redo_slots_offset(redo_blocks, REDO_QUEUE_NOALLOC));
    coop.group_stride = (uint32_t)redo_group_stride(REDO_QUEUE_NOALLOC);
    coop.q_cap = REDO_QUEUE_NOALLOC;
    coop.epoch = epoch;
    const size_t lds = (size_t)BLOCK_THREADS * (ks + km) * sizeof(uint32_t);
    if (lds > 160 * 1024) return set_err(ST_ERR_SCENE, "traversal stacks need %zu B of LDS", lds);
    uint32_t blocks, xcd_blocks = 0;
    size_t lds_primary = lds;
    uint32_t claim_mode = 0;
    if (tune.persistent) {
        const uint32_t want = (uint32_t)g.cus * tune.blocks_per_cu, need = (tiles + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
        blocks = need < want ? need : want;
        claim_mode = (tune.persistent == 2u && tiles >= XCD_MIN_TILES * blocks * WAVES_PER_BLOCK) ? 2u : 1u;
    } else {
        blocks = ((tiles + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK + 7u) & ~7u;
        xcd_blocks = blocks;
        if (tune.blocks_per_cu < 4)
            (void)capped_blocks(reinterpret_cast<const void *>(k_primary), lds, tune.blocks_per_cu, &lds_primary);
    }
    const KReq *reqs_d = (const KReq *)base;
    const StDScene *sc_d = (const StDScene *)(base + o_scene);
    uint32_t *ctr = (uint32_t *)(base + o_ctr);
    hipStream_t stream = g.main.stream;
    /* the request, scene record and row starts; the rects region is filled
       by the device SSB pass */
    HIP_TRY(hipMemcpyAsync(base, g_noalloc_host, up_host, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, stream, cnt_d, 5u);
    hipLaunchKernelGGL(k_zero_u32, dim3(CLAIM_WORDS / 256u), dim3(256), 0, stream, ctr, CLAIM_WORDS);
    hipLaunchKernelGGL(k_ssb, dim3(1), dim3(256), 0, stream, sc_d, reqs_d);
    if (beauty) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, w.ctr, 2u);
    hipLaunchKernelGGL(k_primary_dirs, dim3((height1 * 3 + DIRS_BLOCK - 1) / DIRS_BLOCK), dim3(DIRS_BLOCK), 0, stream,
                       reqs_d, 1u, height1);
    hipLaunchKernelGGL(k_tile_order, dim3((tiles + ORDER_THREADS - 1) / ORDER_THREADS), dim3(ORDER_THREADS), 0, stream,
                       (const uint32_t *)(base + o_cost), (uint32_t *)(base + o_order), tiles,
                       (const unsigned long long *)cnt_d, ctr + ORDER_CLAIM,
                       tune.cost_order == 2u ? ORDER_SUPER : ORDER_TILE, epoch, reqs_d, 1u);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_primary, dim3(blocks), dim3(BLOCK_THREADS), lds_primary, stream, sc_d, reqs_d, 1u, tiles, ctr,
                       claim_mode, xcd_blocks, ks, km, w,
                       (const uint32_t *)(base + o_order), (uint32_t *)(base + o_cost));
    HIP_TRY(hipGetLastError());
    if (beauty && dv.depth > 1) {
        const uint32_t b_bounce = resident_blocks(reinterpret_cast<const void *>(k_bounce), lds);
        hipLaunchKernelGGL(k_bounce, dim3(b_bounce), dim3(BLOCK_THREADS), lds, stream, sc_d, reqs_d, 1u, w,
                           ctr + BOUNCE_CLAIM, ks, km, tune.bounce_order);
    }
    hipLaunchKernelGGL(k_redo, dim3(redo_blocks), dim3(redo_threads), lds_redo, stream, sc_d, reqs_d, 1u, w,
                       s_entries, m_entries, coop);
    if (beauty)
        hipLaunchKernelGGL(k_counts, dim3((tiles + COUNT_TILES - 1) / COUNT_TILES), dim3(256), 0, stream, reqs_d, 1u,
                           (const uint32_t *)w.counts);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(pixels, base + o_pix, px * 24, hipMemcpyDeviceToHost, stream));
    if (hits) HIP_TRY(hipMemcpyAsync(hits, base + o_hits, px * 16, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(g_noalloc_counts, cnt_d, 40, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (g_noalloc_counts[3])
        return set_err(ST_ERR_SCENE, "a traversal ran past its bound (fault bits 0x%llx): malformed BVH",
                       g_noalloc_counts[3]);
    if (counts) {
        counts->primary = px;
        counts->shadow = g_noalloc_counts[1];
        counts->bounce = g_noalloc_counts[2];
    }
    return ST_OK;
}

/* Graph mode for the duration of one API call (ST_BATCH_GRAPH). */
struct GraphScope {
    explicit GraphScope(uint32_t flags) { g.main.graph_mode = (flags & ST_BATCH_GRAPH) != 0; }
    ~GraphScope() { g.main.graph_mode = 0; }
};

extern "C" int st_debug_step_cap(uint32_t cap) {
    RT_LOCK;
    g.step_cap = cap;
    return ST_OK;
}

#ifdef CO_STATS
/* measurement builds: the cooperative traversal's records since the last call */
extern "C" int st_debug_coop_stats(unsigned long long *out, uint32_t cap, uint32_t *n) {
    RT_LOCK;
    HIP_TRY(hipDeviceSynchronize());
    unsigned cnt = 0;
    HIP_TRY(hipMemcpyFromSymbol(&cnt, HIP_SYMBOL(co_stats_n), sizeof cnt));
    if (cnt > 32768u) cnt = 32768u;
    if (cnt > cap) cnt = cap;
    if (cnt) HIP_TRY(hipMemcpyFromSymbol(out, HIP_SYMBOL(co_stats), (size_t)cnt * 32));
    const unsigned zero = 0;
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(co_stats_n), &zero, sizeof zero));
    *n = cnt;
    return ST_OK;
}
#endif

extern "C" int st_debug_redo(uint32_t abandon_steps, uint32_t coop_min_tris) {
    RT_LOCK;
    g.abandon_steps = abandon_steps;
    g.coop_min_tris = coop_min_tris;
    return ST_OK;
}

extern "C" int st_last_redo_count(uint32_t *count) {
    RT_LOCK;
    if (!count) return set_err(ST_ERR_ARG, "null argument");
    *count = g.main.last_redo;
    return ST_OK;
}

extern "C" int st_graph_stats(uint64_t *captures, uint64_t *launches) {
    RT_LOCK;
    if (captures) *captures = g.graph_captures;
    if (launches) *launches = g.graph_launches;
    return ST_OK;
}

extern "C" int st_render_batch(StSceneHandle *const *scenes, const StView *views, uint32_t n,
                               StFloatPixel *const *pixels, StHitRec *const *hits, StRayCounts *counts,
                               uint32_t flags) {
    if (!scenes || !views) return set_err(ST_ERR_ARG, "null argument");
    RT_LOCK;
    GraphScope gs(flags);
    int want_hits = 0;
    if (hits)
        for (uint32_t i = 0; i < n; i++) want_hits |= hits[i] != NULL;
    return launch(scenes, views, n, want_hits, (flags & ST_BATCH_PERSISTENT) ? 1 : 0, counts, pixels, hits, NULL, NULL);
}

extern "C" int st_render_timed(StSceneHandle *const *scenes, const StView *views, uint32_t n, uint32_t flags,
                               float *kernel_ms, StRayCounts *counts) {
    RT_LOCK;
    GraphScope gs(flags);
    g.main.stage_timing = (flags & ST_BATCH_STAGE_TIMES) != 0;
    g.main.stage_n = 0;
    int rc = launch(scenes, views, n, 0, (flags & ST_BATCH_PERSISTENT) ? 1 : 0, counts, NULL, NULL, NULL, kernel_ms);
    g.main.stage_timing = 0;
    return rc;
}

extern "C" int st_last_stage_ms(float *ms, uint32_t cap, uint32_t *count) {
    RT_LOCK;
    if (!count) return set_err(ST_ERR_ARG, "null argument");
    *count = g.main.stage_n;
    for (uint32_t k = 0; k < g.main.stage_n && k < cap && ms; k++) ms[k] = g.main.stage_ms[k];
    return ST_OK;
}

extern "C" int st_render_batch_bgra8(StSceneHandle *const *scenes, const StView *views, uint32_t n,
                                     uint32_t *const *bgra, StRayCounts *counts, uint32_t flags) {
    if (!scenes || !views || !bgra) return set_err(ST_ERR_ARG, "null argument");
    RT_LOCK;
    GraphScope gs(flags);
    return launch(scenes, views, n, 0, (flags & ST_BATCH_PERSISTENT) ? 1 : 0, counts, NULL, NULL, bgra, NULL);
}

extern "C" int st_ssb(StSceneHandle *scene, const StView *view, int32_t *rects, uint8_t *flags) {
    RT_LOCK;
    if (!scene || !view) return set_err(ST_ERR_ARG, "null argument");
    int rc = ensure_init();
    if (rc) return rc;
    if (!scene->host || !scene->dbuf) return set_err(ST_ERR_ARG, "scene not resident");
    StHostView hv;
    if (st_view_prepare_ex(scene->host, view, &hv, 0)) return set_err(ST_ERR_ARG, "invalid view");
    const uint32_t P = scene->host->prim_count;
    const size_t o_sc = round256(sizeof(KReq)), up = o_sc + round256(sizeof(StDScene));
    const size_t gate_b = round256((P + 1) * 16ull);
    if ((rc = arena_reserve(&g.main.work, up + 2 * gate_b + round256(P + 1) + 512)) || (rc = staging_reserve(&g.main, up))) {
        st_view_release(&hv);
        return rc;
    }
    char *blk = (char *)arena_take(&g.main.work, up);
    int32_t *gate_d = (int32_t *)arena_take(&g.main.work, gate_b);
    int32_t *raw_d = (int32_t *)arena_take(&g.main.work, gate_b);
    uint8_t *flags_d = (uint8_t *)arena_take(&g.main.work, P + 1);
    KReq *rq = (KReq *)g.main.staging;
    memset(rq, 0, sizeof(KReq));
    rq->v = hv.dv;
    rq->v.rects = gate_d;
    const sm3 &m = hv.rotation_matrix_inverted;
    const float rmi[9] = {m.X.x, m.X.y, m.X.z, m.Y.x, m.Y.y, m.Y.z, m.Z.x, m.Z.y, m.Z.z};
    memcpy(rq->ssb_rmi, rmi, sizeof rmi);
    rq->focal_length = hv.focal_length;
    rq->device_ssb = 1;
    rq->ssb_raw = raw_d;
    rq->ssb_flags = flags_d;
    memcpy(g.main.staging + o_sc, &scene->dev, sizeof(StDScene));
    st_view_release(&hv);
    if (P == 0) return ST_OK;
    HIP_TRY(hipMemcpyAsync(blk, g.main.staging, up, hipMemcpyHostToDevice, g.main.stream));
    hipLaunchKernelGGL(k_ssb, dim3(1), dim3(256), 0, g.main.stream, (const StDScene *)(blk + o_sc), (const KReq *)blk);
    HIP_TRY(hipGetLastError());
    if (rects) HIP_TRY(hipMemcpyAsync(rects, raw_d, P * 16ull, hipMemcpyDeviceToHost, g.main.stream));
    if (flags) HIP_TRY(hipMemcpyAsync(flags, flags_d, P, hipMemcpyDeviceToHost, g.main.stream));
    HIP_TRY(hipStreamSynchronize(g.main.stream));
    return ST_OK;
}

extern "C" int st_pick(StSceneHandle *scene, const StView *view, uint32_t n, const uint32_t *xy, StPick *out) {
    RT_LOCK;
    if (!scene || !view || (n && (!xy || !out))) return set_err(ST_ERR_ARG, "null argument");
    int rc = ensure_init();
    if (rc) return rc;
    if (!scene->host || !scene->dbuf) return set_err(ST_ERR_ARG, "scene not resident");
    if (n == 0) return ST_OK;
    StHostView hv;
    if (st_view_prepare(scene->host, view, &hv)) return set_err(ST_ERR_ARG, "invalid view");
    const uint32_t P = scene->host->prim_count;
    /* one upload block: request, scene descriptor, rects, points */
    const size_t o_sc = round256(sizeof(KReq)), o_rect = o_sc + round256(sizeof(StDScene)),
                 o_xy = o_rect + round256((P + 1) * 16ull), o_fault = o_xy + round256(n * 8ull),
                 up = o_fault + 256;
    if ((rc = arena_reserve(&g.main.work, up + round256(n * sizeof(StPick)) + 512)) || (rc = staging_reserve(&g.main, up))) {
        st_view_release(&hv);
        return rc;
    }
    char *blk = (char *)arena_take(&g.main.work, up);
    StPick *out_d = (StPick *)arena_take(&g.main.work, n * sizeof(StPick));
    KReq *rq = (KReq *)g.main.staging;
    memset(rq, 0, sizeof(KReq));
    rq->v = hv.dv;
    rq->v.rects = (const int32_t *)(blk + o_rect);
    StDScene dsc = scene->dev;
    dsc.fault = (uint32_t *)(blk + o_fault);
    dsc.step_cap = g.step_cap;
    /* picking traces sequentially: its mesh traversals stop (a fault) at the
       bound no valid tree reaches, the largest mesh's node count */
    uint32_t max_tris = 0;
    for (uint32_t m = 0; m < scene->host->mesh_count; m++)
        if (scene->host->meshes[m].triangle_count > max_tris) max_tris = scene->host->meshes[m].triangle_count;
    dsc.main_step_limit = 2u * max_tris + 2u;
    if (g.step_cap && g.step_cap < dsc.main_step_limit) dsc.main_step_limit = g.step_cap;
    memcpy(g.main.staging + o_sc, &dsc, sizeof(StDScene));
    memset(g.main.staging + o_fault, 0, 4);
    memcpy(g.main.staging + o_rect, hv.rects, P * 16ull);
    memcpy(g.main.staging + o_xy, xy, n * 8ull);
    KPick pk;
    memcpy(pk.start, hv.start, 12);
    memcpy(pk.right, hv.right, 12);
    memcpy(pk.down, hv.down, 12);
    const sm3 &m = hv.rotation_matrix_inverted;
    const float rmi[9] = {m.X.x, m.X.y, m.X.z, m.Y.x, m.Y.y, m.Y.z, m.Z.x, m.Z.y, m.Z.z};
    memcpy(pk.rmi, rmi, sizeof rmi);
    pk.n = n;
    pk.xy = (const uint32_t *)(blk + o_xy);
    pk.out = out_d;
    st_view_release(&hv);
    const uint32_t m_entries = scene->dev.mesh_stack_size > 1 ? scene->dev.mesh_stack_size : 1;
    HIP_TRY(hipMemcpyAsync(blk, g.main.staging, up, hipMemcpyHostToDevice, g.main.stream));
    hipLaunchKernelGGL(k_pick, dim3((n + 63) / 64), dim3(64), (size_t)64 * (1 + m_entries) * 4, g.main.stream,
                       (const StDScene *)(blk + o_sc), (const KReq *)blk, pk, m_entries);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, out_d, n * sizeof(StPick), hipMemcpyDeviceToHost, g.main.stream));
    uint32_t fault = 0;
    HIP_TRY(hipMemcpyAsync(&fault, blk + o_fault, 4, hipMemcpyDeviceToHost, g.main.stream));
    HIP_TRY(hipStreamSynchronize(g.main.stream));
    if (fault) return set_err(ST_ERR_SCENE, "a traversal ran past its bound (fault bits 0x%x): malformed BVH", fault);
    return ST_OK;
}

extern "C" int st_resolve_bgra8(const StFloatPixel *grid, uint32_t grid_w, uint32_t grid_h, uint32_t *out) {
    RT_LOCK;
    if (!grid || !out) return set_err(ST_ERR_ARG, "null argument");
    if (grid_w < 2 || grid_h < 2 || grid_w > 0xFFFF || grid_h > 0xFFFF) return set_err(ST_ERR_ARG, "bad grid size");
    int rc = ensure_init();
    if (rc) return rc;
    size_t gbytes = (size_t)grid_w * grid_h * 24, opx = (size_t)(grid_w - 1) * (grid_h - 1);
    if ((rc = arena_reserve(&g.main.work, round256(sizeof(KReq)) + round256(gbytes) + round256(opx * 4) + 256))) return rc;
    if ((rc = staging_reserve(&g.main, round256(sizeof(KReq))))) return rc;
    KReq *req_d = (KReq *)arena_take(&g.main.work, sizeof(KReq));
    void *grid_d = arena_take(&g.main.work, gbytes);
    uint32_t *out_d = (uint32_t *)arena_take(&g.main.work, opx * 4);
    KReq *req_h = (KReq *)g.main.staging;
    memset(req_h, 0, sizeof(KReq));
    req_h->v.pixels = grid_d;
    req_h->v.width1 = grid_w;
    req_h->v.height1 = grid_h;
    req_h->bgra = out_d;
    HIP_TRY(hipMemcpyAsync(grid_d, grid, gbytes, hipMemcpyHostToDevice, g.main.stream));
    HIP_TRY(hipMemcpyAsync(req_d, req_h, sizeof(KReq), hipMemcpyHostToDevice, g.main.stream));
    hipLaunchKernelGGL(k_resolve, dim3((uint32_t)((opx + 255) / 256)), dim3(256), 0, g.main.stream, (const KReq *)req_d, 1u,
                       (uint32_t)opx);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, out_d, opx * 4, hipMemcpyDeviceToHost, g.main.stream));
    HIP_TRY(hipStreamSynchronize(g.main.stream));
    return ST_OK;
}

extern "C" int st_host_alloc(size_t bytes, void **out) {
    if (!out) return set_err(ST_ERR_ARG, "null argument");
    *out = NULL;
    int rc = ensure_init();
    if (rc) return rc;
    HIP_TRY(hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault));
    return ST_OK;
}

extern "C" void st_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}

/* ---- launch tuning ------------------------------------------------------ */

static int tuning_valid(const StTuning *t) {
    return t->tile_shift <= 6 && t->blocks_per_cu >= 1 && t->blocks_per_cu <= 64 && t->persistent <= 2 &&
           t->cost_order <= 2 && t->ray_sort <= 1 && t->bounce_order <= 1 &&
           (t->stack_entries == 0 || (t->stack_entries >= 4 && t->stack_entries <= 64 &&
                                      (t->stack_entries & (t->stack_entries - 1)) == 0)) &&
           (t->primary_waves == 0 || t->primary_waves == 3 || t->primary_waves == 4);
}

extern "C" int st_set_tuning(const StTuning *t) {
    RT_LOCK;
    if (!t || !tuning_valid(t)) return set_err(ST_ERR_ARG, "invalid tuning");
    int rc = ensure_init();
    if (rc) return rc;
    g.tune = *t;
    return ST_OK;
}

extern "C" int st_get_tuning(StTuning *t) {
    RT_LOCK;
    if (!t) return set_err(ST_ERR_ARG, "null argument");
    int rc = ensure_init();
    if (rc) return rc;
    *t = g.tune;
    return ST_OK;
}

/* Time every candidate launch shape and claim order on the caller's
   requests (median of `reps` render-kernel times from HIP events) and keep
   the fastest. The candidates: 8x8, 16x4 and 32x2 tiles; persistent grids
   of 2, 4 or 8 blocks per CU, or one block per 4 tiles in XCD-aware order
   with at most 2, 3 or any number of blocks per CU resident; tiles claimed
   in tile order or slowest first (that order needs the previous launch's
   tile times: one untimed launch of the shape comes first). Then the best
   with the bounce-ray sort toggled, with the other persistent tile claim
   (global counter or XCD chunks), in super-tile claim order, with each
   traversal-stack window and primary-stage occupancy, and with the bounce
   queue's parts in the other order. */
extern "C" int st_autotune(StSceneHandle *const *scenes, const StView *views, uint32_t n, uint32_t reps,
                           StTuning *best, float *best_ms) {
    RT_LOCK;
    if (!scenes || !views || !n) return set_err(ST_ERR_ARG, "no requests");
    int rc = ensure_init();
    if (rc) return rc;
    if (reps == 0) reps = 3;
    if (reps > 15) reps = 15;
    const StTuning saved = g.tune;
    StTuning cand[36 + 9];
    uint32_t nc = 0;
    for (uint32_t sh = 3; sh <= 5; sh++) {
        const uint32_t bpc[3] = {2, 4, 8}, cap[3] = {2, 3, 8};
        for (uint32_t order = 0; order <= 1; order++) {
            for (int k = 0; k < 3; k++)
                cand[nc++] = StTuning{sh, bpc[k], 1u, order, saved.ray_sort, saved.stack_entries,
                                      saved.primary_waves, saved.bounce_order};
            for (int k = 0; k < 3; k++)
                cand[nc++] = StTuning{sh, cap[k], 0u, order, saved.ray_sort, saved.stack_entries,
                                      saved.primary_waves, saved.bounce_order};
        }
    }
    float best_t = INFINITY;
    StTuning best_c = saved;
    /* the shapes and claim orders; then the best of them with the bounce-ray
       sort toggled; then, when persistent, with the other tile claim (global
       counter or XCD chunks); then in super-tile order; then the best with
       each traversal-stack window and primary-stage occupancy; then with
       the bounce queue's parts in the other order */
    const uint32_t windows[5][2] = {{0u, 3u}, {16u, 3u}, {32u, 3u}, {8u, 4u}, {16u, 4u}};
    for (uint32_t c = 0; c < nc + 9u; c++) {
        if (c == nc) {
            cand[c] = best_c;
            cand[c].ray_sort ^= 1u;
        } else if (c == nc + 8u) {
            cand[c] = best_c;
            cand[c].bounce_order ^= 1u;
        } else if (c == nc + 1u) {
            if (!best_c.persistent) continue;
            cand[c] = best_c;
            cand[c].persistent = best_c.persistent == 2u ? 1u : 2u;
        } else if (c == nc + 2u) {
            if (best_c.cost_order == 2u) continue;
            cand[c] = best_c;
            cand[c].cost_order = 2u;
        } else if (c > nc + 2u && c < nc + 8u) {
            cand[c] = best_c;
            cand[c].stack_entries = windows[c - nc - 3u][0];
            cand[c].primary_waves = windows[c - nc - 3u][1];
            if (cand[c].stack_entries == best_c.stack_entries && cand[c].primary_waves == best_c.primary_waves)
                continue;
        }
        g.tune = cand[c];
        float times[15];
        if (cand[c].cost_order == 1u)   /* slowest first: the shape's tile times first */
            rc = launch(scenes, views, n, 0, (int)cand[c].persistent, NULL, NULL, NULL, NULL, &times[0]);
        for (uint32_t r = 0; r < reps && !rc; r++)
            rc = launch(scenes, views, n, 0, (int)cand[c].persistent, NULL, NULL, NULL, NULL, &times[r]);
        if (rc) break;
        std::sort(times, times + reps);
        if (times[reps / 2] < best_t) {
            best_t = times[reps / 2];
            best_c = cand[c];
        }
    }
    g.tune = rc ? saved : best_c;
    if (best) *best = g.tune;
    if (best_ms) *best_ms = best_t;
    return rc;
}
