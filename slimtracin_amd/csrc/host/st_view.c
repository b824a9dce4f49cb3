/*
 * st_view.c - per-request camera, frame and screen-space-bounds set-up.
 *
 * Restates, for one StView:
 *   rotateXform3 from identity (scene/xform.h:8-20, math/mat3.h:137-183),
 *   updateDimensions with QCAA (core/base.h:344-358),
 *   setViewportProjectionPlane (core/init.h:545-559),
 *   the row-start accumulation of renderSceneOnCPU (render/raytracer.h:156-170:
 *     each row starts at the previous row's start + down, a sequential fp32 sum),
 *   updatePrimitiveSSB / computeSSB (render/SSB.h:5-166) with its (u16) casts.
 * The SSB result is folded into one rect per primitive: a primitive that is
 * not IS_VISIBLE gets an empty rect, which the primary-ray gate rejects for
 * every pixel exactly as the flag test does.
 */
#include <stdlib.h>
#include <string.h>

#include "st_host.h"
#include "slimtracin.h"
#include "../st_ssb.h"

typedef struct {
    uint16_t width, height;
    float f_width, f_height, h_width, h_height, width_over_height, height_over_width;
} Dims;

static void dims_init(Dims *d, uint16_t w, uint16_t h) {
    d->width = (uint16_t)(w + 1);
    d->height = (uint16_t)(h + 1);
    d->f_width = (float)d->width;
    d->f_height = (float)d->height;
    d->h_width = d->f_width / 2;
    d->h_height = d->f_height / 2;
    d->width_over_height = d->f_width / d->f_height;
    d->height_over_width = d->f_height / d->f_width;
}

static sm3 rotation_matrix(float yaw_amt, float pitch_amt, float roll_amt) {
    sm3 yaw = m_identity(), pitch = m_identity(), roll = m_identity();
    if (yaw_amt) {
        sv2 xy = unit_circle(yaw_amt);
        sm3 m = yaw;
        yaw.X.x = xy.x * m.X.x - xy.y * m.X.z;
        yaw.Y.x = xy.x * m.Y.x - xy.y * m.Y.z;
        yaw.Z.x = xy.x * m.Z.x - xy.y * m.Z.z;
        yaw.X.z = xy.x * m.X.z + xy.y * m.X.x;
        yaw.Y.z = xy.x * m.Y.z + xy.y * m.Y.x;
        yaw.Z.z = xy.x * m.Z.z + xy.y * m.Z.x;
    }
    if (pitch_amt) {
        sv2 xy = unit_circle(pitch_amt);
        sm3 m = pitch;
        pitch.X.y = xy.x * m.X.y + xy.y * m.X.z;
        pitch.Y.y = xy.x * m.Y.y + xy.y * m.Y.z;
        pitch.Z.y = xy.x * m.Z.y + xy.y * m.Z.z;
        pitch.X.z = xy.x * m.X.z - xy.y * m.X.y;
        pitch.Y.z = xy.x * m.Y.z - xy.y * m.Y.y;
        pitch.Z.z = xy.x * m.Z.z - xy.y * m.Z.y;
    }
    if (roll_amt) {
        sv2 xy = unit_circle(roll_amt);
        sm3 m = roll;
        roll.X.x = xy.x * m.X.x + xy.y * m.X.y;
        roll.Y.x = xy.x * m.Y.x + xy.y * m.Y.y;
        roll.Z.x = xy.x * m.Z.x + xy.y * m.Z.y;
        roll.X.y = xy.x * m.X.y - xy.y * m.X.x;
        roll.Y.y = xy.x * m.Y.y - xy.y * m.Y.x;
        roll.Z.y = xy.x * m.Z.y - xy.y * m.Z.x;
    }
    return m_mul(m_mul(pitch, yaw), roll);
}

int st_view_prepare(const StScene *sc, const StView *v, StHostView *hv) { return st_view_prepare_ex(sc, v, hv, 1); }

static int view_setup(const StScene *sc, const StView *v, StHostView *hv, int with_ssb, float *rows_buf);

int st_view_prepare_ex(const StScene *sc, const StView *v, StHostView *hv, int with_ssb) {
    return view_setup(sc, v, hv, with_ssb, NULL);
}

int st_view_prepare_rows(const StScene *sc, const StView *v, StHostView *hv, float *row_starts) {
    if (!row_starts) return -1;
    return view_setup(sc, v, hv, 0, row_starts);
}

/* rows_buf: the caller's row-start buffer (4 floats per traced row); then
   nothing is allocated and the gate rects are left to the device SSB pass. */
static int view_setup(const StScene *sc, const StView *v, StHostView *hv, int with_ssb, float *rows_buf) {
    memset(hv, 0, sizeof *hv);
    if (v->width == 0 || v->height == 0 || v->width >= 0xFFFF || v->height >= 0xFFFF) return -1;
    if (v->render_mode > ST_MODE_UVS) return -1;
    sm3 rm = rotation_matrix(v->yaw, v->pitch, v->roll);
    sm3 rmi = m_transpose(rm);
    sq4 rot = q_from_matrix(rm), rot_inv = q_from_matrix(rmi);
    sv3 right_dir = rm.X, up = rm.Y, forward = rm.Z;
    Dims d;
    dims_init(&d, v->width, v->height);

    /* projection plane with QCAA offset 0 */
    const float offset = 0.0f;
    sv3 start = v_scale(forward, d.f_height * v->focal_length);
    sv3 right = v_scale(right_dir, offset - (float)d.width);
    sv3 down = v_scale(up, d.f_height - offset);
    start = v_add(start, right);
    start = v_add(start, down);
    right = v_scale(right_dir, 2);
    down = v_scale(up, -2);

    if (rows_buf) {
        with_ssb = 0;
        hv->row_starts = rows_buf;
    } else {
        hv->row_starts = (float *)malloc(4 * sizeof(float) * d.height);
        hv->rects = (int32_t *)malloc(4 * sizeof(int32_t) * (sc->prim_count + 1));
        hv->flags = (uint8_t *)malloc(sc->prim_count + 1);
        hv->raw_rects = (int32_t *)malloc(4 * sizeof(int32_t) * (sc->prim_count + 1));
        if (!hv->row_starts || !hv->rects || !hv->flags || !hv->raw_rects) { st_view_release(hv); return -2; }
    }
    sv3 s = start;
    for (uint32_t y = 0; y < d.height; y++) {
        hv->row_starts[4 * y] = s.x;
        hv->row_starts[4 * y + 1] = s.y;
        hv->row_starts[4 * y + 2] = s.z;
        hv->row_starts[4 * y + 3] = 0;
        s = v_add(s, down);
    }

    sv3 cam = V3(v->position[0], v->position[1], v->position[2]);
    const StSsbDims sd = st_ssb_dims(d.width, d.height);
    for (uint32_t i = 0; with_ssb && i < sc->prim_count; i++) {
        const StDPrim *p = sc->prims + i;
        uint8_t flags = (uint8_t)(p->flags_mat & 0xFFu);
        const float *mb = p->type == ST_PRIM_MESH ? &sc->mesh_aabbs[p->mesh_id].min.x : NULL;
        st_ssb_prim(p->type, &flags, V3(p->pos[0], p->pos[1], p->pos[2]), V3(p->scale[0], p->scale[1], p->scale[2]),
                    mb, cam, rmi, v->focal_length, &sd, hv->rects + 4 * i, hv->raw_rects + 4 * i);
        hv->flags[i] = flags;
    }

    StDView *dv = &hv->dv;
    memcpy(dv->cam_pos, v->position, 12);
    dv->cone_angle = v->cone_angle;
    dv->cam_rot_inv[0] = rot_inv.v.x;
    dv->cam_rot_inv[1] = rot_inv.v.y;
    dv->cam_rot_inv[2] = rot_inv.v.z;
    dv->cam_rot_inv[3] = rot_inv.w;
    dv->right[0] = right.x;
    dv->right[1] = right.y;
    dv->right[2] = right.z;
    dv->width1 = d.width;
    dv->height1 = d.height;
    dv->mode = v->render_mode;
    dv->depth = (uint8_t)v->depth;   /* Trace.depth is a u8 (core/types.h:199) */
    hv->start[0] = start.x; hv->start[1] = start.y; hv->start[2] = start.z;
    hv->right[0] = right.x; hv->right[1] = right.y; hv->right[2] = right.z;
    hv->down[0] = down.x; hv->down[1] = down.y; hv->down[2] = down.z;
    hv->rot[0] = rot.v.x; hv->rot[1] = rot.v.y; hv->rot[2] = rot.v.z; hv->rot[3] = rot.w;
    memcpy(hv->rot_inv, dv->cam_rot_inv, 16);
    hv->rotation_matrix_inverted = rmi;
    hv->focal_length = v->focal_length;
    hv->device_ssb = !with_ssb;
    return 0;
}

int st_view_prepare_frame(const StScene *sc, const StSeamFrame *f, const int32_t *screen_bounds, StHostView *hv) {
    memset(hv, 0, sizeof *hv);
    if (f->width == 0 || f->height == 0 || f->width == 0xFFFF || f->height == 0xFFFF) return -1;
    if (f->render_mode > ST_MODE_UVS) return -1;
    if (sc->prim_count && !screen_bounds) return -1;
    hv->row_starts = (float *)malloc(4 * sizeof(float) * f->height);
    hv->rects = (int32_t *)malloc(4 * sizeof(int32_t) * (sc->prim_count + 1));
    hv->flags = (uint8_t *)malloc(sc->prim_count + 1);
    hv->raw_rects = (int32_t *)malloc(4 * sizeof(int32_t) * (sc->prim_count + 1));
    if (!hv->row_starts || !hv->rects || !hv->flags || !hv->raw_rects) { st_view_release(hv); return -2; }
    /* start += down per row (render/raytracer.h:156-170) */
    sv3 s = V3(f->start[0], f->start[1], f->start[2]);
    const sv3 down = V3(f->down[0], f->down[1], f->down[2]);
    for (uint32_t y = 0; y < f->height; y++) {
        hv->row_starts[4 * y] = s.x;
        hv->row_starts[4 * y + 1] = s.y;
        hv->row_starts[4 * y + 2] = s.z;
        hv->row_starts[4 * y + 3] = 0;
        s = v_add(s, down);
    }
    for (uint32_t i = 0; i < sc->prim_count; i++) {
        uint8_t flags = (uint8_t)(sc->prims[i].flags_mat & 0xFFu);
        int32_t *rc = hv->rects + 4 * i;
        memcpy(rc, screen_bounds + 4 * i, 16);
        memcpy(hv->raw_rects + 4 * i, rc, 16);
        hv->flags[i] = flags;
        if (!(flags & ST_IS_VISIBLE)) { rc[0] = rc[1] = 1; rc[2] = rc[3] = 0; }
    }
    StDView *dv = &hv->dv;
    memcpy(dv->cam_pos, f->camera_position, 12);
    dv->cone_angle = f->cone_angle;
    memcpy(dv->cam_rot_inv, f->camera_rotation_inverted, 16);
    memcpy(dv->right, f->right, 12);
    dv->width1 = f->width;
    dv->height1 = f->height;
    dv->mode = f->render_mode;
    dv->depth = (uint8_t)f->depth;   /* Trace.depth is a u8 (core/types.h:199) */
    memcpy(hv->start, f->start, 12);
    memcpy(hv->right, f->right, 12);
    memcpy(hv->down, f->down, 12);
    memcpy(hv->rot_inv, f->camera_rotation_inverted, 16);
    return 0;
}

void st_view_release(StHostView *hv) {
    free(hv->row_starts);
    free(hv->rects);
    free(hv->flags);
    free(hv->raw_rects);
    hv->raw_rects = NULL;
    hv->row_starts = NULL;
    hv->rects = NULL;
    hv->flags = NULL;
}

/* Host-only introspection of the set-up a request gets (no device needed):
   SSB rects and flags after updateSceneSSB, and the camera as 17 floats
   (projection-plane start, right, down; rotation_inverted; rotation). */
int st_host_view_info(const void *blob, size_t size, const StView *v, int32_t *rects, uint8_t *flags, float *camera) {
    const char *why = NULL;
    StScene *sc = st_scene_from_blob(blob, size, &why);
    if (!sc) return -2;
    StHostView hv;
    int rc = st_view_prepare(sc, v, &hv);
    if (rc) { st_scene_free(sc); return -1; }
    for (uint32_t i = 0; i < sc->prim_count; i++) {
        if (rects) memcpy(rects + 4 * i, hv.raw_rects + 4 * i, 16);
        if (flags) flags[i] = hv.flags[i];
    }
    if (camera) {
        float c[17] = {hv.start[0], hv.start[1], hv.start[2], hv.right[0], hv.right[1], hv.right[2],
                       hv.down[0], hv.down[1], hv.down[2], hv.rot_inv[0], hv.rot_inv[1], hv.rot_inv[2], hv.rot_inv[3],
                       hv.rot[0], hv.rot[1], hv.rot[2], hv.rot[3]};
        memcpy(camera, c, sizeof c);
    }
    st_view_release(&hv);
    st_scene_free(sc);
    return 0;
}

/* Host-only: the scene BVH the library builds for a blob. */
int st_host_scene_bvh(const void *blob, size_t size, StBVHNodeRec *nodes, uint32_t *leaf_ids, uint32_t *node_count,
                      uint32_t *height) {
    const char *why = NULL;
    StScene *sc = st_scene_from_blob(blob, size, &why);
    if (!sc) return -2;
    memcpy(nodes, sc->scene_nodes, sizeof(StBVHNodeRec) * sc->scene_node_count);
    memcpy(leaf_ids, sc->scene_leaf_ids, 4u * sc->prim_count);
    *node_count = sc->scene_node_count;
    *height = sc->scene_height;
    st_scene_free(sc);
    return 0;
}

/* Host-only: parse + validate a blob; returns 0 or a negative status and
   copies the reason into `why` (up to `cap` bytes). */
int st_host_validate(const void *blob, size_t size, char *why, size_t cap) {
    const char *reason = NULL;
    StScene *sc = st_scene_from_blob(blob, size, &reason);
    if (!sc) {
        if (why && cap) { strncpy(why, reason ? reason : "invalid", cap - 1); why[cap - 1] = 0; }
        return -2;
    }
    st_scene_free(sc);
    if (why && cap) why[0] = 0;
    return 0;
}
