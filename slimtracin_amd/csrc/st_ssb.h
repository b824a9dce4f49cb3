/*
 * st_ssb.h - screen-space bounds of one primitive, shared by the host view
 * set-up and the device SSB kernel so that both run the same operations:
 * updatePrimitiveSSB's bounding radius (render/SSB.h:137-166) and
 * computeSSB (render/SSB.h:5-135) with its (u16) casts through 32-bit ints.
 */
#ifndef ST_SSB_H
#define ST_SSB_H

#include "st_math.h"
#include "st_blob.h"

/* Frame dimensions as updateDimensions keeps them, QCAA on
   (core/base.h:344-358): the traced grid is (w+1) x (h+1). */
typedef struct {
    uint32_t width, height;
    float f_width, f_height, h_width, h_height, width_over_height, height_over_width;
} StSsbDims;

STM StSsbDims st_ssb_dims(uint32_t width1, uint32_t height1) {
    StSsbDims d;
    d.width = width1;
    d.height = height1;
    d.f_width = (float)d.width;
    d.f_height = (float)d.height;
    d.h_width = d.f_width / 2;
    d.h_height = d.f_height / 2;
    d.width_over_height = d.f_width / d.f_height;
    d.height_over_width = d.f_height / d.f_width;
    return d;
}

/* Bounding radius by primitive type; 0 with *known = 0 for other types
   (updatePrimitiveSSB returns early for them). mesh_aabb: 6 floats. */
STM float st_ssb_radius(uint32_t type, sv3 scale, const float *mesh_aabb, int *known) {
    *known = 1;
    switch (type) {
        case ST_PRIM_QUAD:
        case ST_PRIM_BOX: return v_len(scale);
        case ST_PRIM_TETRAHEDRON: return v_maxc(scale) * 1.41421356237f;
        case ST_PRIM_SPHERE: return v_maxc(scale);
        case ST_PRIM_MESH: {
            const float min_r = v_len(v_mul(scale, V3(mesh_aabb[0], mesh_aabb[1], mesh_aabb[2])));
            const float max_r = v_len(v_mul(scale, V3(mesh_aabb[3], mesh_aabb[4], mesh_aabb[5])));
            return max_r > min_r ? max_r : min_r;
        }
        default: *known = 0; return 0;
    }
}

/* computeSSB: b = min.x, min.y, max.x, max.y of the view-space sphere
   (pos, r); returns whether the primitive is visible. */
STM int st_ssb_rect(int32_t *b, sv3 pos, float r, float fl, const StSsbDims *d) {
    b[0] = (int32_t)d->width + 1;
    b[2] = (int32_t)d->width + 1;
    b[1] = (int32_t)d->height + 1;
    b[3] = (int32_t)d->height + 1;
    if (pos.z <= -r) return 0;
    const float x = pos.x, y = pos.y, z = pos.z;
    float left, right, top, bottom, factor;
    if (z <= r) {
        const float h = fabsf(z) / fl;
        top = y + r;
        bottom = y - r;
        if (h < bottom || top < -h) return 0;
        const float w = h * d->width_over_height;
        right = x + r;
        left = x - r;
        if (w < left || right < -w) return 0;
        b[0] = -w < left ? (uint16_t)(int32_t)(d->f_width * (left + w) / (2 * w)) : 0;
        b[2] = right < w ? (uint16_t)(int32_t)(d->f_width * (right + w) / (2 * w)) : (int32_t)d->width;
        b[1] = top < h ? (uint16_t)(int32_t)(d->f_height * (h - top) / (2 * h)) : 0;
        b[3] = -h < bottom ? (uint16_t)(int32_t)(d->f_height * (h - bottom) / (2 * h)) : (int32_t)d->height;
        return 1;
    }
    const float den = z * z - r * r;
    factor = fl / den;
    const float yz = y * z;
    float sqr = y * y + den;
    sqr = r * sqrtf(sqr);
    bottom = factor * (yz - sqr);
    top = factor * (yz + sqr);
    if (bottom < 1 && top > -1) {
        factor *= d->height_over_width;
        const float xz = x * z;
        sqr = r * sqrtf(x * x + den);
        left = factor * (xz - sqr);
        right = factor * (xz + sqr);
        if (left < 1 && right > -1) {
            bottom = bottom > -1 ? bottom : -1;
            bottom += 1;
            top = top < 1 ? top : 1;
            top += 1;
            left = left > -1 ? left : -1;
            left += 1;
            right = right < 1 ? right : 1;
            right += 1;
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

/* One primitive's gate rect as the primary-ray pass uses it: the SSB rect
   when the primitive is visible, else an empty rect (min > max); flags in
   and out (IS_VISIBLE recomputed for the known types). raw receives the
   rect as computeSSB left it (zero for unknown types). */
STM void st_ssb_prim(uint32_t type, uint8_t *flags, sv3 pos, sv3 scale, const float *mesh_aabb, sv3 cam, sm3 rmi,
                     float fl, const StSsbDims *d, int32_t *gate, int32_t *raw) {
    const sv3 vsp = v_mat(v_sub(pos, cam), rmi);
    int known;
    const float radius = st_ssb_radius(type, scale, mesh_aabb, &known);
    uint8_t f = *flags;
    if (known) {
        f &= (uint8_t)~ST_IS_VISIBLE;
        if (st_ssb_rect(raw, vsp, radius, fl, d)) f |= ST_IS_VISIBLE;
    } else {
        raw[0] = raw[1] = raw[2] = raw[3] = 0;
    }
    *flags = f;
    if (f & ST_IS_VISIBLE) {
        gate[0] = raw[0];
        gate[1] = raw[1];
        gate[2] = raw[2];
        gate[3] = raw[3];
    } else {
        gate[0] = gate[1] = 1;
        gate[2] = gate[3] = 0;
    }
}

#endif
