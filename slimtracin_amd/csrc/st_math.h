/*
 * st_math.h - fp32 vector/quaternion helpers shared by the host C code and
 * the gfx950 kernels.
 *
 * The operation order of every helper is the one the reference uses
 * (SlimTracin math/vec3.h:143-266, math/quat.h:15-51, math/mat3.h:79-117,
 * core/base.h:390-411): results must be bit-identical to the reference's CPU
 * renderer, so sums are written left to right, fast_mul_add(a,b,c) is a
 * separate multiply and add, and everything is built with contraction off.
 */
#ifndef ST_MATH_H
#define ST_MATH_H

#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#define STM __host__ __device__ static inline
#else
#define STM static inline
#endif

typedef struct { float x, y, z; } sv3;
typedef struct { float x, y; } sv2;
typedef struct { sv3 v; float w; } sq4;       /* quaternion: axis v, amount w */
typedef struct { sv3 X, Y, Z; } sm3;          /* rows */
typedef struct { sv3 min, max; } sbox;

STM sv3 V3(float x, float y, float z) { sv3 o; o.x = x; o.y = y; o.z = z; return o; }
STM sv3 V3s(float s) { return V3(s, s, s); }
STM sv2 V2(float x, float y) { sv2 o; o.x = x; o.y = y; return o; }

STM sv3 v_add(sv3 a, sv3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
STM sv3 v_sub(sv3 a, sv3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
STM sv3 v_mul(sv3 a, sv3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
STM sv3 v_scale(sv3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
STM sv3 v_neg(sv3 a) { return V3(-a.x, -a.y, -a.z); }
/* mulAddVec3: v*f + a per component */
STM sv3 v_mad(sv3 v, sv3 f, sv3 a) { return V3(v.x * f.x + a.x, v.y * f.y + a.y, v.z * f.z + a.z); }
/* scaleAddVec3: v*s + a per component */
STM sv3 v_smad(sv3 v, float s, sv3 a) { return V3(v.x * s + a.x, v.y * s + a.y, v.z * s + a.z); }
STM float v_dot(sv3 a, sv3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
STM sv3 v_cross(sv3 a, sv3 b) {
    return V3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
STM float v_sqlen(sv3 v) { return (v.x * v.x) + (v.y * v.y) + (v.z * v.z); }
STM float v_len(sv3 v) { return sqrtf(v_sqlen(v)); }
STM sv3 v_norm(sv3 v) { return v_scale(v, 1.0f / v_len(v)); }
STM sv3 v_rcp(sv3 v) { return V3(1.0f / v.x, 1.0f / v.y, 1.0f / v.z); }
STM sv3 v_one_minus(sv3 v) { return V3(1.0f - v.x, 1.0f - v.y, 1.0f - v.z); }
STM float f_clamp01(float v) { float mn = v < 1.0f ? v : 1.0f; return mn > 0.0f ? mn : 0.0f; }
STM float v_dotc(sv3 a, sv3 b) { return f_clamp01(v_dot(a, b)); }
STM float v_maxc(sv3 v) { float o = v.x; if (v.y > o) o = v.y; if (v.z > o) o = v.z; return o; }
STM sv3 v_mat(sv3 in, sm3 m) {
    return V3(in.x * m.X.x + in.y * m.Y.x + in.z * m.Z.x,
              in.x * m.X.y + in.y * m.Y.y + in.z * m.Z.y,
              in.x * m.X.z + in.y * m.Y.z + in.z * m.Z.z);
}

STM sq4 q_norm(sq4 q) {
    float f = 1.0f / sqrtf(q.v.x * q.v.x + q.v.y * q.v.y + q.v.z * q.v.z + q.w * q.w);
    sq4 o; o.v = v_scale(q.v, f); o.w = q.w * f; return o;
}
STM sq4 q_conj(sq4 q) { sq4 o; o.w = q.w; o.v = v_neg(q.v); return o; }
/* mulVec3Quat: v + 2*(w*(a x v) + a x (a x v)) in the reference's order */
STM sv3 q_rot(sv3 v, sq4 q) {
    sv3 out = v_cross(q.v, v);
    sv3 qqv = v_cross(q.v, out);
    out = v_smad(out, q.w, qqv);
    return v_smad(out, 2.0f, v);
}
STM sq4 q_mul(sq4 a, sq4 b) {
    sq4 o;
    o.w = a.w * b.w - a.v.x * b.v.x - a.v.y * b.v.y - a.v.z * b.v.z;
    o.v.x = a.w * b.v.x + a.v.x * b.w + a.v.y * b.v.z - a.v.z * b.v.y;
    o.v.y = a.w * b.v.y - a.v.x * b.v.z + a.v.y * b.w + a.v.z * b.v.x;
    o.v.z = a.w * b.v.z + a.v.x * b.v.y - a.v.y * b.v.x + a.v.z * b.w;
    return o;
}
/* getPointOnUnitCircle: rational parametrisation */
STM sv2 unit_circle(float t) {
    float t2 = t * t;
    float f = 1 / (1 + t2);
    return V2(f - f * t2, f * 2 * t);
}
/* getRotationAroundAxis */
STM sq4 q_axis(sv3 axis, float amount) {
    sv2 sc = unit_circle(amount);
    sq4 o; o.v = v_scale(axis, sc.y); o.w = sc.x;
    return q_norm(o);
}

STM sbox box_merge(sbox a, sbox b) {
    sbox m;
    m.min.x = a.min.x < b.min.x ? a.min.x : b.min.x;
    m.min.y = a.min.y < b.min.y ? a.min.y : b.min.y;
    m.min.z = a.min.z < b.min.z ? a.min.z : b.min.z;
    m.max.x = a.max.x > b.max.x ? a.max.x : b.max.x;
    m.max.y = a.max.y > b.max.y ? a.max.y : b.max.y;
    m.max.z = a.max.z > b.max.z ? a.max.z : b.max.z;
    return m;
}
/* getSurfaceAreaOfAABB (half area) */
STM float box_area(sbox b) {
    sv3 e = v_sub(b.max, b.min);
    return e.x * e.y + e.y * e.z + e.z * e.x;
}

STM sm3 m_identity(void) { sm3 m; m.X = V3(1, 0, 0); m.Y = V3(0, 1, 0); m.Z = V3(0, 0, 1); return m; }
STM sm3 m_mul(sm3 a, sm3 b) {
    sm3 o;
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
STM sm3 m_transpose(sm3 m) {
    sm3 o;
    o.X = V3(m.X.x, m.Y.x, m.Z.x);
    o.Y = V3(m.X.y, m.Y.y, m.Z.y);
    o.Z = V3(m.X.z, m.Y.z, m.Z.z);
    return o;
}
/* invMat3: cofactors over the determinant, in the reference's order */
STM sm3 m_inverse(sm3 m) {
    sm3 o;
    float d = 1.0f / (+m.X.x * (m.Y.y * m.Z.z - m.Z.y * m.Y.z)
                      - m.Y.x * (m.X.y * m.Z.z - m.Z.y * m.X.z)
                      + m.Z.x * (m.X.y * m.Y.z - m.Y.y * m.X.z));
    o.X.x = +(m.Y.y * m.Z.z - m.Z.y * m.Y.z) * d;
    o.Y.x = -(m.Y.x * m.Z.z - m.Z.x * m.Y.z) * d;
    o.Z.x = +(m.Y.x * m.Z.y - m.Z.x * m.Y.y) * d;
    o.X.y = -(m.X.y * m.Z.z - m.Z.y * m.X.z) * d;
    o.Y.y = +(m.X.x * m.Z.z - m.Z.x * m.X.z) * d;
    o.Z.y = -(m.X.x * m.Z.y - m.Z.x * m.X.y) * d;
    o.X.z = +(m.X.y * m.Y.z - m.Y.y * m.X.z) * d;
    o.Y.z = -(m.X.x * m.Y.z - m.Y.x * m.X.z) * d;
    o.Z.z = +(m.X.x * m.Y.y - m.Y.x * m.X.y) * d;
    return o;
}
/* convertRotationMatrixToQuaternion */
STM sq4 q_from_matrix(sm3 r) {
    sq4 o;
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
    if (bi == 0) { o.w = bv; o.v.x = (r.Y.z - r.Z.y) * mult; o.v.y = (r.Z.x - r.X.z) * mult; o.v.z = (r.X.y - r.Y.x) * mult; }
    else if (bi == 1) { o.w = (r.Y.z - r.Z.y) * mult; o.v.x = bv; o.v.y = (r.X.y + r.Y.x) * mult; o.v.z = (r.Z.x + r.X.z) * mult; }
    else if (bi == 2) { o.w = (r.Z.x - r.X.z) * mult; o.v.x = (r.X.y + r.Y.x) * mult; o.v.y = bv; o.v.z = (r.Y.z + r.Z.y) * mult; }
    else { o.w = (r.X.y - r.Y.x) * mult; o.v.x = (r.Z.x + r.X.z) * mult; o.v.y = (r.Y.z + r.Z.y) * mult; o.v.z = bv; }
    return o;
}

#endif
