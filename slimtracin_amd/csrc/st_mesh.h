/*
 * st_mesh.h - per-triangle build steps shared by the host SAH mesh build
 * and the device mesh LBVH build, so both produce the same bits.
 *
 * updateMeshBVH (builder_top_down.h:272-364): a triangle's leaf box is the
 * min/max of its vertices (v2,v3 first, then v1), padded by EPS = 1e-4 on
 * each axis where it is flat; its record holds world_to_tangent =
 * inverse([v3-v1, v2-v1, n]) with n the unit normal, v1, n, the vertex
 * normals and uvs, the parallelogram area |(v3-v1) x (v2-v1)| and the uv
 * parallelogram area.
 */
#ifndef ST_MESH_H
#define ST_MESH_H

#include <string.h>

#include "st_blob.h"
#include "st_layout.h"
#include "st_math.h"

STM sbox st_tri_box(const float *v1, const float *v2, const float *v3) {
    const float EPS = 0.0001f;
    sbox bx;
    bx.min.x = v2[0] < v3[0] ? v2[0] : v3[0];
    bx.min.y = v2[1] < v3[1] ? v2[1] : v3[1];
    bx.min.z = v2[2] < v3[2] ? v2[2] : v3[2];
    bx.max.x = v2[0] > v3[0] ? v2[0] : v3[0];
    bx.max.y = v2[1] > v3[1] ? v2[1] : v3[1];
    bx.max.z = v2[2] > v3[2] ? v2[2] : v3[2];
    bx.min.x = bx.min.x < v1[0] ? bx.min.x : v1[0];
    bx.min.y = bx.min.y < v1[1] ? bx.min.y : v1[1];
    bx.min.z = bx.min.z < v1[2] ? bx.min.z : v1[2];
    bx.max.x = bx.max.x > v1[0] ? bx.max.x : v1[0];
    bx.max.y = bx.max.y > v1[1] ? bx.max.y : v1[1];
    bx.max.z = bx.max.z > v1[2] ? bx.max.z : v1[2];
    float d;
    d = bx.max.x - bx.min.x; if (d < 0) d = -d;
    if (d < EPS) { bx.min.x -= EPS; bx.max.x += EPS; }
    d = bx.max.y - bx.min.y; if (d < 0) d = -d;
    if (d < EPS) { bx.min.y -= EPS; bx.max.y += EPS; }
    d = bx.max.z - bx.min.z; if (d < 0) d = -d;
    if (d < EPS) { bx.min.z -= EPS; bx.max.z += EPS; }
    return bx;
}

/* Record of triangle `tid` of an indexed mesh (nrm/uvs may be NULL). */
STM void st_tri_record(uint32_t tid, const float *pos, const uint32_t *pidx, const float *nrm, const uint32_t *nidx,
                       const float *uvs, const uint32_t *uidx, StTriangleRec *t) {
    memset(t, 0, sizeof *t);
    const float *a = pos + 3 * (size_t)pidx[3 * (size_t)tid];
    const float *bb = pos + 3 * (size_t)pidx[3 * (size_t)tid + 1];
    const float *c = pos + 3 * (size_t)pidx[3 * (size_t)tid + 2];
    sv3 v1 = V3(a[0], a[1], a[2]), v2 = V3(bb[0], bb[1], bb[2]), v3 = V3(c[0], c[1], c[2]);
    sm3 m;
    m.X = v_sub(v3, v1);
    m.Y = v_sub(v2, v1);
    m.Z = v_cross(m.X, m.Y);
    t->area_of_parallelogram = v_len(m.Z);
    m.Z = v_scale(m.Z, 1.0f / t->area_of_parallelogram);
    sm3 w = m_inverse(m);
    t->world_to_tangent[0] = w.X.x; t->world_to_tangent[1] = w.X.y; t->world_to_tangent[2] = w.X.z;
    t->world_to_tangent[3] = w.Y.x; t->world_to_tangent[4] = w.Y.y; t->world_to_tangent[5] = w.Y.z;
    t->world_to_tangent[6] = w.Z.x; t->world_to_tangent[7] = w.Z.y; t->world_to_tangent[8] = w.Z.z;
    t->normal[0] = m.Z.x; t->normal[1] = m.Z.y; t->normal[2] = m.Z.z;
    t->position[0] = v1.x; t->position[1] = v1.y; t->position[2] = v1.z;
    if (nrm) {
        for (int k = 0; k < 3; k++) {
            const float *n = nrm + 3 * (size_t)nidx[3 * (size_t)tid + k];
            t->vertex_normals[3 * k] = n[0];
            t->vertex_normals[3 * k + 1] = n[1];
            t->vertex_normals[3 * k + 2] = n[2];
        }
    }
    if (uvs) {
        const float *p0 = uvs + 2 * (size_t)uidx[3 * (size_t)tid];
        const float *p1 = uvs + 2 * (size_t)uidx[3 * (size_t)tid + 1];
        const float *p2 = uvs + 2 * (size_t)uidx[3 * (size_t)tid + 2];
        t->uvs[0] = p0[0]; t->uvs[1] = p0[1];
        t->uvs[2] = p1[0]; t->uvs[3] = p1[1];
        t->uvs[4] = p2[0]; t->uvs[5] = p2[1];
        t->area_of_uv = fabsf((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]));
    }
}

/* Split a Triangle record into the device's hot (intersection) and cold
   (hit attribute) records. */
STM void st_tri_split(const StTriangleRec *t, StDTriHot *h, StDTriCold *c) {
    memset(h, 0, sizeof *h);
    memset(c, 0, sizeof *c);
    for (int i = 0; i < 9; i++) h->w2t[i] = t->world_to_tangent[i];
    for (int i = 0; i < 3; i++) { h->pos[i] = t->position[i]; h->normal[i] = t->normal[i]; }
    for (int i = 0; i < 9; i++) c->vn[i] = t->vertex_normals[i];
    for (int i = 0; i < 6; i++) c->uv[i] = t->uvs[i];
    c->area = t->area_of_parallelogram;
    c->uv_area = t->area_of_uv;
}

/* Inverse of st_tri_split (device records back to the reference layout). */
STM void st_tri_join(const StDTriHot *h, const StDTriCold *c, StTriangleRec *t) {
    for (int i = 0; i < 9; i++) t->world_to_tangent[i] = h->w2t[i];
    for (int i = 0; i < 3; i++) { t->position[i] = h->pos[i]; t->normal[i] = h->normal[i]; }
    for (int i = 0; i < 9; i++) t->vertex_normals[i] = c->vn[i];
    for (int i = 0; i < 6; i++) t->uvs[i] = c->uv[i];
    t->area_of_parallelogram = c->area;
    t->area_of_uv = c->uv_area;
}

#endif
