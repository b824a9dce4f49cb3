/*
 * st_scene.c - scene blob -> device-ready StScene.
 *
 * Mirrors the reference's scene initialisation: primitives and materials get
 * their fields from the blob (defaults of initPrimitive / initMaterial are
 * applied by whoever writes the blob), mesh BVHs and triangle records come
 * from updateMeshBVH (or a prebuilt .mesh BVH), and the scene BVH from
 * updateSceneBVH (builder_top_down.h:366-381) unless the blob carries one.
 */
#include <stdlib.h>
#include <string.h>

#include "st_host.h"
#include "../st_mesh.h"

typedef struct { const uint8_t *p, *end; int bad; } Rd;

static const void *take(Rd *r, size_t n) {
    if (r->bad || (size_t)(r->end - r->p) < n) { r->bad = 1; return NULL; }
    const void *o = r->p;
    size_t adv = (n + 3) & ~(size_t)3;
    r->p = (size_t)(r->end - r->p) < adv ? r->end : r->p + adv;
    return o;
}

void st_scene_free(StScene *sc) {
    if (!sc) return;
    free(sc->prims); free(sc->mats); free(sc->lights);
    free(sc->scene_nodes); free(sc->scene_leaf_ids);
    free(sc->meshes); free(sc->mesh_aabbs); free(sc->mesh_nodes);
    free(sc->tri_hot); free(sc->tri_cold);
    free(sc->textures); free(sc->mips); free(sc->texels);
    free(sc->mesh_on_device); free(sc->deferred);
    free(sc);
}


/* Reject BVHs the device traversal could not walk safely: every inner
   node's two children must exist, no node may be reached twice from the root
   (a tree: no cycle, no shared subtree, so every descent ends), and leaves
   must index inside [0, leaf_limit). Returns the tree's real height or -1. */
int st_bvh_validate(const StBVHNodeRec *nodes, uint32_t count, uint64_t leaf_limit) {
    if (!count) return -1;
    uint8_t *seen = (uint8_t *)calloc(count, 1);
    uint32_t *stack = (uint32_t *)malloc(2 * sizeof(uint32_t) * ((size_t)count + 1));
    if (!seen || !stack) { free(seen); free(stack); return -1; }
    int max_depth = 0, ok = 1;
    size_t top = 0;
    stack[top++] = 0; stack[top++] = 1;
    seen[0] = 1;
    while (top && ok) {
        uint32_t depth = stack[--top];
        uint32_t n = stack[--top];
        if ((int)depth > max_depth) max_depth = (int)depth;
        const StBVHNodeRec *nd = nodes + n;
        if (nd->child_count) {
            if ((uint64_t)nd->first_child_id + nd->child_count > leaf_limit) ok = 0;
            continue;
        }
        uint32_t c = nd->first_child_id;
        if ((uint64_t)c + 1 >= count || seen[c] || seen[c + 1]) { ok = 0; break; }
        seen[c] = seen[c + 1] = 1;
        stack[top++] = c; stack[top++] = depth + 1;
        stack[top++] = c + 1; stack[top++] = depth + 1;
    }
    free(seen);
    free(stack);
    return ok ? max_depth : -1;
}

StDPrim st_prim_from_rec(const StPrimRec *r) {
    StDPrim p;
    memset(&p, 0, sizeof p);
    memcpy(p.rot, r->rotation, 16);
    memcpy(p.pos, r->position, 12);
    memcpy(p.scale, r->scale, 12);
    p.flags_mat = (r->flags & 0xFFu) | ((r->material_id & 0xFFu) << 8);
    p.type = r->type;
    /* oneOverVec3(scale), computed once: the same division the reference does per use */
    p.inv_scale[0] = 1.0f / r->scale[0];
    p.inv_scale[1] = 1.0f / r->scale[1];
    p.inv_scale[2] = 1.0f / r->scale[2];
    p.mesh_id = r->mesh_id;
    return p;
}

static sq4 prim_rot(const StDPrim *p) { sq4 q; q.v = V3(p->rot[0], p->rot[1], p->rot[2]); q.w = p->rot[3]; return q; }

/* scene/primitive.h:53-59 + render/AABB.h:44-84 */
sbox st_prim_world_box(const StScene *sc, uint32_t i) {
    const StDPrim *p = sc->prims + i;
    uint32_t flags = p->flags_mat & 0xFFu;
    sbox bx;
    if (p->type == ST_PRIM_MESH && p->mesh_id < sc->mesh_count) bx = sc->mesh_aabbs[p->mesh_id];
    else {
        bx.max = V3s(p->type == ST_PRIM_TETRAHEDRON ? (1.73205080757f / 3) : 1);
        bx.min = v_neg(bx.max);
    }
    float x = bx.min.x, y = bx.min.y, z = bx.min.z, X = bx.max.x, Y = bx.max.y, Z = bx.max.z;
    sv3 c[8] = {V3(x, y, z), V3(x, y, Z), V3(x, Y, z), V3(x, Y, Z), V3(X, y, z), V3(X, y, Z), V3(X, Y, z), V3(X, Y, Z)};
    x = y = z = +INFINITY;
    X = Y = Z = -INFINITY;
    sv3 scale = V3(p->scale[0], p->scale[1], p->scale[2]);
    sv3 pos = V3(p->pos[0], p->pos[1], p->pos[2]);
    sq4 rot = prim_rot(p);
    for (int k = 0; k < 8; k++) {
        sv3 q = c[k];
        if (flags & ST_IS_SCALED) q = v_mul(q, scale);
        if (flags & ST_IS_ROTATED) q = q_rot(q, rot);
        if (flags & ST_IS_TRANSLATED) q = v_add(q, pos);
        if (q.x > X) X = q.x;
        if (q.y > Y) Y = q.y;
        if (q.z > Z) Z = q.z;
        if (q.x < x) x = q.x;
        if (q.y < y) y = q.y;
        if (q.z < z) z = q.z;
    }
    bx.min = V3(x, y, z);
    bx.max = V3(X, Y, Z);
    return bx;
}

int st_scene_rebuild_sah(StScene *sc) {
    uint32_t P = sc->prim_count;
    if (!P) return 0;
    sbox *boxes = (sbox *)malloc(P * sizeof(sbox));
    if (!boxes) return -1;
    for (uint32_t i = 0; i < P; i++) boxes[i] = st_prim_world_box(sc, i);
    sc->scene_node_count = st_sah_build(boxes, P, 2, sc->scene_nodes, sc->scene_leaf_ids, &sc->scene_height);
    free(boxes);
    return 0;
}

StDMat st_mat_from_rec(const StMaterialRec *q) {
    StDMat m;
    memset(&m, 0, sizeof m);
    memcpy(m.albedo, q->albedo, 12);
    memcpy(m.reflectivity, q->reflectivity, 12);
    memcpy(m.emission, q->emission, 12);
    memcpy(m.uv_repeat, q->uv_repeat, 8);
    m.metallic = q->metallic;
    m.roughness = q->roughness;
    m.n1_over_n2 = q->n1_over_n2;
    m.n2_over_n1 = q->n2_over_n1;
    m.brdf = q->brdf;
    m.bits = (uint32_t)q->is | ((uint32_t)q->use << 8) | ((uint32_t)q->texture_count << 16);
    m.tex0 = q->texture_ids[0];
    m.tex1 = q->texture_ids[1];
    return m;
}

#define FAIL(msg) do { *err = (msg); st_scene_free(sc); return NULL; } while (0)

StScene *st_scene_from_parts(const StSceneParts *pp, const char **err) {
    const char *dummy;
    if (!err) err = &dummy;
    *err = NULL;
    StScene *sc = (StScene *)calloc(1, sizeof(StScene));
    if (!sc) { *err = "out of host memory"; return NULL; }
    const uint32_t P = pp->prim_count;
    sc->prim_count = P;
    sc->material_count = pp->material_count;
    sc->light_count = pp->light_count;
    sc->mesh_count = pp->mesh_count;
    sc->texture_count = pp->texture_count;
    memcpy(sc->ambient, pp->ambient, 12);
    if ((P && !pp->prims) || (pp->material_count && !pp->mats) || (pp->light_count && !pp->lights) ||
        (pp->mesh_count && !pp->meshes) || (pp->texture_count && !pp->textures))
        FAIL("missing scene array");

    sc->prims = (StDPrim *)calloc((size_t)P + 1, sizeof(StDPrim));
    sc->mats = (StDMat *)calloc((size_t)pp->material_count + 1, sizeof(StDMat));
    sc->lights = (StDLight *)calloc((size_t)pp->light_count + 1, sizeof(StDLight));
    sc->meshes = (StDMesh *)calloc((size_t)pp->mesh_count + 1, sizeof(StDMesh));
    sc->mesh_aabbs = (sbox *)calloc((size_t)pp->mesh_count + 1, sizeof(sbox));
    sc->textures = (StDTex *)calloc((size_t)pp->texture_count + 1, sizeof(StDTex));
    if (!sc->prims || !sc->mats || !sc->lights || !sc->meshes || !sc->mesh_aabbs || !sc->textures)
        FAIL("out of host memory");
    for (uint32_t i = 0; i < P; i++) {
        const StPrimRec *r = pp->prims + i;
        if (r->material_id >= pp->material_count || r->material_id > 255) FAIL("primitive material_id out of range");
        if (r->type == ST_PRIM_MESH && r->mesh_id >= pp->mesh_count) FAIL("primitive mesh_id out of range");
        sc->prims[i] = st_prim_from_rec(r);
    }
    for (uint32_t i = 0; i < pp->material_count; i++) {
        const StMaterialRec *q = pp->mats + i;
        sc->mats[i] = st_mat_from_rec(q);
        if (q->texture_count && q->use) {
            uint32_t need = q->texture_count > 1 ? 2 : 1;
            for (uint32_t t = 0; t < need; t++)
                if (q->texture_ids[t] >= pp->texture_count) FAIL("material texture id out of range");
        }
    }
    for (uint32_t i = 0; i < pp->light_count; i++) {
        memcpy(sc->lights[i].pos, pp->lights[i].position, 12);
        sc->lights[i].intensity = pp->lights[i].intensity;
        memcpy(sc->lights[i].color, pp->lights[i].color, 12);
    }

    /* meshes: offsets first, so node/triangle totals are known */
    uint64_t node_total = 1, tri_total = 0;
    for (uint32_t m = 0; m < pp->mesh_count; m++) {
        const StMeshPart *s = pp->meshes + m;
        const StMeshRec *q = &s->rec;
        if (!q->triangle_count) FAIL("mesh without triangles");
        if (q->prebuilt) {
            if (!s->nodes || !s->tris || !q->node_count) FAIL("prebuilt mesh without BVH or triangles");
        } else if (!s->pos || !s->pidx) {
            FAIL("mesh without geometry");
        }
        if (s->pidx)
            for (uint64_t t = 0; t < 3ull * q->triangle_count; t++)
                if (s->pidx[t] >= q->vertex_count || (s->nidx && s->nidx[t] >= q->normals_count) ||
                    (s->uidx && s->uidx[t] >= q->uvs_count))
                    FAIL("mesh index out of range");
        uint64_t nodes = q->prebuilt ? q->node_count : 2ull * q->triangle_count;
        if (!(node_total & 1)) node_total++;   /* odd offset: child pairs start 64-byte aligned */
        sc->meshes[m].node_offset = (uint32_t)node_total;
        node_total += nodes;
        sc->meshes[m].tri_offset = (uint32_t)tri_total;
        sc->meshes[m].triangle_count = q->triangle_count;
        sc->meshes[m].attrs = (q->normals_count ? 1u : 0u) | (q->uvs_count ? 2u : 0u);
        tri_total += q->triangle_count;
        sc->mesh_aabbs[m].min = V3(q->aabb[0], q->aabb[1], q->aabb[2]);
        sc->mesh_aabbs[m].max = V3(q->aabb[3], q->aabb[4], q->aabb[5]);
        if (node_total > 0xFFFFFFF0ull || tri_total > 0xFFFFFFF0ull) FAIL("scene too large");
    }
    sc->mesh_node_total = (uint32_t)node_total;
    sc->tri_total = (uint32_t)tri_total;
    sc->mesh_on_device = (uint8_t *)calloc((size_t)pp->mesh_count + 1, 1);
    sc->deferred = (StMeshPart *)calloc((size_t)pp->mesh_count + 1, sizeof(StMeshPart));
    if (!sc->mesh_on_device || !sc->deferred) FAIL("out of host memory");
    /* host copies of the mesh records only when some mesh is built or
       loaded on the host; device-built meshes never have one */
    sc->mesh_host_data = pp->mesh_count ? 0u : 1u;
    for (uint32_t m = 0; m < pp->mesh_count; m++)
        if (pp->meshes[m].rec.prebuilt || pp->mesh_builder != ST_MESH_BUILD_DEVICE) sc->mesh_host_data = 1;
    sc->mesh_nodes = (StBVHNodeRec *)calloc(sc->mesh_host_data ? node_total + 1 : 1, sizeof(StBVHNodeRec));
    sc->tri_hot = (StDTriHot *)calloc(sc->mesh_host_data ? tri_total + 1 : 1, sizeof(StDTriHot));
    sc->tri_cold = (StDTriCold *)calloc(sc->mesh_host_data ? tri_total + 1 : 1, sizeof(StDTriCold));
    if (!sc->mesh_nodes || !sc->tri_hot || !sc->tri_cold) FAIL("out of host memory");
    for (uint32_t m = 0; m < pp->mesh_count; m++) {
        const StMeshPart *s = pp->meshes + m;
        const StMeshRec *q = &s->rec;
        StDMesh *dm = sc->meshes + m;
        uint32_t height = 0, count = 0;
        if (!q->prebuilt && pp->mesh_builder == ST_MESH_BUILD_DEVICE) {
            sc->mesh_on_device[m] = 1;
            sc->deferred[m] = *s;
            dm->height = 0;
            continue;
        }
        if (q->prebuilt) {
            memcpy(sc->mesh_nodes + dm->node_offset, s->nodes, sizeof(StBVHNodeRec) * q->node_count);
            count = q->node_count;
            height = q->height;
            for (uint32_t t = 0; t < q->triangle_count; t++)
                st_tri_split(s->tris + t, sc->tri_hot + dm->tri_offset + t, sc->tri_cold + dm->tri_offset + t);
        } else {
            StTriangleRec *tris = (StTriangleRec *)malloc(sizeof(StTriangleRec) * q->triangle_count);
            uint32_t *leaf = (uint32_t *)malloc(4u * (size_t)q->triangle_count);
            int (*build)(uint32_t, const float *, const uint32_t *, const float *, const uint32_t *, uint32_t,
                         const float *, const uint32_t *, uint32_t, StBVHNodeRec *, uint32_t *, uint32_t *,
                         StTriangleRec *, uint32_t *) =
                pp->mesh_builder == ST_MESH_BUILD_BINNED ? st_build_mesh_binned : st_build_mesh;
            if (!tris || !leaf || build(q->triangle_count, s->pos, s->pidx, s->nrm, s->nidx, q->normals_count, s->uvs,
                                        s->uidx, q->uvs_count, sc->mesh_nodes + dm->node_offset, &count, &height,
                                        tris, leaf)) {
                free(tris); free(leaf);
                FAIL("mesh BVH build failed");
            }
            for (uint32_t t = 0; t < q->triangle_count; t++)
                st_tri_split(tris + t, sc->tri_hot + dm->tri_offset + t, sc->tri_cold + dm->tri_offset + t);
            free(tris);
            free(leaf);
        }
        dm->height = height;
        if (height > sc->max_mesh_height) sc->max_mesh_height = height;
        if (st_bvh_validate(sc->mesh_nodes + dm->node_offset, count, q->triangle_count) < 0) FAIL("malformed mesh BVH");
    }
    if (pp->mesh_count) {
        sc->mesh_stack_size = (uint8_t)(sc->max_mesh_height + 2); /* core/init.h:523-527 */
        if (sc->max_mesh_height + 2 > ST_MESH_STACK_MAX) FAIL("mesh BVH deeper than the device traversal stack");
    }

    /* textures */
    uint64_t quads = 0, mips = 0;
    for (uint32_t t = 0; t < pp->texture_count; t++) {
        const StTexturePart *q = pp->textures + t;
        if (!q->rec.mip_count || !q->mips) FAIL("texture without mips");
        for (uint32_t i = 0; i < q->rec.mip_count; i++) {
            if (!q->mips[i].quads) FAIL("texture mip without texels");
            quads += (uint64_t)(q->mips[i].width + 1) * (q->mips[i].height + 1);
            mips++;
        }
    }
    sc->mips = (StDMip *)calloc(mips + 1, sizeof(StDMip));
    sc->texel_bytes = 12 * quads;
    sc->texels = (uint8_t *)malloc(sc->texel_bytes + 16);
    if (!sc->mips || !sc->texels) FAIL("out of host memory");
    sc->mip_total = (uint32_t)mips;
    uint64_t qoff = 0;
    uint32_t mip_index = 0;
    for (uint32_t t = 0; t < pp->texture_count; t++) {
        const StTexturePart *q = pp->textures + t;
        StDTex *dt = sc->textures + t;
        dt->width = q->rec.width;
        dt->height = q->rec.height;
        dt->mip_count = q->rec.mip_count;
        dt->mipmap = q->rec.mipmap;
        dt->mip_base = mip_index;
        for (uint32_t i = 0; i < q->rec.mip_count; i++) {
            uint64_t nq = (uint64_t)(q->mips[i].width + 1) * (q->mips[i].height + 1);
            sc->mips[mip_index].width = q->mips[i].width;
            sc->mips[mip_index].height = q->mips[i].height;
            sc->mips[mip_index].quad_offset = qoff;
            memcpy(sc->texels + 12 * qoff, q->mips[i].quads, 12 * nq);
            qoff += nq;
            mip_index++;
        }
    }

    /* scene BVH */
    sc->scene_nodes = (StBVHNodeRec *)calloc(2 * (size_t)P + 2, sizeof(StBVHNodeRec));
    sc->scene_leaf_ids = (uint32_t *)calloc((size_t)P + 1, sizeof(uint32_t));
    if (!sc->scene_nodes || !sc->scene_leaf_ids) FAIL("out of host memory");
    if (pp->scene_nodes) {
        if (!pp->scene_leaf_ids || pp->scene_node_count > 2 * (size_t)P + 1) FAIL("bad scene BVH");
        memcpy(sc->scene_nodes, pp->scene_nodes, sizeof(StBVHNodeRec) * pp->scene_node_count);
        memcpy(sc->scene_leaf_ids, pp->scene_leaf_ids, 4u * P);
        sc->scene_node_count = pp->scene_node_count;
        sc->scene_height = pp->scene_height;
        for (uint32_t i = 0; i < P; i++)
            if (pp->scene_leaf_ids[i] >= P) FAIL("scene BVH leaf id out of range");
    } else if (P) {
        if (st_scene_rebuild_sah(sc)) FAIL("scene BVH build failed");
    }
    if (P) {
        int real_height = st_bvh_validate(sc->scene_nodes, sc->scene_node_count, P);
        if (real_height < 0) FAIL("malformed scene BVH");
        sc->scene_real_height = (uint32_t)real_height;
        if (real_height + 1 > ST_SCENE_STACK_MAX) FAIL("scene BVH deeper than the device traversal stack");
    }
    for (uint32_t i = 0; i < P; i++) {
        uint32_t mid = (sc->prims[i].flags_mat >> 8) & 0xFFu;
        if (sc->prims[i].type == ST_PRIM_QUAD && (sc->mats[mid].bits & ST_EMISSIVE)) sc->has_emissive_quads = 1;
    }
    return sc;
}

#undef FAIL

StScene *st_scene_from_blob(const void *blob, size_t size, const char **err) {
    return st_scene_from_blob_ex(blob, size, 0, err);
}

StScene *st_scene_from_blob_ex(const void *blob, size_t size, uint32_t mesh_builder, const char **err) {
    const char *dummy;
    if (!err) err = &dummy;
    *err = NULL;
    Rd r = {(const uint8_t *)blob, (const uint8_t *)blob + size, 0};
    const StBlobHeader *h = (const StBlobHeader *)take(&r, sizeof *h);
    if (!h || h->magic != ST_BLOB_MAGIC || h->version != ST_BLOB_VERSION) { *err = "bad blob header"; return NULL; }
    StSceneParts pp;
    memset(&pp, 0, sizeof pp);
    pp.prim_count = h->prim_count;
    pp.material_count = h->material_count;
    pp.light_count = h->light_count;
    pp.mesh_count = h->mesh_count;
    pp.texture_count = h->texture_count;
    memcpy(pp.ambient, h->ambient, 12);
    pp.prims = (const StPrimRec *)take(&r, sizeof(StPrimRec) * h->prim_count);
    pp.mats = (const StMaterialRec *)take(&r, sizeof(StMaterialRec) * h->material_count);
    pp.lights = (const StLightRec *)take(&r, sizeof(StLightRec) * h->light_count);
    if (r.bad) { *err = "truncated blob"; return NULL; }

    StMeshPart *meshes = (StMeshPart *)calloc((size_t)h->mesh_count + 1, sizeof(StMeshPart));
    StTexturePart *texs = (StTexturePart *)calloc((size_t)h->texture_count + 1, sizeof(StTexturePart));
    StMipPart **mip_arrays = (StMipPart **)calloc((size_t)h->texture_count + 1, sizeof(StMipPart *));
    StScene *sc = NULL;
    if (!meshes || !texs || !mip_arrays) { *err = "out of host memory"; goto out; }
    for (uint32_t m = 0; m < h->mesh_count; m++) {
        StMeshPart *s = meshes + m;
        const StMeshRec *q = (const StMeshRec *)take(&r, sizeof(StMeshRec));
        if (!q) break;
        s->rec = *q;
        s->pos = (const float *)take(&r, 12u * (size_t)q->vertex_count);
        s->pidx = (const uint32_t *)take(&r, 12u * (size_t)q->triangle_count);
        if (q->normals_count) {
            s->nrm = (const float *)take(&r, 12u * (size_t)q->normals_count);
            s->nidx = (const uint32_t *)take(&r, 12u * (size_t)q->triangle_count);
        }
        if (q->uvs_count) {
            s->uvs = (const float *)take(&r, 8u * (size_t)q->uvs_count);
            s->uidx = (const uint32_t *)take(&r, 12u * (size_t)q->triangle_count);
        }
        if (q->prebuilt) {
            s->nodes = (const StBVHNodeRec *)take(&r, sizeof(StBVHNodeRec) * (size_t)q->node_count);
            s->tris = (const StTriangleRec *)take(&r, sizeof(StTriangleRec) * (size_t)q->triangle_count);
            take(&r, 4u * (size_t)q->triangle_count);
        }
        if (r.bad) break;
    }
    if (r.bad) { *err = "truncated mesh section"; goto out; }
    for (uint32_t t = 0; t < h->texture_count && !r.bad; t++) {
        const StTextureRec *q = (const StTextureRec *)take(&r, sizeof *q);
        if (!q) break;
        texs[t].rec = *q;
        if (!q->mip_count) { r.bad = 1; break; }
        StMipPart *mp = (StMipPart *)calloc(q->mip_count, sizeof(StMipPart));
        if (!mp) { *err = "out of host memory"; goto out; }
        mip_arrays[t] = mp;
        texs[t].mips = mp;
        for (uint32_t i = 0; i < q->mip_count; i++) {
            const uint16_t *d = (const uint16_t *)take(&r, 4);
            if (!d) break;
            mp[i].width = d[0];
            mp[i].height = d[1];
            mp[i].quads = (const uint8_t *)take(&r, 12 * (size_t)(d[0] + 1) * (d[1] + 1));
            if (!mp[i].quads) break;
        }
    }
    if (r.bad) { *err = "truncated texture section"; goto out; }
    pp.meshes = meshes;
    pp.textures = texs;
    if (h->flags & ST_BLOB_HAS_SCENE_BVH) {
        const uint32_t *cnt = (const uint32_t *)take(&r, 8);
        if (!cnt || cnt[0] > 2 * (size_t)h->prim_count + 1) { *err = "bad scene BVH section"; goto out; }
        pp.scene_node_count = cnt[0];
        pp.scene_height = cnt[1];
        pp.scene_nodes = (const StBVHNodeRec *)take(&r, sizeof(StBVHNodeRec) * cnt[0]);
        pp.scene_leaf_ids = (const uint32_t *)take(&r, 4u * (size_t)h->prim_count);
        if (r.bad) { *err = "truncated scene BVH"; goto out; }
    }
    pp.mesh_builder = mesh_builder;
    sc = st_scene_from_parts(&pp, err);
out:
    for (uint32_t t = 0; t < h->texture_count; t++) free(mip_arrays ? mip_arrays[t] : NULL);
    free(mip_arrays);
    free(texs);
    free(meshes);
    return sc;
}

uint64_t st_scene_device_bytes(const StScene *sc) {
    uint64_t b = 0;
    b += (uint64_t)(sc->prim_count + 1) * sizeof(StDPrim);
    b += (uint64_t)(sc->material_count + 1) * sizeof(StDMat);
    b += (uint64_t)(sc->light_count + 1) * sizeof(StDLight);
    b += (uint64_t)(2 * sc->prim_count + 4) * sizeof(StBVHNodeRec) + 4ull * (sc->prim_count + 1);
    b += (uint64_t)(sc->mesh_count + 1) * sizeof(StDMesh);
    b += (uint64_t)(sc->mesh_node_total + 2) * sizeof(StBVHNodeRec);
    b += (uint64_t)(sc->tri_total + 1) * (sizeof(StDTriHot) + sizeof(StDTriCold));
    b += (uint64_t)(sc->texture_count + 1) * sizeof(StDTex) + (uint64_t)(sc->mip_total + 1) * sizeof(StDMip);
    b += sc->texel_bytes + 16;
    return b;
}
