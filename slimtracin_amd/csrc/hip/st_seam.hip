/*
 * st_seam.hip - the reference's device-mirror seam (core/init.h:36-126,
 * render/raytracer.h:258-299) over the library's resident scenes.
 *
 * The caller uploads arrays the way the reference's CUDA build does; this
 * file keeps host copies of what was uploaded and brings the resident scene
 * up to date before a frame:
 *   - the first frame, and any frame after meshes or textures were uploaded,
 *     builds the device-ready scene from the copies and copies it to HBM;
 *   - primitives, materials, lights and the scene BVH uploaded later go
 *     straight to their device arrays (uploadPrimitives every frame is the
 *     reference's animation pattern, 02_Geometry.c:70-75).
 * A primitive upload without a scene BVH upload leaves the old tree in
 * place, as in the reference. If no scene BVH was ever uploaded, the scene
 * BVH is the SAH build over the primitives (updateSceneBVH).
 */
#include <hip/hip_runtime.h>

#include <stdlib.h>
#include <string.h>

#include <vector>

#include "st_internal.h"
#include "slimtracin_seam.h"

struct SeamMesh {
    StSeamMesh m;
    std::vector<StBVHNodeRec> nodes;
    std::vector<StTriangleRec> tris;
};

struct SeamTexture {
    StTextureRec rec;
    std::vector<StMipPart> mips;
    std::vector<std::vector<uint8_t>> texels;
};

struct StSeamState {
    StSeamSizes sz;
    std::vector<StPrimRec> prims;
    std::vector<int32_t> bounds;
    std::vector<StMaterialRec> mats;
    std::vector<StLightRec> lights;
    std::vector<SeamMesh> meshes;
    std::vector<SeamTexture> textures;
    std::vector<StBVHNodeRec> scene_nodes;
    std::vector<uint32_t> scene_leaf;
    uint32_t scene_height = 0;
    float ambient[3] = {0, 0, 0};
    bool have_prims = false, have_mats = false, have_lights = false, have_meshes = false, have_textures = false;
    bool have_scene_bvh = false;
    bool dirty_full = true, dirty_prims = false, dirty_mats = false, dirty_lights = false, dirty_scene_bvh = false;
};

void st_seam_state_free(StSeamState *s) { delete s; }

static int fail(int code, const char *msg) { return st_rt_error(code, msg); }

#define HIP_OK(call)                                                                \
    do {                                                                            \
        if ((call) != hipSuccess) return fail(ST_ERR_HIP, #call " failed");         \
    } while (0)

extern "C" int st_seam_allocate(const StSeamSizes *sizes, StSceneHandle **out) {
    if (!sizes || !out) return fail(ST_ERR_ARG, "null argument");
    *out = NULL;
    int rc = st_rt_ensure_init();
    if (rc) return rc;
    StSceneHandle *h = (StSceneHandle *)calloc(1, sizeof(StSceneHandle));
    if (!h) return fail(ST_ERR_MEMORY, "out of host memory");
    h->seam = new StSeamState();
    h->seam->sz = *sizes;
    *out = h;
    return ST_OK;
}

static StSeamState *state(StSceneHandle *h) { return h ? h->seam : NULL; }

extern "C" int st_seam_upload_primitives(StSceneHandle *h, const StPrimRec *prims, const int32_t *screen_bounds) {
    StSeamState *s = state(h);
    if (!s) return fail(ST_ERR_ARG, "not a seam scene");
    uint32_t P = s->sz.primitives;
    if (P && (!prims || !screen_bounds)) return fail(ST_ERR_ARG, "null primitives or screen bounds");
    for (uint32_t i = 0; i < P; i++) {
        if (prims[i].material_id >= s->sz.materials || prims[i].material_id > 255)
            return fail(ST_ERR_ARG, "primitive material_id out of range");
        if (prims[i].type == ST_PRIM_MESH && prims[i].mesh_id >= s->sz.meshes)
            return fail(ST_ERR_ARG, "primitive mesh_id out of range");
    }
    s->prims.assign(prims, prims + P);
    s->bounds.assign(screen_bounds, screen_bounds + 4 * (size_t)P);
    s->have_prims = true;
    s->dirty_prims = true;
    return ST_OK;
}

extern "C" int st_seam_upload_materials(StSceneHandle *h, const StMaterialRec *materials) {
    StSeamState *s = state(h);
    if (!s) return fail(ST_ERR_ARG, "not a seam scene");
    if (s->sz.materials && !materials) return fail(ST_ERR_ARG, "null materials");
    for (uint32_t i = 0; i < s->sz.materials; i++)
        if (materials[i].texture_count && materials[i].use) {
            uint32_t need = materials[i].texture_count > 1 ? 2 : 1;
            for (uint32_t t = 0; t < need; t++)
                if (materials[i].texture_ids[t] >= s->sz.textures) return fail(ST_ERR_ARG, "material texture id out of range");
        }
    s->mats.assign(materials, materials + s->sz.materials);
    s->have_mats = true;
    s->dirty_mats = true;
    return ST_OK;
}

extern "C" int st_seam_upload_lights(StSceneHandle *h, const StLightRec *lights) {
    StSeamState *s = state(h);
    if (!s) return fail(ST_ERR_ARG, "not a seam scene");
    if (s->sz.lights && !lights) return fail(ST_ERR_ARG, "null lights");
    s->lights.assign(lights, lights + s->sz.lights);
    s->have_lights = true;
    s->dirty_lights = true;
    return ST_OK;
}

extern "C" int st_seam_upload_scene(StSceneHandle *h, const StLightRec *lights, const StPrimRec *prims,
                                    const int32_t *screen_bounds, const StMaterialRec *materials,
                                    const StSeamTexture *textures) {
    StSeamState *s = state(h);
    if (!s) return fail(ST_ERR_ARG, "not a seam scene");
    if (s->sz.textures && !textures) return fail(ST_ERR_ARG, "null textures");
    std::vector<SeamTexture> texs(s->sz.textures);
    for (uint32_t t = 0; t < s->sz.textures; t++) {
        const StSeamTexture *q = textures + t;
        if (!q->mip_count || !q->mips) return fail(ST_ERR_ARG, "texture without mips");
        SeamTexture &d = texs[t];
        d.rec.width = q->width;
        d.rec.height = q->height;
        d.rec.mip_count = q->mip_count;
        d.rec.wrap = q->wrap;
        d.rec.mipmap = q->mipmap;
        d.rec.pad = 0;
        d.mips.resize(q->mip_count);
        d.texels.resize(q->mip_count);
        for (uint32_t i = 0; i < q->mip_count; i++) {
            const StSeamMip *m = q->mips + i;
            if (!m->texel_quads) return fail(ST_ERR_ARG, "texture mip without texels");
            size_t bytes = 12 * (size_t)(m->width + 1) * (m->height + 1);
            d.texels[i].assign(m->texel_quads, m->texel_quads + bytes);
            d.mips[i].width = m->width;
            d.mips[i].height = m->height;
            d.mips[i].quads = d.texels[i].data();
        }
    }
    int rc;
    if ((rc = st_seam_upload_lights(h, lights))) return rc;
    if ((rc = st_seam_upload_primitives(h, prims, screen_bounds))) return rc;
    if ((rc = st_seam_upload_materials(h, materials))) return rc;
    s->textures.swap(texs);
    s->have_textures = true;
    s->dirty_full = true;
    return ST_OK;
}

extern "C" int st_seam_upload_scene_bvh(StSceneHandle *h, const StBVHNodeRec *nodes, uint32_t node_count,
                                        uint32_t height, const uint32_t *leaf_ids) {
    StSeamState *s = state(h);
    if (!s) return fail(ST_ERR_ARG, "not a seam scene");
    uint32_t P = s->sz.primitives;
    if (!P) return ST_OK;
    if (!nodes || !leaf_ids || !node_count || node_count > 2 * P) return fail(ST_ERR_ARG, "bad scene BVH arguments");
    if (st_bvh_validate(nodes, node_count, P) < 0) return fail(ST_ERR_SCENE, "malformed scene BVH");
    for (uint32_t i = 0; i < P; i++)
        if (leaf_ids[i] >= P) return fail(ST_ERR_SCENE, "scene BVH leaf id out of range");
    s->scene_nodes.assign(nodes, nodes + node_count);
    s->scene_leaf.assign(leaf_ids, leaf_ids + P);
    s->scene_height = height;
    s->have_scene_bvh = true;
    s->dirty_scene_bvh = true;
    return ST_OK;
}

extern "C" int st_seam_upload_mesh_bvhs(StSceneHandle *h, const StSeamMesh *meshes) {
    StSeamState *s = state(h);
    if (!s) return fail(ST_ERR_ARG, "not a seam scene");
    if (s->sz.meshes && !meshes) return fail(ST_ERR_ARG, "null meshes");
    std::vector<SeamMesh> ms(s->sz.meshes);
    for (uint32_t m = 0; m < s->sz.meshes; m++) {
        const StSeamMesh *q = meshes + m;
        if (!q->triangle_count || !q->node_count || !q->nodes || !q->triangles)
            return fail(ST_ERR_ARG, "mesh without BVH or triangles");
        if (q->node_count > 2ull * q->triangle_count) return fail(ST_ERR_ARG, "mesh BVH larger than 2 nodes per triangle");
        ms[m].m = *q;
        ms[m].nodes.assign(q->nodes, q->nodes + q->node_count);
        ms[m].tris.assign(q->triangles, q->triangles + q->triangle_count);
        ms[m].m.leaf_ids = NULL;
    }
    s->meshes.swap(ms);
    s->have_meshes = true;
    s->dirty_full = true;
    return ST_OK;
}

/* Full build of the resident scene from the host copies. */
static int seam_build(StSceneHandle *h) {
    StSeamState *s = h->seam;
    const StSeamSizes &z = s->sz;
    if ((z.primitives && !s->have_prims) || (z.materials && !s->have_mats) || (z.lights && !s->have_lights) ||
        (z.meshes && !s->have_meshes) || (z.textures && !s->have_textures))
        return fail(ST_ERR_ARG, "scene arrays missing: upload the scene, its mesh BVHs and primitives first");
    std::vector<StMeshPart> mparts(z.meshes + 1);
    for (uint32_t m = 0; m < z.meshes; m++) {
        const SeamMesh &q = s->meshes[m];
        StMeshPart &p = mparts[m];
        memset(&p, 0, sizeof p);
        p.rec.triangle_count = q.m.triangle_count;
        p.rec.normals_count = q.m.normals_count;
        p.rec.uvs_count = q.m.uvs_count;
        p.rec.prebuilt = 1;
        p.rec.node_count = q.m.node_count;
        p.rec.height = q.m.height;
        memcpy(p.rec.aabb, q.m.aabb, 24);
        p.nodes = q.nodes.data();
        p.tris = q.tris.data();
    }
    std::vector<StTexturePart> tparts(z.textures + 1);
    for (uint32_t t = 0; t < z.textures; t++) {
        tparts[t].rec = s->textures[t].rec;
        tparts[t].mips = s->textures[t].mips.data();
    }
    StSceneParts pp;
    memset(&pp, 0, sizeof pp);
    pp.prim_count = z.primitives;
    pp.material_count = z.materials;
    pp.light_count = z.lights;
    pp.mesh_count = z.meshes;
    pp.texture_count = z.textures;
    memcpy(pp.ambient, s->ambient, 12);
    pp.prims = s->prims.data();
    pp.mats = s->mats.data();
    pp.lights = s->lights.data();
    pp.meshes = mparts.data();
    pp.textures = tparts.data();
    if (s->have_scene_bvh) {
        pp.scene_nodes = s->scene_nodes.data();
        pp.scene_node_count = (uint32_t)s->scene_nodes.size();
        pp.scene_height = s->scene_height;
        pp.scene_leaf_ids = s->scene_leaf.data();
    }
    const char *why = NULL;
    StScene *sc = st_scene_from_parts(&pp, &why);
    if (!sc) return fail(ST_ERR_SCENE, why ? why : "invalid scene");
    StScene *old = h->host;
    h->host = sc;
    int rc = st_scene_upload(h);
    if (rc) {
        h->host = old;
        st_scene_free(sc);
        return rc;
    }
    st_scene_free(old);
    s->dirty_full = s->dirty_prims = s->dirty_mats = s->dirty_lights = s->dirty_scene_bvh = false;
    return ST_OK;
}

static void refresh_emissive(StSceneHandle *h) {
    StScene *sc = h->host;
    sc->has_emissive_quads = 0;
    for (uint32_t i = 0; i < sc->prim_count; i++) {
        uint32_t mid = (sc->prims[i].flags_mat >> 8) & 0xFFu;
        if (sc->prims[i].type == ST_PRIM_QUAD && (sc->mats[mid].bits & ST_EMISSIVE)) sc->has_emissive_quads = 1;
    }
    h->dev.has_emissive_quads = sc->has_emissive_quads;
}

namespace {
struct RtLock {
    RtLock() { st_rt_lock(); }
    ~RtLock() { st_rt_unlock(); }
};
}  // namespace

extern "C" int st_seam_sync(StSceneHandle *h) {
    RtLock lock;
    StSeamState *s = state(h);
    if (!s) return fail(ST_ERR_ARG, "not a seam scene");
    if (!h->host || !h->dbuf || s->dirty_full) return seam_build(h);
    StScene *sc = h->host;
    hipStream_t stream = st_rt_stream();
    if (s->dirty_prims) {
        for (uint32_t i = 0; i < sc->prim_count; i++) sc->prims[i] = st_prim_from_rec(&s->prims[i]);
        if (sc->prim_count)
            HIP_OK(hipMemcpyAsync((void *)h->dev.prims, sc->prims, sc->prim_count * sizeof(StDPrim),
                                  hipMemcpyHostToDevice, stream));
    }
    if (s->dirty_mats) {
        for (uint32_t i = 0; i < sc->material_count; i++) sc->mats[i] = st_mat_from_rec(&s->mats[i]);
        if (sc->material_count)
            HIP_OK(hipMemcpyAsync((void *)h->dev.mats, sc->mats, sc->material_count * sizeof(StDMat),
                                  hipMemcpyHostToDevice, stream));
    }
    if (s->dirty_prims || s->dirty_mats) refresh_emissive(h);
    if (s->dirty_lights) {
        for (uint32_t i = 0; i < sc->light_count; i++) {
            memcpy(sc->lights[i].pos, s->lights[i].position, 12);
            sc->lights[i].intensity = s->lights[i].intensity;
            memcpy(sc->lights[i].color, s->lights[i].color, 12);
        }
        if (sc->light_count)
            HIP_OK(hipMemcpyAsync((void *)h->dev.lights, sc->lights, sc->light_count * sizeof(StDLight),
                                  hipMemcpyHostToDevice, stream));
    }
    if (s->dirty_scene_bvh && sc->prim_count) {
        int height = st_bvh_validate(s->scene_nodes.data(), (uint32_t)s->scene_nodes.size(), sc->prim_count);
        if (height < 0 || height + 1 > ST_SCENE_STACK_MAX) return fail(ST_ERR_SCENE, "scene BVH too deep");
        memcpy(sc->scene_nodes, s->scene_nodes.data(), s->scene_nodes.size() * sizeof(StBVHNodeRec));
        memcpy(sc->scene_leaf_ids, s->scene_leaf.data(), 4u * sc->prim_count);
        sc->scene_node_count = (uint32_t)s->scene_nodes.size();
        sc->scene_height = s->scene_height;
        sc->scene_real_height = (uint32_t)height;
        HIP_OK(hipMemcpyAsync((void *)h->dev.scene_nodes, sc->scene_nodes, sc->scene_node_count * sizeof(StBVHNodeRec),
                              hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemcpyAsync((void *)h->dev.scene_leaf_ids, sc->scene_leaf_ids, 4u * sc->prim_count,
                              hipMemcpyHostToDevice, stream));
        h->dev.scene_stack_need = sc->scene_real_height + 1;
        h->bvh_on_device_only = 0;
    }
    HIP_OK(hipStreamSynchronize(stream));
    s->dirty_prims = s->dirty_mats = s->dirty_lights = s->dirty_scene_bvh = false;
    return ST_OK;
}

extern "C" int st_seam_render(StSceneHandle *h, const StSeamFrame *frame, StFloatPixel *pixels, StHitRec *hits,
                              StRayCounts *counts) {
    RtLock lock;
    StSeamState *s = state(h);
    if (!s || !frame) return fail(ST_ERR_ARG, "not a seam scene or null frame");
    memcpy(s->ambient, frame->ambient, 12);
    int rc = st_seam_sync(h);
    if (rc) return rc;
    memcpy(h->host->ambient, frame->ambient, 12);
    memcpy(h->dev.ambient, frame->ambient, 12);
    StHostView hv;
    if (st_view_prepare_frame(h->host, frame, s->bounds.data(), &hv))
        return fail(ST_ERR_ARG, "invalid frame (size or render mode)");
    rc = st_rt_launch_prepared(&h, &hv, 1, hits != NULL, 1, counts, &pixels, &hits, NULL, NULL);
    st_view_release(&hv);
    return rc;
}
