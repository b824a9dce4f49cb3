/*
 * init_amd.h - SlimTracin's device mirror over libslimtracin.so (MI355X).
 *
 * The reference builds its GPU path only under nvcc: core/init.h:10-126
 * defines allocateDeviceScene, uploadPrimitives, uploadMaterials,
 * uploadLights, uploadScene, uploadSceneBVH and uploadMeshBVHs over CUDA
 * buffers, render/raytracer.h:258-299 defines renderSceneOnGPU, and
 * renderScene (raytracer.h:302-309) picks it when viewport->settings.use_GPU
 * is set. This header is the same function set over the library's C ABI
 * (include/slimtracin_seam.h), for a C build of the reference. A maintainer
 * adds it to core/ and changes two guards:
 *
 *   core/init.h         #ifdef __CUDACC__ ... #else (stubs) #endif
 *                    -> #ifdef __CUDACC__ ... #elif defined(SLIMTRACIN_AMD)
 *                       #include "./init_amd.h"  #else (stubs) #endif
 *   render/raytracer.h  renderScene's #ifdef __CUDACC__
 *                    -> #if defined(__CUDACC__) || defined(SLIMTRACIN_AMD)
 *                       (and the same guard around nothing else: the CUDA
 *                       kernel d_render stays nvcc-only)
 *
 * then builds with -DSLIMTRACIN_AMD -I<this repo>/include and links
 * libslimtracin.so. It must be included after core/types.h (it reads the
 * reference's Scene, Viewport, Primitive, Material, Light, Mesh, Triangle,
 * BVHNode and Texture).
 *
 * The reference's structs are converted field by field into the records of
 * st_blob.h: their layouts differ (the reference's u32 is an unsigned long,
 * 8 bytes on LP64 Linux, and its enums and bools sit between fields), so
 * nothing is memcpy'd across except FloatPixel, whose layout (vec3 color, f32
 * opacity, f64 depth: 24 bytes) the library shares and this header checks.
 *
 * Semantics kept: every upload copies the host state at call time;
 * uploadPrimitives without uploadSceneBVH renders the new transforms over the
 * old scene BVH (examples/02_Geometry.c:70-75); primitives carry the screen
 * bounds updateSceneSSB left in them (render/SSB.h:137-173), so a primitive
 * is gated exactly as the reference's renderer gates it; the frame is the
 * CPU renderer's frame, (W+1)(H+1) FloatPixels under QCAA. The reference's
 * error handling (gpuErrchk: print and exit in debug builds, nothing in
 * release; core/base.h:19-38) becomes SLIMTRACIN_AMD_ON_ERROR, by default
 * print st_last_error() and exit(1).
 */
#pragma once

#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "slimtracin_seam.h"

#ifndef SLIMTRACIN_AMD_ON_ERROR
#define SLIMTRACIN_AMD_ON_ERROR(rc, what)                                                   \
    do {                                                                                   \
        fprintf(stderr, "slimtracin: %s failed (%d): %s\n", (what), (rc), st_last_error()); \
        exit(1);                                                                           \
    } while (0)
#endif

_Static_assert(sizeof(FloatPixel) == sizeof(StFloatPixel), "FloatPixel is 24 bytes in both layouts");
_Static_assert(offsetof(FloatPixel, opacity) == offsetof(StFloatPixel, opacity), "FloatPixel opacity offset");
_Static_assert(offsetof(FloatPixel, depth) == offsetof(StFloatPixel, depth), "FloatPixel depth offset");

static StSceneHandle *amd_scene;
static StSeamSizes amd_sizes;
static StPrimRec *amd_prims;
static int32_t *amd_bounds;
static StMaterialRec *amd_materials;
static StLightRec *amd_lights;

#define AMD_CALL(call, what)                              \
    do {                                                  \
        int amd_rc_ = (call);                             \
        if (amd_rc_) SLIMTRACIN_AMD_ON_ERROR(amd_rc_, what); \
    } while (0)

static void amdVec3(float *out, vec3 v) {
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}

/* Primitive (core/types.h:144-152) -> StPrimRec, screen_bounds -> 4 ints. */
static void amdConvertPrimitives(Scene *scene) {
    for (u32 i = 0; i < scene->settings.primitives; i++) {
        const Primitive *p = scene->primitives + i;
        StPrimRec *r = amd_prims + i;
        amdVec3(r->position, p->position);
        amdVec3(r->scale, p->scale);
        r->rotation[0] = p->rotation.axis.x;
        r->rotation[1] = p->rotation.axis.y;
        r->rotation[2] = p->rotation.axis.z;
        r->rotation[3] = p->rotation.amount;
        r->type = (uint32_t)p->type;
        r->mesh_id = (uint32_t)p->id;
        r->material_id = p->material_id;
        r->flags = p->flags;
        amd_bounds[4 * i + 0] = p->screen_bounds.min.x;
        amd_bounds[4 * i + 1] = p->screen_bounds.min.y;
        amd_bounds[4 * i + 2] = p->screen_bounds.max.x;
        amd_bounds[4 * i + 3] = p->screen_bounds.max.y;
    }
}

/* Material (core/types.h:320-326) -> StMaterialRec. */
static void amdConvertMaterials(Scene *scene) {
    for (u32 i = 0; i < scene->settings.materials; i++) {
        const Material *m = scene->materials + i;
        StMaterialRec *r = amd_materials + i;
        memset(r, 0, sizeof *r);
        amdVec3(r->albedo, m->albedo);
        amdVec3(r->reflectivity, m->reflectivity);
        amdVec3(r->emission, m->emission);
        r->uv_repeat[0] = m->uv_repeat.u;
        r->uv_repeat[1] = m->uv_repeat.v;
        r->metallic = m->metallic;
        r->roughness = m->roughness;
        r->n1_over_n2 = m->n1_over_n2;
        r->n2_over_n1 = m->n2_over_n1;
        r->brdf = (uint32_t)m->brdf;
        r->is = m->is;
        r->use = m->use;
        r->texture_count = m->texture_count;
        memcpy(r->texture_ids, m->texture_ids, 16);
    }
}

/* Light (core/types.h:300-304) -> StLightRec; the renderer reads position,
   colour and intensity (shading/surface.h:9-35, closest_hit/lights.h). */
static void amdConvertLights(Scene *scene) {
    for (u32 i = 0; i < scene->settings.lights; i++) {
        const Light *l = scene->lights + i;
        StLightRec *r = amd_lights + i;
        amdVec3(r->position, l->position_or_direction);
        amdVec3(r->color, l->color);
        r->intensity = l->intensity;
    }
}

static void amdConvertNode(StBVHNodeRec *r, const BVHNode *n) {
    amdVec3(r->min, n->aabb.min);
    amdVec3(r->max, n->aabb.max);
    r->first_child_id = (uint32_t)n->first_child_id;
    r->child_count = n->child_count;
    r->depth = n->depth;
}

/* Triangle (core/types.h:272-278) -> StTriangleRec. */
static void amdConvertTriangle(StTriangleRec *r, const Triangle *t) {
    amdVec3(r->world_to_tangent + 0, t->world_to_tangent.X);
    amdVec3(r->world_to_tangent + 3, t->world_to_tangent.Y);
    amdVec3(r->world_to_tangent + 6, t->world_to_tangent.Z);
    amdVec3(r->position, t->position);
    amdVec3(r->normal, t->normal);
    for (int k = 0; k < 3; k++) {
        amdVec3(r->vertex_normals + 3 * k, t->vertex_normals[k]);
        r->uvs[2 * k] = t->uvs[k].x;
        r->uvs[2 * k + 1] = t->uvs[k].y;
    }
    r->area_of_parallelogram = t->area_of_parallelogram;
    r->area_of_uv = t->area_of_uv;
}

/* core/init.h:36 - capacities from scene->settings, as the CUDA build
   sizes its device buffers once. */
void allocateDeviceScene(Scene *scene) {
    StSeamSizes z = {scene->settings.primitives, scene->settings.materials, scene->settings.lights,
                     scene->settings.meshes, scene->settings.textures};
    amd_sizes = z;
    free(amd_prims);
    free(amd_bounds);
    free(amd_materials);
    free(amd_lights);
    amd_prims = (StPrimRec *)calloc(z.primitives + 1, sizeof(StPrimRec));
    amd_bounds = (int32_t *)calloc(4 * ((size_t)z.primitives + 1), sizeof(int32_t));
    amd_materials = (StMaterialRec *)calloc(z.materials + 1, sizeof(StMaterialRec));
    amd_lights = (StLightRec *)calloc(z.lights + 1, sizeof(StLightRec));
    if (!amd_prims || !amd_bounds || !amd_materials || !amd_lights)
        SLIMTRACIN_AMD_ON_ERROR(ST_ERR_MEMORY, "allocateDeviceScene");
    if (amd_scene) st_scene_destroy(amd_scene);
    amd_scene = NULL;
    AMD_CALL(st_seam_allocate(&z, &amd_scene), "allocateDeviceScene");
}

/* core/init.h:74 */
void uploadPrimitives(Scene *scene) {
    amdConvertPrimitives(scene);
    AMD_CALL(st_seam_upload_primitives(amd_scene, amd_prims, amd_bounds), "uploadPrimitives");
}

/* core/init.h:77 */
void uploadMaterials(Scene *scene) {
    amdConvertMaterials(scene);
    AMD_CALL(st_seam_upload_materials(amd_scene, amd_materials), "uploadMaterials");
}

/* core/init.h:80 (area lights are not read by the renderer: emissive quads
   are primitives) */
void uploadLights(Scene *scene) {
    amdConvertLights(scene);
    AMD_CALL(st_seam_upload_lights(amd_scene, amd_lights), "uploadLights");
}

/* core/init.h:85 - lights, primitives, materials and every texture's mip
   chain (the .texture layout: (w+1)(h+1) texel quads per mip). */
void uploadScene(Scene *scene) {
    const u32 T = scene->settings.textures;
    StSeamTexture *tex = (StSeamTexture *)calloc(T + 1, sizeof(StSeamTexture));
    StSeamMip **mips = (StSeamMip **)calloc(T + 1, sizeof(StSeamMip *));
    if (!tex || !mips) SLIMTRACIN_AMD_ON_ERROR(ST_ERR_MEMORY, "uploadScene");
    for (u32 t = 0; t < T; t++) {
        const Texture *x = scene->textures + t;
        mips[t] = (StSeamMip *)calloc(x->mip_count + 1, sizeof(StSeamMip));
        if (!mips[t]) SLIMTRACIN_AMD_ON_ERROR(ST_ERR_MEMORY, "uploadScene");
        for (u32 m = 0; m < x->mip_count; m++) {
            mips[t][m].width = x->mips[m].width;
            mips[t][m].height = x->mips[m].height;
            mips[t][m].texel_quads = (const uint8_t *)x->mips[m].texel_quads;
        }
        tex[t].width = x->width;
        tex[t].height = x->height;
        tex[t].mip_count = x->mip_count;
        tex[t].wrap = x->wrap ? 1 : 0;
        tex[t].mipmap = x->mipmap ? 1 : 0;
        tex[t].pad = 0;
        tex[t].mips = mips[t];
    }
    amdConvertPrimitives(scene);
    amdConvertMaterials(scene);
    amdConvertLights(scene);
    AMD_CALL(st_seam_upload_scene(amd_scene, amd_lights, amd_prims, amd_bounds, amd_materials, T ? tex : NULL),
             "uploadScene");
    for (u32 t = 0; t < T; t++) free(mips[t]);
    free(mips);
    free(tex);
}

/* core/init.h:106 */
void uploadSceneBVH(Scene *scene) {
    const u32 N = scene->bvh.node_count, P = scene->settings.primitives;
    StBVHNodeRec *nodes = (StBVHNodeRec *)calloc(N + 1, sizeof(StBVHNodeRec));
    uint32_t *leaf = (uint32_t *)calloc(P + 1, sizeof(uint32_t));
    if (!nodes || !leaf) SLIMTRACIN_AMD_ON_ERROR(ST_ERR_MEMORY, "uploadSceneBVH");
    for (u32 i = 0; i < N; i++) amdConvertNode(nodes + i, scene->bvh.nodes + i);
    for (u32 i = 0; i < P; i++) leaf[i] = (uint32_t)scene->bvh.leaf_ids[i];
    AMD_CALL(st_seam_upload_scene_bvh(amd_scene, nodes, N, (uint32_t)scene->bvh.height, leaf), "uploadSceneBVH");
    free(nodes);
    free(leaf);
}

/* core/init.h:111 - every mesh's BVH nodes, its triangles in leaf order and
   its leaf ids. */
void uploadMeshBVHs(Scene *scene) {
    const u32 M = scene->settings.meshes;
    StSeamMesh *out = (StSeamMesh *)calloc(M + 1, sizeof(StSeamMesh));
    StBVHNodeRec **nodes = (StBVHNodeRec **)calloc(M + 1, sizeof(StBVHNodeRec *));
    StTriangleRec **tris = (StTriangleRec **)calloc(M + 1, sizeof(StTriangleRec *));
    uint32_t **leaf = (uint32_t **)calloc(M + 1, sizeof(uint32_t *));
    if (!out || !nodes || !tris || !leaf) SLIMTRACIN_AMD_ON_ERROR(ST_ERR_MEMORY, "uploadMeshBVHs");
    for (u32 m = 0; m < M; m++) {
        const Mesh *mesh = scene->meshes + m;
        const u32 N = mesh->bvh.node_count, T = mesh->triangle_count;
        nodes[m] = (StBVHNodeRec *)calloc(N + 1, sizeof(StBVHNodeRec));
        tris[m] = (StTriangleRec *)calloc(T + 1, sizeof(StTriangleRec));
        leaf[m] = (uint32_t *)calloc(T + 1, sizeof(uint32_t));
        if (!nodes[m] || !tris[m] || !leaf[m]) SLIMTRACIN_AMD_ON_ERROR(ST_ERR_MEMORY, "uploadMeshBVHs");
        for (u32 i = 0; i < N; i++) amdConvertNode(nodes[m] + i, mesh->bvh.nodes + i);
        for (u32 t = 0; t < T; t++) {
            amdConvertTriangle(tris[m] + t, mesh->triangles + t);
            leaf[m][t] = (uint32_t)mesh->bvh.leaf_ids[t];
        }
        amdVec3(out[m].aabb, mesh->aabb.min);
        amdVec3(out[m].aabb + 3, mesh->aabb.max);
        out[m].triangle_count = T;
        out[m].normals_count = mesh->normals_count;
        out[m].uvs_count = mesh->uvs_count;
        out[m].node_count = N;
        out[m].height = (uint32_t)mesh->bvh.height;
        out[m].nodes = nodes[m];
        out[m].triangles = tris[m];
        out[m].leaf_ids = leaf[m];
    }
    AMD_CALL(st_seam_upload_mesh_bvhs(amd_scene, out), "uploadMeshBVHs");
    for (u32 m = 0; m < M; m++) {
        free(nodes[m]);
        free(tris[m]);
        free(leaf[m]);
    }
    free(nodes);
    free(tris);
    free(leaf);
    free(out);
}

/* render/raytracer.h:258-299 - one frame into the viewport's FloatPixels
   (blocking, as the reference's download is). */
void renderSceneOnGPU(Scene *scene, Viewport *viewport) {
    const ProjectionPlane *pp = &viewport->projection_plane;
    const xform3 *cam = &viewport->camera->transform;
    StSeamFrame f;
    memset(&f, 0, sizeof f);
    amdVec3(f.start, pp->start);
    amdVec3(f.right, pp->right);
    amdVec3(f.down, pp->down);
    f.cone_angle = pp->cone_angle;
    amdVec3(f.camera_position, cam->position);
    f.camera_rotation_inverted[0] = cam->rotation_inverted.axis.x;
    f.camera_rotation_inverted[1] = cam->rotation_inverted.axis.y;
    f.camera_rotation_inverted[2] = cam->rotation_inverted.axis.z;
    f.camera_rotation_inverted[3] = cam->rotation_inverted.amount;
    amdVec3(f.ambient, scene->ambient_light.color);
    f.width = (uint16_t)viewport->frame_buffer->dimensions.width;
    f.height = (uint16_t)viewport->frame_buffer->dimensions.height;
    f.render_mode = (uint32_t)viewport->settings.render_mode;
    f.depth = viewport->trace.depth;
    AMD_CALL(st_seam_render(amd_scene, &f, (StFloatPixel *)viewport->frame_buffer->float_pixels, NULL, NULL),
             "renderSceneOnGPU");
}

#undef AMD_CALL
