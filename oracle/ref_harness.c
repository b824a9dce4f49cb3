/*
 * ref_harness.c - headless driver around the reference CPU renderer.
 *
 * TEST INFRASTRUCTURE ONLY. This file is compiled together with the
 * reference's own headers, read in place from the reference checkout
 * (SlimTracin/src/SlimTracin, never copied), into oracle/_ref/libslimref.so.
 * It plays the part of the reference's Win32 platform layer and example app:
 * it turns a scene blob (include/st_blob.h) into the reference's Scene /
 * Viewport structs through the reference's own init functions, then calls
 * the reference's BVH builder, SSB pass and CPU renderer. Nothing here
 * re-implements reference arithmetic except the ray set-up loop that mirrors
 * renderSceneOnCPU (render/raytracer.h:136-171) so that the primary-hit pass
 * and the multi-threaded row split see the same rays.
 *
 * Init order follows app.h:_initApp (app.h:201-308) and the resize step
 * (_windowResize, app.h:31-39) that a headless run must reproduce.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "app.h"
#include "render/raytracer.h"
#include "scene/xform.h"
#include "viewport/manipulation.h"

#include "st_blob.h"

void initApp(Defaults *defaults) { (void)defaults; }

typedef struct RefScene {
    Scene scene;
    Memory memory;
    u8 *arena;
    BVHBuilder builder;
    u32 max_leaves;
} RefScene;

typedef struct Reader {
    const uint8_t *p, *end;
    int bad;
} Reader;

static const void *take(Reader *r, size_t n) {
    if (r->bad || (size_t)(r->end - r->p) < n) { r->bad = 1; return NULL; }
    const void *out = r->p;
    r->p += (n + 3) & ~(size_t)3;
    if (r->p > r->end) r->p = r->end;
    return out;
}

static void fillVec3(vec3 *v, const float *f) { v->x = f[0]; v->y = f[1]; v->z = f[2]; }

static void refFree(RefScene *rs) {
    if (!rs) return;
    free(rs->arena);
    free(rs);
}

/* Build the reference scene. `build_scene_bvh` runs updateSceneBVH. */
static RefScene *refLoad(const void *blob, size_t size) {
    Reader r = {(const uint8_t *)blob, (const uint8_t *)blob + size, 0};
    const StBlobHeader *h = (const StBlobHeader *)take(&r, sizeof(StBlobHeader));
    if (!h || h->magic != ST_BLOB_MAGIC || h->version != ST_BLOB_VERSION) return NULL;

    RefScene *rs = (RefScene *)calloc(1, sizeof(RefScene));
    /* Arena sized generously from the blob size: all reference allocations are
       bump allocations (core/base.h:275-294). */
    u64 cap = (u64)size * 64 + (4u << 20) + (u64)(h->prim_count + 16) * 4096;
    rs->arena = (u8 *)calloc(1, cap);
    initMemory(&rs->memory, rs->arena, cap);
    Memory *mem = &rs->memory;
    Scene *s = &rs->scene;

    s->settings.cameras = 1;
    s->settings.primitives = h->prim_count;
    s->settings.materials = h->material_count;
    s->settings.lights = h->light_count;
    s->settings.meshes = h->mesh_count;
    s->settings.textures = h->texture_count;
    s->ambient_light.color = Vec3(h->ambient[0], h->ambient[1], h->ambient[2]);

    const StPrimRec *prims = (const StPrimRec *)take(&r, sizeof(StPrimRec) * h->prim_count);
    const StMaterialRec *mats = (const StMaterialRec *)take(&r, sizeof(StMaterialRec) * h->material_count);
    const StLightRec *lights = (const StLightRec *)take(&r, sizeof(StLightRec) * h->light_count);
    if (r.bad) { refFree(rs); return NULL; }

    s->primitives = (Primitive *)allocateMemory(mem, sizeof(Primitive) * (h->prim_count ? h->prim_count : 1));
    for (u32 i = 0; i < h->prim_count; i++) {
        Primitive *p = s->primitives + i;
        initPrimitive(p);
        fillVec3(&p->position, prims[i].position);
        fillVec3(&p->scale, prims[i].scale);
        p->rotation.axis = Vec3(prims[i].rotation[0], prims[i].rotation[1], prims[i].rotation[2]);
        p->rotation.amount = prims[i].rotation[3];
        p->type = (enum PrimitiveType)prims[i].type;
        p->id = prims[i].mesh_id;
        p->material_id = (u8)prims[i].material_id;
        p->flags = (u8)prims[i].flags;
    }
    s->materials = (Material *)allocateMemory(mem, sizeof(Material) * (h->material_count ? h->material_count : 1));
    for (u32 i = 0; i < h->material_count; i++) {
        Material *m = s->materials + i;
        const StMaterialRec *q = mats + i;
        initMaterial(m);
        fillVec3(&m->albedo, q->albedo);
        fillVec3(&m->reflectivity, q->reflectivity);
        fillVec3(&m->emission, q->emission);
        m->uv_repeat.u = q->uv_repeat[0];
        m->uv_repeat.v = q->uv_repeat[1];
        m->metallic = q->metallic;
        m->roughness = q->roughness;
        m->n1_over_n2 = q->n1_over_n2;
        m->n2_over_n1 = q->n2_over_n1;
        m->brdf = (enum BRDF)q->brdf;
        m->is = q->is;
        m->use = q->use;
        m->texture_count = q->texture_count;
        for (int t = 0; t < 16; t++) m->texture_ids[t] = q->texture_ids[t];
    }
    if (h->light_count) {
        s->lights = (Light *)allocateMemory(mem, sizeof(Light) * h->light_count);
        for (u32 i = 0; i < h->light_count; i++) {
            Light *l = s->lights + i;
            fillVec3(&l->position_or_direction, lights[i].position);
            fillVec3(&l->color, lights[i].color);
            l->attenuation = getVec3Of(1);
            l->intensity = lights[i].intensity;
            l->is_directional = false;
        }
    }

    u32 max_tris = 0;
    if (h->mesh_count) {
        s->meshes = (Mesh *)allocateMemory(mem, sizeof(Mesh) * h->mesh_count);
        s->mesh_bvh_node_counts = (u32 *)allocateMemory(mem, sizeof(u32) * h->mesh_count);
        s->mesh_triangle_counts = (u32 *)allocateMemory(mem, sizeof(u32) * h->mesh_count);
    }
    /* First pass over meshes: geometry. BVHs are built after the builder exists. */
    const StMeshRec **mrecs = (const StMeshRec **)calloc(h->mesh_count + 1, sizeof(void *));
    const StBVHNodeRec **pre_nodes = (const StBVHNodeRec **)calloc(h->mesh_count + 1, sizeof(void *));
    const StTriangleRec **pre_tris = (const StTriangleRec **)calloc(h->mesh_count + 1, sizeof(void *));
    const uint32_t **pre_leaf = (const uint32_t **)calloc(h->mesh_count + 1, sizeof(void *));
    for (u32 m = 0; m < h->mesh_count; m++) {
        const StMeshRec *mr = (const StMeshRec *)take(&r, sizeof(StMeshRec));
        if (!mr) break;
        mrecs[m] = mr;
        Mesh *mesh = s->meshes + m;
        memset(mesh, 0, sizeof(Mesh));
        mesh->aabb.min = Vec3(mr->aabb[0], mr->aabb[1], mr->aabb[2]);
        mesh->aabb.max = Vec3(mr->aabb[3], mr->aabb[4], mr->aabb[5]);
        mesh->vertex_count = mr->vertex_count;
        mesh->triangle_count = mr->triangle_count;
        mesh->normals_count = mr->normals_count;
        mesh->uvs_count = mr->uvs_count;
        if (mr->triangle_count > max_tris) max_tris = mr->triangle_count;
        initBVH(&mesh->bvh, mr->triangle_count, mem);

        const float *pos = (const float *)take(&r, 12u * mr->vertex_count);
        const uint32_t *pidx = (const uint32_t *)take(&r, 12u * mr->triangle_count);
        mesh->vertex_positions = (vec3 *)allocateMemory(mem, sizeof(vec3) * (mr->vertex_count + 1));
        mesh->vertex_position_indices = (TriangleVertexIndices *)allocateMemory(mem, sizeof(TriangleVertexIndices) * (mr->triangle_count + 1));
        mesh->triangles = (Triangle *)allocateMemory(mem, sizeof(Triangle) * (mr->triangle_count + 1));
        if (r.bad) break;
        for (u32 v = 0; v < mr->vertex_count; v++) fillVec3(mesh->vertex_positions + v, pos + 3 * v);
        for (u32 t = 0; t < mr->triangle_count * 3; t++) mesh->vertex_position_indices[t / 3].ids[t % 3] = pidx[t];
        if (mr->normals_count) {
            const float *nrm = (const float *)take(&r, 12u * mr->normals_count);
            const uint32_t *nidx = (const uint32_t *)take(&r, 12u * mr->triangle_count);
            if (r.bad) break;
            mesh->vertex_normals = (vec3 *)allocateMemory(mem, sizeof(vec3) * mr->normals_count);
            mesh->vertex_normal_indices = (TriangleVertexIndices *)allocateMemory(mem, sizeof(TriangleVertexIndices) * mr->triangle_count);
            for (u32 v = 0; v < mr->normals_count; v++) fillVec3(mesh->vertex_normals + v, nrm + 3 * v);
            for (u32 t = 0; t < mr->triangle_count * 3; t++) mesh->vertex_normal_indices[t / 3].ids[t % 3] = nidx[t];
        }
        if (mr->uvs_count) {
            const float *uv = (const float *)take(&r, 8u * mr->uvs_count);
            const uint32_t *uidx = (const uint32_t *)take(&r, 12u * mr->triangle_count);
            if (r.bad) break;
            mesh->vertex_uvs = (vec2 *)allocateMemory(mem, sizeof(vec2) * mr->uvs_count);
            mesh->vertex_uvs_indices = (TriangleVertexIndices *)allocateMemory(mem, sizeof(TriangleVertexIndices) * mr->triangle_count);
            for (u32 v = 0; v < mr->uvs_count; v++) { mesh->vertex_uvs[v].x = uv[2 * v]; mesh->vertex_uvs[v].y = uv[2 * v + 1]; }
            for (u32 t = 0; t < mr->triangle_count * 3; t++) mesh->vertex_uvs_indices[t / 3].ids[t % 3] = uidx[t];
        }
        if (mr->prebuilt) {
            pre_nodes[m] = (const StBVHNodeRec *)take(&r, sizeof(StBVHNodeRec) * mr->node_count);
            pre_tris[m] = (const StTriangleRec *)take(&r, sizeof(StTriangleRec) * mr->triangle_count);
            pre_leaf[m] = (const uint32_t *)take(&r, 4u * mr->triangle_count);
            if (r.bad) break;
        }
    }
    if (h->texture_count && !r.bad) {
        s->textures = (Texture *)allocateMemory(mem, sizeof(Texture) * h->texture_count);
        for (u32 t = 0; t < h->texture_count && !r.bad; t++) {
            const StTextureRec *tr = (const StTextureRec *)take(&r, sizeof(StTextureRec));
            if (!tr) break;
            Texture *tex = s->textures + t;
            tex->width = tr->width;
            tex->height = tr->height;
            tex->mip_count = tr->mip_count;
            tex->wrap = tr->wrap;
            tex->mipmap = tr->mipmap;
            tex->mips = (TextureMip *)allocateMemory(mem, sizeof(TextureMip) * (tr->mip_count ? tr->mip_count : 1));
            for (u32 mi = 0; mi < tr->mip_count; mi++) {
                const uint16_t *dims = (const uint16_t *)take(&r, 4);
                if (!dims) break;
                tex->mips[mi].width = dims[0];
                tex->mips[mi].height = dims[1];
                size_t nq = (size_t)(dims[0] + 1) * (dims[1] + 1);
                const uint8_t *q = (const uint8_t *)take(&r, 12 * nq);
                if (!q) break;
                tex->mips[mi].texel_quads = (TexelQuad *)allocateMemory(mem, sizeof(TexelQuad) * nq);
                memcpy(tex->mips[mi].texel_quads, q, 12 * nq);
            }
        }
    }
    /* Optional scene BVH built elsewhere (the Morton LBVH): loaded into the
       reference's Scene.bvh in place of updateSceneBVH. */
    const uint32_t *pre_scene_head = NULL;
    const StBVHNodeRec *pre_scene_nodes = NULL;
    const uint32_t *pre_scene_leaf = NULL;
    if ((h->flags & ST_BLOB_HAS_SCENE_BVH) && !r.bad) {
        pre_scene_head = (const uint32_t *)take(&r, 8);
        if (pre_scene_head && pre_scene_head[0] <= 2u * h->prim_count) {
            pre_scene_nodes = (const StBVHNodeRec *)take(&r, sizeof(StBVHNodeRec) * pre_scene_head[0]);
            pre_scene_leaf = (const uint32_t *)take(&r, 4u * h->prim_count);
        } else {
            r.bad = 1;
        }
    }
    if (r.bad) { free(mrecs); free(pre_nodes); free(pre_tris); free(pre_leaf); refFree(rs); return NULL; }

    /* Builder sized like _initApp (app.h:277-279) then initBVHBuilder (builder_top_down.h:42-64). */
    rs->max_leaves = h->prim_count > max_tris ? h->prim_count : max_tris;
#ifdef REF_HARNESS_BEFORE_BVH
    /* a build with a device mirror: allocateDeviceScene and uploadScene come
       before the BVH builds, in _initApp's order (app.h:160-199, 296-298) */
    REF_HARNESS_BEFORE_BVH(s);
#endif
    initBVHBuilder(&rs->builder, s, mem);
    initBVH(&s->bvh, h->prim_count, mem);

    for (u32 m = 0; m < h->mesh_count; m++) {
        Mesh *mesh = s->meshes + m;
        if (pre_nodes[m]) {
            mesh->bvh.node_count = mrecs[m]->node_count;
            mesh->bvh.height = mrecs[m]->height;
            for (u32 n = 0; n < mrecs[m]->node_count; n++) {
                const StBVHNodeRec *q = pre_nodes[m] + n;
                BVHNode *node = mesh->bvh.nodes + n;
                node->aabb.min = Vec3(q->min[0], q->min[1], q->min[2]);
                node->aabb.max = Vec3(q->max[0], q->max[1], q->max[2]);
                node->first_child_id = q->first_child_id;
                node->child_count = q->child_count;
                node->depth = q->depth;
            }
            for (u32 t = 0; t < mesh->triangle_count; t++) {
                const StTriangleRec *q = pre_tris[m] + t;
                Triangle *tri = mesh->triangles + t;
                tri->world_to_tangent.X = Vec3(q->world_to_tangent[0], q->world_to_tangent[1], q->world_to_tangent[2]);
                tri->world_to_tangent.Y = Vec3(q->world_to_tangent[3], q->world_to_tangent[4], q->world_to_tangent[5]);
                tri->world_to_tangent.Z = Vec3(q->world_to_tangent[6], q->world_to_tangent[7], q->world_to_tangent[8]);
                fillVec3(&tri->position, q->position);
                fillVec3(&tri->normal, q->normal);
                for (int k = 0; k < 3; k++) fillVec3(&tri->vertex_normals[k], q->vertex_normals + 3 * k);
                for (int k = 0; k < 3; k++) { tri->uvs[k].x = q->uvs[2 * k]; tri->uvs[k].y = q->uvs[2 * k + 1]; }
                tri->area_of_parallelogram = q->area_of_parallelogram;
                tri->area_of_uv = q->area_of_uv;
                mesh->bvh.leaf_ids[t] = pre_leaf[m][t];
            }
        } else {
            updateMeshBVH(mesh, &rs->builder);
        }
        s->mesh_bvh_node_counts[m] = mesh->bvh.node_count;
        s->mesh_triangle_counts[m] = mesh->triangle_count;
    }
    free(mrecs); free(pre_nodes); free(pre_tris); free(pre_leaf);

    if (pre_scene_nodes) {
        s->bvh.node_count = pre_scene_head[0];
        s->bvh.height = pre_scene_head[1];
        for (u32 n = 0; n < pre_scene_head[0]; n++) {
            const StBVHNodeRec *q = pre_scene_nodes + n;
            BVHNode *node = s->bvh.nodes + n;
            node->aabb.min = Vec3(q->min[0], q->min[1], q->min[2]);
            node->aabb.max = Vec3(q->max[0], q->max[1], q->max[2]);
            node->first_child_id = q->first_child_id;
            node->child_count = q->child_count;
            node->depth = q->depth;
        }
        for (u32 i = 0; i < h->prim_count; i++) s->bvh.leaf_ids[i] = pre_scene_leaf[i];
    } else if (h->prim_count) {
        updateSceneBVH(s, &rs->builder);
    }
    return rs;
}

typedef struct RefView {
    Viewport viewport;
    Camera camera;
    PixelGrid frame;
    Trace trace;
} RefView;

/* Camera, frame and trace set-up for one view: _initApp + _windowResize + beginDrawing. */
static RefView *refViewInit(RefScene *rs, const StView *v) {
    RefView *rv = (RefView *)calloc(1, sizeof(RefView));
    Scene *s = &rs->scene;
    initCamera(&rv->camera);
    rv->camera.transform.position = Vec3(v->position[0], v->position[1], v->position[2]);
    rotateXform3(&rv->camera.transform, v->yaw, v->pitch, v->roll);
    rv->camera.focal_length = v->focal_length;

    rv->frame.QCAA = true;
    updateDimensions(&rv->frame.dimensions, v->width, v->height, true);
    rv->frame.float_pixels = (FloatPixel *)calloc((size_t)rv->frame.dimensions.width_times_height, sizeof(FloatPixel));

    ViewportSettings vs;
    memset(&vs, 0, sizeof(vs));
    setDefaultViewportSettings(&vs);
    vs.render_mode = (enum RenderMode)v->render_mode;
    NavigationSettings ns;
    setDefaultNavigationSettings(&ns);
    initViewport(&rv->viewport, &vs, &ns, &rv->camera, &rv->frame);

    initTrace(&rv->viewport.trace, s, &rs->memory);
    rv->viewport.trace.depth = (u8)v->depth;
    rv->viewport.projection_plane.cone_angle = v->cone_angle;

    updateSceneSSB(s, &rv->viewport);
    beginDrawing(&rv->viewport);
    return rv;
}

static void refViewFree(RefView *rv) {
    if (!rv) return;
    free(rv->frame.float_pixels);
    free(rv);
}

static void copyOutPixels(RefView *rv, StFloatPixel *out) {
    u32 n = rv->frame.dimensions.width_times_height;
    for (u32 i = 0; i < n; i++) {
        FloatPixel *p = rv->frame.float_pixels + i;
        out[i].r = p->color.r;
        out[i].g = p->color.g;
        out[i].b = p->color.b;
        out[i].opacity = p->opacity;
        out[i].depth = p->depth;
    }
}

/* Row-interleaved multi-threaded variant of renderSceneOnCPU: one Trace per
   thread, row starts accumulated sequentially exactly as the single-threaded
   loop does (raytracer.h:156-170). */
typedef struct RowJob {
    RefScene *rs;
    RefView *rv;
    int tid, nthreads;
    u32 *stack_mem;
} RowJob;

static void *rowWorker(void *arg) {
    RowJob *job = (RowJob *)arg;
    Scene *scene = &job->rs->scene;
    Viewport *viewport = &job->rv->viewport;
    Dimensions *dim = &viewport->frame_buffer->dimensions;
    Trace trace;
    memset(&trace, 0, sizeof(trace));
    trace.closest_mesh_hit.object_type = PrimitiveType_Mesh;
    trace.mesh_stack_size = viewport->trace.mesh_stack_size;
    trace.scene_stack_size = viewport->trace.scene_stack_size;
    trace.depth = viewport->trace.depth;
    trace.scene_stack = job->stack_mem;
    trace.mesh_stack = job->stack_mem + 256;

    quat camera_rotation = viewport->camera->transform.rotation_inverted;
    vec3 camera_position = viewport->camera->transform.position;
    vec3 start = viewport->projection_plane.start;
    vec3 right = viewport->projection_plane.right;
    vec3 down = viewport->projection_plane.down;
    enum RenderMode mode = viewport->settings.render_mode;
    Ray ray;
    for (u16 y = 0; y < dim->height; y++) {
        if (y % job->nthreads == job->tid) {
            vec3 current = start;
            FloatPixel *pixel = viewport->frame_buffer->float_pixels + (u32)y * dim->width;
            for (u16 x = 0; x < dim->width; x++, pixel++) {
                ray.origin = camera_position;
                ray.direction = normVec3(current);
                ray.direction_reciprocal = oneOverVec3(ray.direction);
                trace.closest_hit.distance = trace.closest_hit.distance_squared = INFINITY;
                trace.closest_hit.cone_angle = viewport->projection_plane.cone_angle;
                trace.closest_hit.cone_width = 0;
                rayTrace(&ray, &trace, scene, mode, pixel, x, y, camera_position, camera_rotation);
                current = addVec3(current, right);
            }
        }
        start = addVec3(start, down);
    }
    return NULL;
}

/* ---- exported API ------------------------------------------------------ */

/* Render one frame with the reference CPU renderer. threads <= 1 calls
   renderSceneOnCPU itself. pixels: (width+1)*(height+1) records. */
int ref_render(const void *blob, size_t size, const StView *view, StFloatPixel *pixels, int threads) {
    RefScene *rs = refLoad(blob, size);
    if (!rs) return -1;
    RefView *rv = refViewInit(rs, view);
    if (threads <= 1) {
        renderSceneOnCPU(&rs->scene, &rv->viewport);
    } else {
        pthread_t th[256];
        RowJob jobs[256];
        if (threads > 256) threads = 256;
        u32 *stacks = (u32 *)calloc((size_t)threads * 512, sizeof(u32));
        for (int t = 0; t < threads; t++) {
            jobs[t].rs = rs; jobs[t].rv = rv; jobs[t].tid = t; jobs[t].nthreads = threads;
            jobs[t].stack_mem = stacks + (size_t)t * 512;
            pthread_create(&th[t], NULL, rowWorker, &jobs[t]);
        }
        for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
        free(stacks);
    }
    copyOutPixels(rv, pixels);
    refViewFree(rv);
    refFree(rs);
    return 0;
}

/* Timed render: loads once, renders `frames` times, returns the per-frame
   render-call seconds in `seconds` (BVH build and set-up excluded). */
#include <time.h>
static double nowSeconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
int ref_render_timed(const void *blob, size_t size, const StView *view, StFloatPixel *pixels,
                     int threads, int frames, double *seconds) {
    RefScene *rs = refLoad(blob, size);
    if (!rs) return -1;
    RefView *rv = refViewInit(rs, view);
    for (int f = 0; f < frames; f++) {
        beginDrawing(&rv->viewport);
        double t0 = nowSeconds();
        if (threads <= 1) {
            renderSceneOnCPU(&rs->scene, &rv->viewport);
        } else {
            pthread_t th[256];
            RowJob jobs[256];
            int n = threads > 256 ? 256 : threads;
            u32 *stacks = (u32 *)calloc((size_t)n * 512, sizeof(u32));
            for (int t = 0; t < n; t++) {
                jobs[t].rs = rs; jobs[t].rv = rv; jobs[t].tid = t; jobs[t].nthreads = n;
                jobs[t].stack_mem = stacks + (size_t)t * 512;
                pthread_create(&th[t], NULL, rowWorker, &jobs[t]);
            }
            for (int t = 0; t < n; t++) pthread_join(th[t], NULL);
            free(stacks);
        }
        seconds[f] = nowSeconds() - t0;
    }
    if (pixels) copyOutPixels(rv, pixels);
    refViewFree(rv);
    refFree(rs);
    return 0;
}

/* Primary-ray hit records, one per traced pixel, from the reference's
   hitPrimitives with the ray set-up of renderSceneOnCPU. For mesh hits the
   triangle index is recovered by re-running traceMesh on the winning
   primitive with an unbounded distance. */
int ref_primary_hits(const void *blob, size_t size, const StView *view, StHitRec *hits) {
    RefScene *rs = refLoad(blob, size);
    if (!rs) return -1;
    RefView *rv = refViewInit(rs, view);
    Scene *scene = &rs->scene;
    Viewport *viewport = &rv->viewport;
    Trace *trace = &viewport->trace;
    Dimensions *dim = &viewport->frame_buffer->dimensions;
    vec3 camera_position = viewport->camera->transform.position;
    vec3 start = viewport->projection_plane.start;
    vec3 right = viewport->projection_plane.right;
    vec3 down = viewport->projection_plane.down;
    vec3 current = start;
    Ray ray;
    u32 i = 0;
    for (u16 y = 0; y < dim->height; y++) {
        for (u16 x = 0; x < dim->width; x++, i++) {
            ray.origin = camera_position;
            ray.direction = normVec3(current);
            ray.direction_reciprocal = oneOverVec3(ray.direction);
            trace->closest_hit.distance = trace->closest_hit.distance_squared = INFINITY;
            trace->closest_hit.cone_angle = viewport->projection_plane.cone_angle;
            trace->closest_hit.cone_width = 0;
            bool found = hitPrimitives(&ray, trace, scene, scene->bvh.leaf_ids, scene->settings.primitives, false, true, x, y);
            hits[i].prim_id = found ? (int32_t)trace->closest_hit.object_id : -1;
            hits[i].object_type = found ? (int32_t)trace->closest_hit.object_type : 0;
            hits[i].distance = found ? trace->closest_hit.distance : INFINITY;
            hits[i].triangle_id = -1;
            if (found && trace->closest_hit.object_type == PrimitiveType_Mesh) {
                Primitive *p = scene->primitives + trace->closest_hit.object_id;
                vec3 *Rd = &trace->local_space_ray.direction;
                vec3 *Ro = &trace->local_space_ray.origin;
                convertPositionAndDirectionToObjectSpace(ray.origin, ray.direction, p, Ro, Rd);
                *Ro = scaleAddVec3(*Rd, TRACE_OFFSET, *Ro);
                trace->local_space_ray.direction_reciprocal = oneOverVec3(*Rd);
                trace->closest_mesh_hit.distance = INFINITY;
                prePrepRay(&trace->local_space_ray);
                if (traceMesh(trace, scene->meshes + p->id, false))
                    hits[i].triangle_id = (int32_t)trace->closest_mesh_hit.object_id;
            }
            current = addVec3(current, right);
        }
        current = start = addVec3(start, down);
    }
    refViewFree(rv);
    refFree(rs);
    return 0;
}

/* Scene BVH from the reference SAH builder (updateSceneBVH). */
int ref_scene_bvh(const void *blob, size_t size, StBVHNodeRec *nodes, uint32_t *leaf_ids,
                  uint32_t *node_count, uint32_t *height) {
    RefScene *rs = refLoad(blob, size);
    if (!rs) return -1;
    BVH *b = &rs->scene.bvh;
    for (u32 n = 0; n < b->node_count; n++) {
        BVHNode *q = b->nodes + n;
        nodes[n].min[0] = q->aabb.min.x; nodes[n].min[1] = q->aabb.min.y; nodes[n].min[2] = q->aabb.min.z;
        nodes[n].max[0] = q->aabb.max.x; nodes[n].max[1] = q->aabb.max.y; nodes[n].max[2] = q->aabb.max.z;
        nodes[n].first_child_id = (uint32_t)q->first_child_id;
        nodes[n].child_count = q->child_count;
        nodes[n].depth = q->depth;
    }
    for (u32 i = 0; i < rs->scene.settings.primitives; i++) leaf_ids[i] = (uint32_t)b->leaf_ids[i];
    *node_count = (uint32_t)b->node_count;
    *height = (uint32_t)b->height;
    refFree(rs);
    return 0;
}

/* Mesh BVH + triangle records of mesh `m` from updateMeshBVH. */
int ref_mesh_bvh(const void *blob, size_t size, uint32_t m, StBVHNodeRec *nodes, StTriangleRec *tris,
                 uint32_t *leaf_ids, uint32_t *node_count, uint32_t *height) {
    RefScene *rs = refLoad(blob, size);
    if (!rs || m >= rs->scene.settings.meshes) { refFree(rs); return -1; }
    Mesh *mesh = rs->scene.meshes + m;
    for (u32 n = 0; n < mesh->bvh.node_count; n++) {
        BVHNode *q = mesh->bvh.nodes + n;
        nodes[n].min[0] = q->aabb.min.x; nodes[n].min[1] = q->aabb.min.y; nodes[n].min[2] = q->aabb.min.z;
        nodes[n].max[0] = q->aabb.max.x; nodes[n].max[1] = q->aabb.max.y; nodes[n].max[2] = q->aabb.max.z;
        nodes[n].first_child_id = (uint32_t)q->first_child_id;
        nodes[n].child_count = q->child_count;
        nodes[n].depth = q->depth;
    }
    for (u32 t = 0; t < mesh->triangle_count; t++) {
        Triangle *q = mesh->triangles + t;
        StTriangleRec *o = tris + t;
        const mat3 *w = &q->world_to_tangent;
        float wt[9] = {w->X.x, w->X.y, w->X.z, w->Y.x, w->Y.y, w->Y.z, w->Z.x, w->Z.y, w->Z.z};
        memcpy(o->world_to_tangent, wt, sizeof wt);
        o->position[0] = q->position.x; o->position[1] = q->position.y; o->position[2] = q->position.z;
        o->normal[0] = q->normal.x; o->normal[1] = q->normal.y; o->normal[2] = q->normal.z;
        for (int k = 0; k < 3; k++) {
            o->vertex_normals[3 * k] = q->vertex_normals[k].x;
            o->vertex_normals[3 * k + 1] = q->vertex_normals[k].y;
            o->vertex_normals[3 * k + 2] = q->vertex_normals[k].z;
            o->uvs[2 * k] = q->uvs[k].x;
            o->uvs[2 * k + 1] = q->uvs[k].y;
        }
        o->area_of_parallelogram = q->area_of_parallelogram;
        o->area_of_uv = q->area_of_uv;
        leaf_ids[t] = (uint32_t)mesh->bvh.leaf_ids[t];
    }
    *node_count = (uint32_t)mesh->bvh.node_count;
    *height = (uint32_t)mesh->bvh.height;
    refFree(rs);
    return 0;
}

/* Screen-space bounds and visibility flags after updateSceneSSB. rects: 4 per primitive. */
int ref_ssb(const void *blob, size_t size, const StView *view, int32_t *rects, uint8_t *flags) {
    RefScene *rs = refLoad(blob, size);
    if (!rs) return -1;
    RefView *rv = refViewInit(rs, view);
    for (u32 i = 0; i < rs->scene.settings.primitives; i++) {
        Primitive *p = rs->scene.primitives + i;
        rects[4 * i + 0] = (int32_t)p->screen_bounds.min.x;
        rects[4 * i + 1] = (int32_t)p->screen_bounds.min.y;
        rects[4 * i + 2] = (int32_t)p->screen_bounds.max.x;
        rects[4 * i + 3] = (int32_t)p->screen_bounds.max.y;
        flags[i] = p->flags;
    }
    refViewFree(rv);
    refFree(rs);
    return 0;
}

/* Picking: manipulateSelection (viewport/manipulation.h:33-240) called with
   the left button freshly pressed at each window point, no modifier and no
   mouse movement, so only its pick branch runs. Each pick starts from the
   hit state initTrace leaves (zeroed hits, closest_mesh_hit typed as mesh). */
int ref_pick(const void *blob, size_t size, const StView *view, uint32_t n, const uint32_t *xy, StPick *out) {
    RefScene *rs = refLoad(blob, size);
    if (!rs) return -1;
    RefView *rv = refViewInit(rs, view);
    Scene *scene = &rs->scene;
    Selection selection;
    Selection *saved = scene->selection;
    scene->selection = &selection;
    Controls controls;
    Trace *t = &rv->viewport.trace;
    for (uint32_t k = 0; k < n; k++) {
        memset(&selection, 0, sizeof selection);
        memset(&controls, 0, sizeof controls);
        memset(&t->closest_hit, 0, sizeof t->closest_hit);
        memset(&t->closest_mesh_hit, 0, sizeof t->closest_mesh_hit);
        memset(&t->current_hit, 0, sizeof t->current_hit);
        memset(&t->sphere_hit, 0, sizeof t->sphere_hit);
        if (scene->settings.meshes) t->closest_mesh_hit.object_type = PrimitiveType_Mesh;
        controls.mouse.pos.x = (i32)xy[2 * k] + rv->viewport.settings.position.x;
        controls.mouse.pos.y = (i32)xy[2 * k + 1] + rv->viewport.settings.position.y;
        controls.mouse.left_button.is_pressed = true;
        manipulateSelection(scene, &rv->viewport, &controls);
        StPick *o = out + k;
        memset(o, 0, sizeof *o);
        if (selection.object_type) {
            o->object_type = (int32_t)selection.object_type;
            o->object_id = selection.object_id;
            o->position[0] = selection.transformation_plane_origin.x;
            o->position[1] = selection.transformation_plane_origin.y;
            o->position[2] = selection.transformation_plane_origin.z;
            o->object_distance = selection.object_distance;
        }
    }
    scene->selection = saved;
    refViewFree(rv);
    refFree(rs);
    return 0;
}

/* Projection plane and camera quaternion/position for a view (A8). out: 3x3 floats
   start,right,down then rotation_inverted quat (axis xyz, amount), then rotation quat. */
int ref_camera(const StView *view, float *out) {
    Camera cam;
    initCamera(&cam);
    cam.transform.position = Vec3(view->position[0], view->position[1], view->position[2]);
    rotateXform3(&cam.transform, view->yaw, view->pitch, view->roll);
    cam.focal_length = view->focal_length;
    PixelGrid frame;
    memset(&frame, 0, sizeof frame);
    frame.QCAA = true;
    updateDimensions(&frame.dimensions, view->width, view->height, true);
    Viewport vp;
    memset(&vp, 0, sizeof vp);
    vp.camera = &cam;
    vp.frame_buffer = &frame;
    setViewportProjectionPlane(&vp);
    float v[17] = {vp.projection_plane.start.x, vp.projection_plane.start.y, vp.projection_plane.start.z,
                   vp.projection_plane.right.x, vp.projection_plane.right.y, vp.projection_plane.right.z,
                   vp.projection_plane.down.x, vp.projection_plane.down.y, vp.projection_plane.down.z,
                   cam.transform.rotation_inverted.axis.x, cam.transform.rotation_inverted.axis.y,
                   cam.transform.rotation_inverted.axis.z, cam.transform.rotation_inverted.amount,
                   cam.transform.rotation.axis.x, cam.transform.rotation.axis.y,
                   cam.transform.rotation.axis.z, cam.transform.rotation.amount};
    memcpy(out, v, sizeof v);
    return 0;
}

/* Quaternion helpers used by the example scenes and the seeded generators:
   normQuat(mulQuat(a, b)) and getRotationAroundAxis (math/quat.h:15-42,132-139). */
void ref_quat_mul_norm(const float *a, const float *b, float *out) {
    quat qa = {{{a[0], a[1], a[2]}}, a[3]}, qb = {{{b[0], b[1], b[2]}}, b[3]};
    quat q = normQuat(mulQuat(qa, qb));
    out[0] = q.axis.x; out[1] = q.axis.y; out[2] = q.axis.z; out[3] = q.amount;
}
void ref_quat_axis(const float *axis, float amount, float *out) {
    quat q = getRotationAroundAxis(Vec3(axis[0], axis[1], axis[2]), amount);
    out[0] = q.axis.x; out[1] = q.axis.y; out[2] = q.axis.z; out[3] = q.amount;
}

/* The reference's display resolve (drawViewportToWindowContent,
   viewport/viewport.h:12-66) over a caller's traced grid. */
int ref_resolve(const StFloatPixel *grid, uint32_t grid_w, uint32_t grid_h, uint32_t *out) {
    if (grid_w < 2 || grid_h < 2 || grid_w > 0xFFFF || grid_h > 0xFFFF) return -1;
    PixelGrid frame;
    memset(&frame, 0, sizeof frame);
    frame.QCAA = true;
    frame.dimensions.width = (u16)grid_w;
    frame.dimensions.height = (u16)grid_h;
    frame.dimensions.width_times_height = grid_w * grid_h;
    frame.float_pixels = (FloatPixel *)calloc((size_t)grid_w * grid_h, sizeof(FloatPixel));
    /* Pixel holds a u32, which is 8 bytes in an LP64 build of the reference:
       resolve into the reference's own Pixel array, then keep the RGBA bytes */
    size_t npx = (size_t)(grid_w - 1) * (grid_h - 1);
    frame.pixels = (Pixel *)calloc(npx, sizeof(Pixel));
    for (size_t i = 0; i < (size_t)grid_w * grid_h; i++) {
        frame.float_pixels[i].color = Vec3(grid[i].r, grid[i].g, grid[i].b);
        frame.float_pixels[i].opacity = grid[i].opacity;
        frame.float_pixels[i].depth = grid[i].depth;
    }
    Viewport vp;
    memset(&vp, 0, sizeof vp);
    vp.frame_buffer = &frame;
    drawViewportToWindowContent(&vp);
    for (size_t i = 0; i < npx; i++) {
        RGBA c = frame.pixels[i].color;
        out[i] = (uint32_t)c.B | ((uint32_t)c.G << 8) | ((uint32_t)c.R << 16) | ((uint32_t)c.A << 24);
    }
    free(frame.pixels);
    free(frame.float_pixels);
    return 0;
}
