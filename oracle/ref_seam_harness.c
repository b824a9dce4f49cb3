/*
 * ref_seam_harness.c - the reference's own renderer driving the MI355X
 * library through integration/init_amd.h.
 *
 * TEST INFRASTRUCTURE ONLY. Compiled with the reference's headers, read in
 * place from the reference checkout (never copied), and the reference-side
 * binding integration/init_amd.h into oracle/_ref/libslimref_seam.so, linked
 * against libslimtracin.so. It builds the C version of the reference the way
 * init_amd.h tells a maintainer to patch it:
 *   - core/init.h's branch without __CUDACC__ defines empty device-mirror
 *     stubs; here they compile under other names and init_amd.h supplies
 *     allocateDeviceScene, upload* and renderSceneOnGPU;
 *   - renderScene (render/raytracer.h:302-309) picks renderSceneOnGPU when
 *     viewport->settings.use_GPU is set, as its __CUDACC__ branch does.
 * Then it replays _initApp's order (app.h:160-199, 285-298): the scene is
 * loaded (ref_harness.c's refLoad), allocateDeviceScene and uploadScene run
 * before the BVH builds, updateSceneBVH uploads the scene BVH itself
 * (builder_top_down.h:380), uploadMeshBVHs follows, and the window set-up's
 * updateSceneSSB uploads the primitives (render/SSB.h:172). A frame is then
 * rendered twice by the reference's renderScene: use_GPU set (through the
 * library) and clear (renderSceneOnCPU).
 */
#include <stdio.h>

static int amd_harness_rc;
static char amd_harness_what[96];
static void amdHarnessFail(int rc, const char *what) {
    if (!amd_harness_rc) {
        amd_harness_rc = rc;
        snprintf(amd_harness_what, sizeof amd_harness_what, "%s", what);
    }
}
/* record the first failure and leave the seam function (all are void) */
#define SLIMTRACIN_AMD_ON_ERROR(rc, what) \
    do {                                  \
        amdHarnessFail((rc), (what));     \
        return;                           \
    } while (0)

/* the C build's stubs of the device mirror, under other names */
#define allocateDeviceScene refCStub_allocateDeviceScene
#define uploadLights refCStub_uploadLights
#define uploadPrimitives refCStub_uploadPrimitives
#define uploadMaterials refCStub_uploadMaterials
#define uploadScene refCStub_uploadScene
#define uploadMeshBVHs refCStub_uploadMeshBVHs
#define uploadSceneBVH refCStub_uploadSceneBVH
#define renderOnGPU refCStub_renderOnGPU
#include "core/init.h"
#undef allocateDeviceScene
#undef uploadLights
#undef uploadPrimitives
#undef uploadMaterials
#undef uploadScene
#undef uploadMeshBVHs
#undef uploadSceneBVH
#undef renderOnGPU

#include "init_amd.h"
#include "app.h"

/* renderScene as the patched guard compiles it (the __CUDACC__ branch's
   choice by use_GPU) */
#define renderScene refCBuild_renderScene
#include "render/raytracer.h"
#undef renderScene
void renderScene(Scene *scene, Viewport *viewport) {
    if (viewport->settings.use_GPU) renderSceneOnGPU(scene, viewport);
    else renderSceneOnCPU(scene, viewport);
}

#define REF_HARNESS_BEFORE_BVH(s) \
    do {                          \
        allocateDeviceScene(s);   \
        uploadScene(s);           \
    } while (0)
#include "ref_harness.c"

const char *ref_seam_error(void) { return amd_harness_what; }

/* One frame of a scene blob, rendered by the reference's renderScene with
   use_GPU set into gpu[] and clear into cpu[] ((width+1)*(height+1) records
   each). frames > 1: after each frame the primitives from index 1 on turn
   as examples/02_Geometry.c:70-75 turns them (rotation = normQuat(rotation *
   step)) and only uploadPrimitives runs (the scene BVH and the SSB rects
   stay as they were), and the next frame is rendered both ways; the last
   frame's pixels are returned. Returns 0, or the library's status (see
   ref_seam_error). */
int ref_seam_render(const void *blob, size_t size, const StView *view, int frames, StFloatPixel *gpu,
                    StFloatPixel *cpu) {
    amd_harness_rc = 0;
    amd_harness_what[0] = 0;
    RefScene *rs = refLoad(blob, size);
    if (!rs) return -100;
    Scene *s = &rs->scene;
    if (!amd_harness_rc) uploadMeshBVHs(s);
    RefView *rv = refViewInit(rs, view);
    const quat step = normQuat(Quat(normVec3(Vec3(0.3f, 1.0f, 0.2f)), 0.05f));
    for (int f = 0; f < frames && !amd_harness_rc; f++) {
        if (f) {
            for (u32 i = 1; i < s->settings.primitives; i++)
                s->primitives[i].rotation = normQuat(mulQuat(s->primitives[i].rotation, step));
            uploadPrimitives(s);
        }
        rv->viewport.settings.use_GPU = true;
        beginDrawing(&rv->viewport);
        renderScene(s, &rv->viewport);
        copyOutPixels(rv, gpu);
        rv->viewport.settings.use_GPU = false;
        beginDrawing(&rv->viewport);
        renderScene(s, &rv->viewport);
        copyOutPixels(rv, cpu);
    }
    int rc = amd_harness_rc;
    refViewFree(rv);
    refFree(rs);
    if (amd_scene) {
        st_scene_destroy(amd_scene);
        amd_scene = NULL;
    }
    return rc;
}
