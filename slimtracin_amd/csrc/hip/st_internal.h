/*
 * st_internal.h - what the library's translation units share: the resident
 * scene handle and the launch entry over prepared host views.
 */
#ifndef ST_INTERNAL_H
#define ST_INTERNAL_H

#include <hip/hip_runtime.h>

#include "../host/st_host.h"
#include "slimtracin.h"

struct StSeamState;

struct StSceneHandle {
    StScene *host;
    StDScene dev;
    char *dbuf;
    uint64_t dbytes;
    float *d_mesh_aabbs;       /* 6 floats per mesh, for device BVH rebuilds */
    char *lbvh_work;           /* scratch of device LBVH rebuilds */
    size_t lbvh_work_bytes;    /* its size */
    uint32_t *d_height;        /* height written by the device builder */
    int bvh_on_device_only;    /* host copy of the scene BVH is stale */
    struct StSeamState *seam;  /* host copies of a scene fed through slimtracin_seam.h */
};

int st_rt_ensure_init(void);
/* The runtime's lock over the shared stream and buffers (recursive). */
void st_rt_lock(void);
void st_rt_unlock(void);
hipStream_t st_rt_stream(void);
/* Record `msg` as the last error and return `code`. */
int st_rt_error(int code, const char *msg);

/* One launch over prepared views on the library's main context (holds the
   library lock). */
int st_rt_launch_prepared(StSceneHandle *const *scenes, StHostView *hv, uint32_t n, int want_hits, int persistent,
                          StRayCounts *counts, StFloatPixel *const *pixels_host, StHitRec *const *hits_host,
                          uint32_t *const *bgra_host, float *kernel_ms);

/* A launch context of its own (stream, per-launch buffers, staging, tile
   history): launches on different contexts run concurrently on the device.
   A context serves one launch at a time; its owner calls st_rt_launch_ctx
   from one thread at a time, without the library lock. */
struct LaunchCtx;
int st_rt_ctx_create(LaunchCtx **out);
void st_rt_ctx_destroy(LaunchCtx *c);
int st_rt_launch_ctx(LaunchCtx *c, StSceneHandle *const *scenes, StHostView *hv, uint32_t n, int want_hits,
                     int persistent, StRayCounts *counts, StFloatPixel *const *pixels_host,
                     StHitRec *const *hits_host, uint32_t *const *bgra_host, float *kernel_ms);

void st_seam_state_free(struct StSeamState *s);

#endif
