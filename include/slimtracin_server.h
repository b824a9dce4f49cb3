/*
 * slimtracin_server.h - serving loop over the tracing path: resident scenes
 * admitted against an HBM budget, render requests queued with tickets, and
 * continuous batching: up to max_in_flight launches run at once, each on a
 * launch context of its own (stream, buffers); a free worker fuses
 * everything queued into one persistent launch, so requests that arrive
 * while launches run start at once on another stream and their work
 * overlaps on the device with the launches already there.
 *
 * The reference has no server: each frame is one blocking renderSceneOnGPU
 * call from the window loop (render/raytracer.h:258-299, win32.h:296-303).
 * A request here is exactly such a frame - its pixels equal what st_render
 * returns for the same scene and view - delivered as the window image
 * (drawViewportToWindowContent, viewport/viewport.h:12-66, done on the
 * device) or as the traced FloatPixel grid.
 *
 * Thread safety: every function may be called from any thread. The
 * server's launches run on their own streams; launches made directly with
 * slimtracin.h run on the library's main stream and may overlap with them.
 */
#ifndef SLIMTRACIN_SERVER_H
#define SLIMTRACIN_SERVER_H

#include <stddef.h>
#include <stdint.h>

#include "slimtracin.h"
#include "st_blob.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct StServer StServer;

#define ST_OUTPUT_BGRA8 0u         /* width*height BGRA8 (4 B per pixel) */
#define ST_OUTPUT_FLOAT_PIXELS 1u  /* (width+1)*(height+1) FloatPixels (24 B per pixel) */

#define ST_SERVER_MAX_IN_FLIGHT 8

typedef struct StServerConfig {
    uint64_t hbm_budget;           /* bytes of HBM for resident scenes; 0 = 85% of free memory at creation */
    uint32_t max_batch_requests;   /* requests fused into one launch; 0 = 256 */
    uint32_t output;               /* ST_OUTPUT_* */
    uint64_t max_batch_pixels;     /* traced pixels fused into one launch; 0 = 64M */
    uint32_t max_in_flight;        /* launches in flight at once, each with its own stream and launch
                                      buffers (grown to the largest batch it ran); 0 = 3, at most
                                      ST_SERVER_MAX_IN_FLIGHT */
    uint32_t reserved;
} StServerConfig;

typedef struct StServerStats {
    uint64_t submitted, completed, launches, rays;
    uint64_t resident_bytes, budget_bytes;
    uint32_t resident_scenes, queued;
    double latency_p50_ms, latency_p99_ms, latency_max_ms;  /* submit -> frame ready, completed requests */
    double mean_batch;                                       /* requests per launch */
    double prep_ms, launch_ms, kernel_ms, alloc_ms;          /* totals: host set-up, launch wall (uploads,
                                                                kernels, downloads), device work of the launches, output
                                                                buffer allocation at submit */
    uint64_t overlapped_launches;  /* launches started while another launch was in flight */
    uint32_t max_in_flight;        /* most launches in flight at once so far */
    uint32_t reserved;
} StServerStats;

int st_server_create(const StServerConfig *cfg, StServer **out);
/* Change the batching limits of a running server: requests fused into one
   launch and launches in flight (0: the configured value; max_in_flight at
   most the configured one). (1, 1) serves one request per launch, one
   launch at a time: plain per-request launches over the same resident
   scenes. */
int st_server_set_limits(StServer *srv, uint32_t max_batch_requests, uint32_t max_in_flight);
/* Finishes queued requests, then frees everything the server holds. */
void st_server_destroy(StServer *srv);

/* Admit a scene (st_scene_create) if it fits the remaining HBM budget;
   ST_ERR_BUSY when it does not (remove scenes first). */
int st_server_add_scene(StServer *srv, const void *blob, size_t size, uint32_t *scene_id);
/* The same with st_scene_create_ex flags (ST_SCENE_MESH_LBVH: build meshes
   on the device, for scenes of millions of triangles). */
int st_server_add_scene_ex(StServer *srv, const void *blob, size_t size, uint32_t flags, uint32_t *scene_id);
/* Waits until no queued request uses the scene, then frees it. */
int st_server_remove_scene(StServer *srv, uint32_t scene_id);

/* Queue one request; returns at once. */
int st_server_submit(StServer *srv, uint32_t scene_id, const StView *view, uint64_t *ticket);
/* *ready = 1 once the ticket's frame is done (or failed). */
int st_server_poll(StServer *srv, uint64_t ticket, int *ready);
/* Block until the frame is ready (timeout_ms < 0: no limit; ST_ERR_BUSY on
   timeout). *image points into server-owned pinned memory, valid until
   st_server_release; counts and latency_ms may be NULL. Returns the
   request's own status if it failed. */
int st_server_wait(StServer *srv, uint64_t ticket, int timeout_ms, const void **image, size_t *bytes,
                   StRayCounts *counts, double *latency_ms);
int st_server_release(StServer *srv, uint64_t ticket);
int st_server_stats(StServer *srv, StServerStats *stats);

#ifdef __cplusplus
}
#endif
#endif
