/*
 * st_server.hip - serving loop (include/slimtracin_server.h).
 *
 * Requests are queued FIFO with a ticket. The server runs max_in_flight
 * workers, each with a launch context of its own (stream, per-launch
 * buffers, pinned staging): a free worker takes everything queued (up to
 * the batch limits), prepares each request's camera set-up on the host and
 * runs the batch as one persistent launch on its context (one tile list over
 * all requests, one upload transfer, device display resolve, downloads into
 * pinned per-ticket buffers). Requests that arrive while launches run are
 * taken by the next free worker and launched at once on its own stream, so
 * their work overlaps on the device with the launches already there; a
 * launch that runs long holds only its own requests. Scenes are admitted
 * against an HBM byte budget.
 */
#include <hip/hip_runtime.h>

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "st_internal.h"
#include "slimtracin_server.h"

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

/* Pinned host buffers recycled by size class (powers of two from 64 KiB). */
class PinnedPool {
  public:
    ~PinnedPool() {
        for (auto &kv : free_)
            for (void *p : kv.second) (void)hipHostFree(p);
    }
    void *get(size_t bytes, size_t *cls_out) {
        size_t cls = 65536;
        while (cls < bytes) cls <<= 1;
        *cls_out = cls;
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto &v = free_[cls];
            if (!v.empty()) {
                void *p = v.back();
                v.pop_back();
                return p;
            }
        }
        void *p = NULL;
        if (hipHostMalloc(&p, cls, hipHostMallocDefault) != hipSuccess) return NULL;
        return p;
    }
    void put(void *p, size_t cls) {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu_);
        free_[cls].push_back(p);
    }

  private:
    std::mutex mu_;
    std::map<size_t, std::vector<void *>> free_;
};

struct Ticket {
    uint32_t scene = 0;
    StView view;
    int status = ST_OK;
    bool done = false;
    StRayCounts counts = {0, 0, 0};
    double t_submit = 0, t_done = 0;
    void *buf = NULL;
    size_t cls = 0, bytes = 0;
};

struct Resident {
    StSceneHandle *h = NULL;
    uint64_t bytes = 0;
    uint32_t users = 0;     /* queued or in-flight requests */
    bool removing = false;
};

}  // namespace

struct StServer {
    StServerConfig cfg;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, Ticket> tickets;
    uint64_t next_ticket = 1;
    std::vector<Resident> scenes;
    uint64_t resident = 0;
    std::vector<std::thread> workers;
    std::vector<LaunchCtx *> ctxs;    /* one per worker */
    uint32_t batch_limit = 0;         /* st_server_set_limits: requests per launch */
    uint32_t active_limit = 0;        /* st_server_set_limits: workers that take work */
    uint32_t in_flight = 0, max_in_flight_seen = 0;
    bool stop = false;
    PinnedPool pool;
    uint64_t submitted = 0, completed = 0, launches = 0, rays = 0, batched = 0, overlapped = 0;
    double prep_ms = 0, launch_ms = 0, kernel_ms = 0, alloc_ms = 0;
    std::vector<float> latencies;     /* ms, completed requests */
};

static int fail(int code, const char *msg) { return st_rt_error(code, msg); }

static size_t output_bytes(const StServerConfig &cfg, const StView &v) {
    if (cfg.output == ST_OUTPUT_FLOAT_PIXELS) return (size_t)(v.width + 1) * (v.height + 1) * 24;
    return (size_t)v.width * v.height * 4;
}

static void worker_loop(StServer *srv, uint32_t index) {
    LaunchCtx *ctx = srv->ctxs[index];
    const uint64_t max_px = srv->cfg.max_batch_pixels;
    for (;;) {
        std::vector<uint64_t> ids;
        std::vector<StSceneHandle *> handles;
        std::vector<StView> views;
        std::vector<void *> bufs;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_work.wait(lk, [&] {
                return (srv->stop && (srv->queue.empty() || index >= srv->active_limit)) ||
                       (!srv->queue.empty() && index < srv->active_limit);
            });
            if (srv->queue.empty() || index >= srv->active_limit) break;  /* stop requested and drained */
            const uint32_t max_req = srv->batch_limit;
            uint64_t px = 0;
            while (!srv->queue.empty() && ids.size() < max_req) {
                Ticket &t = srv->tickets[srv->queue.front()];
                uint64_t p = (uint64_t)(t.view.width + 1) * (t.view.height + 1);
                if (!ids.empty() && px + p > max_px) break;
                ids.push_back(srv->queue.front());
                srv->queue.pop_front();
                handles.push_back(srv->scenes[t.scene].h);
                views.push_back(t.view);
                bufs.push_back(t.buf);
                px += p;
            }
            if (srv->in_flight) srv->overlapped++;
            srv->in_flight++;
            if (srv->in_flight > srv->max_in_flight_seen) srv->max_in_flight_seen = srv->in_flight;
        }
        /* another worker may take what this one left queued. Every waiter is
           woken: with st_server_set_limits below the worker count only
           workers 0..active_limit-1 may take work, and a single wake-up
           could go to one that may not, which would leave the queue
           waiting for the next submission (for the last request of a
           stream of arrivals, forever). */
        srv->cv_work.notify_all();
        const size_t n = ids.size();
        const double t_prep = now_ms();
        std::vector<StHostView> hv(n);
        std::vector<int> status(n, ST_OK);
        std::vector<uint32_t> live;
        for (size_t i = 0; i < n; i++) {
            if (!bufs[i]) status[i] = fail(ST_ERR_MEMORY, "no pinned output buffer");
            else if (st_view_prepare_ex(handles[i]->host, &views[i], &hv[i], 0))
                status[i] = fail(ST_ERR_ARG, "invalid view");
            else live.push_back((uint32_t)i);
        }
        std::vector<StRayCounts> counts(n);
        int rc = ST_OK;
        const double t_launch = now_ms();
        float kms = 0;
        if (!live.empty()) {
            const size_t m = live.size();
            std::vector<StSceneHandle *> lh(m);
            std::vector<StHostView> lv(m);
            std::vector<StRayCounts> lc(m);
            std::vector<StFloatPixel *> fp(m, NULL);
            std::vector<uint32_t *> bp(m, NULL);
            for (size_t k = 0; k < m; k++) {
                lh[k] = handles[live[k]];
                lv[k] = hv[live[k]];
                if (srv->cfg.output == ST_OUTPUT_FLOAT_PIXELS) fp[k] = (StFloatPixel *)bufs[live[k]];
                else bp[k] = (uint32_t *)bufs[live[k]];
            }
            rc = st_rt_launch_ctx(ctx, lh.data(), lv.data(), (uint32_t)m, 0, 1, lc.data(),
                                  srv->cfg.output == ST_OUTPUT_FLOAT_PIXELS ? fp.data() : NULL, NULL,
                                  srv->cfg.output == ST_OUTPUT_FLOAT_PIXELS ? NULL : bp.data(), &kms);
            for (size_t k = 0; k < m; k++) {
                counts[live[k]] = lc[k];
                if (rc) status[live[k]] = rc;
            }
        }
        for (size_t i = 0; i < n; i++) st_view_release(&hv[i]);
        const double t = now_ms();
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->in_flight--;
            if (!live.empty()) {
                srv->launches++;
                srv->batched += live.size();
                srv->prep_ms += t_launch - t_prep;
                srv->launch_ms += t - t_launch;
                srv->kernel_ms += kms;
            }
            for (size_t i = 0; i < n; i++) {
                Ticket &tk = srv->tickets[ids[i]];
                tk.status = status[i];
                tk.counts = counts[i];
                tk.t_done = t;
                tk.done = true;
                srv->scenes[tk.scene].users--;
                srv->completed++;
                srv->rays += counts[i].primary + counts[i].shadow + counts[i].bounce;
                if (srv->latencies.size() < (1u << 22)) srv->latencies.push_back((float)(t - tk.t_submit));
            }
        }
        srv->cv_done.notify_all();
    }
}

extern "C" int st_server_create(const StServerConfig *cfg, StServer **out) {
    if (!out) return fail(ST_ERR_ARG, "null argument");
    *out = NULL;
    int rc = st_rt_ensure_init();
    if (rc) return rc;
    StServer *srv = new StServer();
    if (cfg) srv->cfg = *cfg;
    else memset(&srv->cfg, 0, sizeof srv->cfg);
    if (srv->cfg.output > ST_OUTPUT_FLOAT_PIXELS) {
        delete srv;
        return fail(ST_ERR_ARG, "unknown output kind");
    }
    if (!srv->cfg.max_batch_requests) srv->cfg.max_batch_requests = 256;
    if (!srv->cfg.max_batch_pixels) srv->cfg.max_batch_pixels = 64ull << 20;
    if (!srv->cfg.max_in_flight) srv->cfg.max_in_flight = 3;
    if (srv->cfg.max_in_flight > ST_SERVER_MAX_IN_FLIGHT) {
        delete srv;
        return fail(ST_ERR_ARG, "max_in_flight above ST_SERVER_MAX_IN_FLIGHT");
    }
    srv->batch_limit = srv->cfg.max_batch_requests;
    srv->active_limit = srv->cfg.max_in_flight;
    if (!srv->cfg.hbm_budget) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            delete srv;
            return fail(ST_ERR_HIP, "hipMemGetInfo failed");
        }
        srv->cfg.hbm_budget = (uint64_t)(free_b * 0.85);
    }
    for (uint32_t i = 0; i < srv->cfg.max_in_flight; i++) {
        LaunchCtx *c = NULL;
        if ((rc = st_rt_ctx_create(&c))) {
            for (LaunchCtx *x : srv->ctxs) st_rt_ctx_destroy(x);
            delete srv;
            return rc;
        }
        srv->ctxs.push_back(c);
    }
    for (uint32_t i = 0; i < srv->cfg.max_in_flight; i++) srv->workers.emplace_back(worker_loop, srv, i);
    *out = srv;
    return ST_OK;
}

extern "C" int st_server_set_limits(StServer *srv, uint32_t max_batch_requests, uint32_t max_in_flight) {
    if (!srv) return fail(ST_ERR_ARG, "null server");
    if (max_in_flight > srv->cfg.max_in_flight) return fail(ST_ERR_ARG, "max_in_flight above the server's");
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->batch_limit = max_batch_requests ? max_batch_requests : srv->cfg.max_batch_requests;
        srv->active_limit = max_in_flight ? max_in_flight : srv->cfg.max_in_flight;
    }
    srv->cv_work.notify_all();
    return ST_OK;
}

extern "C" void st_server_destroy(StServer *srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto &w : srv->workers)
        if (w.joinable()) w.join();
    for (LaunchCtx *c : srv->ctxs) st_rt_ctx_destroy(c);
    for (auto &kv : srv->tickets) srv->pool.put(kv.second.buf, kv.second.cls);
    for (auto &r : srv->scenes)
        if (r.h) st_scene_destroy(r.h);
    delete srv;
}

extern "C" int st_server_add_scene(StServer *srv, const void *blob, size_t size, uint32_t *scene_id) {
    return st_server_add_scene_ex(srv, blob, size, 0, scene_id);
}

extern "C" int st_server_add_scene_ex(StServer *srv, const void *blob, size_t size, uint32_t flags,
                                      uint32_t *scene_id) {
    if (!srv || !blob || !scene_id) return fail(ST_ERR_ARG, "null argument");
    const char *why = NULL;
    /* size the scene without building any BVH (mesh builds are deferred) */
    StScene *probe = st_scene_from_blob_ex(blob, size, ST_MESH_BUILD_DEVICE, &why);
    if (!probe) return fail(ST_ERR_SCENE, why ? why : "invalid scene");
    uint64_t need = st_scene_device_bytes(probe) + 64 * 1024;
    st_scene_free(probe);
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->resident + need > srv->cfg.hbm_budget) return fail(ST_ERR_BUSY, "scene does not fit the HBM budget");
        srv->resident += need;   /* reserve while building */
    }
    StSceneHandle *h = NULL;
    int rc = st_scene_create_ex(blob, size, flags, &h);
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->resident -= need;
    if (rc) return rc;
    uint64_t bytes = h->dbytes;
    if (srv->resident + bytes > srv->cfg.hbm_budget) {
        st_scene_destroy(h);
        return fail(ST_ERR_BUSY, "scene does not fit the HBM budget");
    }
    srv->resident += bytes;
    uint32_t id = 0;
    while (id < srv->scenes.size() && srv->scenes[id].h) id++;
    if (id == srv->scenes.size()) srv->scenes.emplace_back();
    srv->scenes[id].h = h;
    srv->scenes[id].bytes = bytes;
    srv->scenes[id].users = 0;
    srv->scenes[id].removing = false;
    *scene_id = id;
    return ST_OK;
}

extern "C" int st_server_remove_scene(StServer *srv, uint32_t scene_id) {
    if (!srv) return fail(ST_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    if (scene_id >= srv->scenes.size() || !srv->scenes[scene_id].h || srv->scenes[scene_id].removing)
        return fail(ST_ERR_ARG, "no such scene");
    srv->scenes[scene_id].removing = true;
    srv->cv_done.wait(lk, [&] { return srv->scenes[scene_id].users == 0; });
    StSceneHandle *h = srv->scenes[scene_id].h;
    srv->resident -= srv->scenes[scene_id].bytes;
    srv->scenes[scene_id] = Resident();
    lk.unlock();
    st_scene_destroy(h);
    return ST_OK;
}

extern "C" int st_server_submit(StServer *srv, uint32_t scene_id, const StView *view, uint64_t *ticket) {
    if (!srv || !view || !ticket) return fail(ST_ERR_ARG, "null argument");
    if (view->width == 0 || view->height == 0 || view->width >= 0xFFFF || view->height >= 0xFFFF ||
        view->render_mode > ST_MODE_UVS)
        return fail(ST_ERR_ARG, "invalid view");
    Ticket t;
    t.scene = scene_id;
    t.view = *view;
    t.bytes = output_bytes(srv->cfg, *view);
    t.t_submit = now_ms();
    t.buf = srv->pool.get(t.bytes, &t.cls);
    if (!t.buf) return fail(ST_ERR_MEMORY, "pinned output allocation failed");
    const double t_alloc = now_ms() - t.t_submit;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (scene_id >= srv->scenes.size() || !srv->scenes[scene_id].h || srv->scenes[scene_id].removing) {
            srv->pool.put(t.buf, t.cls);
            return fail(ST_ERR_ARG, "no such scene");
        }
        srv->scenes[scene_id].users++;
        uint64_t id = srv->next_ticket++;
        srv->tickets.emplace(id, t);
        srv->queue.push_back(id);
        srv->submitted++;
        srv->alloc_ms += t_alloc;
        *ticket = id;
    }
    /* all waiters, so that an eligible worker sees the request (see the
       worker loop's wake-up) */
    srv->cv_work.notify_all();
    return ST_OK;
}

extern "C" int st_server_poll(StServer *srv, uint64_t ticket, int *ready) {
    if (!srv || !ready) return fail(ST_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return fail(ST_ERR_ARG, "unknown ticket");
    *ready = it->second.done ? 1 : 0;
    return ST_OK;
}

extern "C" int st_server_wait(StServer *srv, uint64_t ticket, int timeout_ms, const void **image, size_t *bytes,
                              StRayCounts *counts, double *latency_ms) {
    if (!srv) return fail(ST_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return fail(ST_ERR_ARG, "unknown ticket");
    auto ready = [&] {
        auto f = srv->tickets.find(ticket);
        return f == srv->tickets.end() || f->second.done;
    };
    if (timeout_ms < 0) srv->cv_done.wait(lk, ready);
    else if (!srv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return fail(ST_ERR_BUSY, "timed out");
    auto f = srv->tickets.find(ticket);
    if (f == srv->tickets.end()) return fail(ST_ERR_ARG, "ticket released while waiting");
    const Ticket &t = f->second;
    if (image) *image = t.buf;
    if (bytes) *bytes = t.bytes;
    if (counts) *counts = t.counts;
    if (latency_ms) *latency_ms = t.t_done - t.t_submit;
    return t.status;
}

extern "C" int st_server_release(StServer *srv, uint64_t ticket) {
    if (!srv) return fail(ST_ERR_ARG, "null server");
    Ticket t;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->tickets.find(ticket);
        if (it == srv->tickets.end()) return fail(ST_ERR_ARG, "unknown ticket");
        if (!it->second.done) return fail(ST_ERR_BUSY, "ticket still in flight");
        t = it->second;
        srv->tickets.erase(it);
    }
    srv->pool.put(t.buf, t.cls);
    return ST_OK;
}

extern "C" int st_server_stats(StServer *srv, StServerStats *s) {
    if (!srv || !s) return fail(ST_ERR_ARG, "null argument");
    memset(s, 0, sizeof *s);
    std::vector<float> lat;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        s->submitted = srv->submitted;
        s->completed = srv->completed;
        s->launches = srv->launches;
        s->rays = srv->rays;
        s->resident_bytes = srv->resident;
        s->budget_bytes = srv->cfg.hbm_budget;
        for (auto &r : srv->scenes) s->resident_scenes += r.h ? 1 : 0;
        s->queued = (uint32_t)srv->queue.size();
        s->mean_batch = srv->launches ? (double)srv->batched / srv->launches : 0.0;
        s->prep_ms = srv->prep_ms;
        s->launch_ms = srv->launch_ms;
        s->kernel_ms = srv->kernel_ms;
        s->alloc_ms = srv->alloc_ms;
        s->overlapped_launches = srv->overlapped;
        s->max_in_flight = srv->max_in_flight_seen;
        lat = srv->latencies;
    }
    if (!lat.empty()) {
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double p) {
            size_t i = (size_t)(p * (lat.size() - 1) + 0.5);
            return (double)lat[std::min(i, lat.size() - 1)];
        };
        s->latency_p50_ms = pct(0.50);
        s->latency_p99_ms = pct(0.99);
        s->latency_max_ms = lat.back();
    }
    return ST_OK;
}
