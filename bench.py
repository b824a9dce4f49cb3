#!/usr/bin/env python3
"""Benchmark of the MI355X tracing path (one JSON line on stdout).

Headline workload (one step = one launch): the batched request mix - 256
render requests over 8 resident scenes of 0.2-5M triangles (4 of the
single-mesh PBR type: a noise-displaced icosphere with Cook-Torrance
materials, a floor quad, a glass sphere, a mirror quad, 4 point lights and
2 emissive quads; 4 of 64 instances of 3 textured meshes, sampled with ray
cones; mesh BVHs built on the host by the binned SAH builder),
resolutions log-uniform 128x128..1920x1080, trace depth uniform 1-6, mode
uniform over Beauty/Depth/Normal/UV, cameras on a sphere around each scene
(seed 7). A step fuses all 256 requests into one persistent launch: the
per-request host set-up (camera basis, SSB rects, row starts), one upload,
and every stage kernel; the frames stay in HBM. `value` = rays of the 256
requests / wall time of the step.

Beside it, in the same run:
  - plain launches ("batched.plain_*"): the same 256 requests as 256
    sequential single-request launches of the same kernels, without
    downloads (like for like with the headline) and with blocking
    FloatPixel downloads into page-locked memory against the batched launch
    with the same downloads; the batched frames are checked equal to the
    plain ones byte for byte;
  - parity of the headline workload against the CPU oracle: every hit id,
    triangle id and hit distance of all 256 requests bit for bit, and every
    request's ray counts;
  - the reference CPU renderer (compiled from the reference's own sources,
    or the oracle restatement when that build is absent) on a seeded
    8-request subset of the same mix, all host cores and one pinned core,
    with the CPU model;
  - "frame_c1": the 640x480 implicit-primitive frame (16 seeded quads,
    boxes, spheres and tetrahedra, 3 point lights, Blinn-Phong, depth 2),
    GPU and CPU;
  - "frame_c2": the single-mesh PBR frame (200,000-triangle icosphere,
    1920x1080, depth 4), with per-stage times, HIP-graph replays and the
    PCIe-inclusive frame;
  - "instanced" (C3): 64 instances of 50k/200k/500k-triangle meshes moving
    every frame, scene LBVH rebuilt on the device per frame, 2048^2 Perlin
    albedo/normal maps with mips, ray cones, depth 3;
  - "served" (C5): 40 resident scenes of 1-20M triangles admitted against a
    200 GB HBM budget, Poisson arrivals of the same request mix served with
    continuous batching; Mrays/s with p50/p99 submit-to-image latency, and a
    plain one-request-per-launch baseline on the same scenes;
  - roofline of the headline step's dominant kernel: algorithmic bytes from
    the CPU oracle's per-ray-class event counters over all 256 requests
    (bytes per event at the reference's record sizes) over the kernel's
    mean duration from HIP events recorded around each stage launch on the
    launch stream during the timed steps; HBM traffic per launch from the
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary in profiles/.
Every leg picks its own launch shape with st_autotune on its own requests
(the headline unless --fixed-shape), as a caller tuning once per
configuration would; the frame legs put the headline's shape back when
they end, so the served leg runs at it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import slimtracin_amd as st  # noqa: E402
from slimtracin_amd.scene import MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS, make_view  # noqa: E402

HBM_PEAK_GBS = 8000.0
PBR_K = 100            # 20 * 100^2 = 200,000 triangles
PBR_SEED = 2024
PBR_CAM = dict(position=(0.0, 6.0, -14.0), pitch=-0.15)
# The 8 resident scenes of the batched workload: 4 of the single-mesh PBR
# type (noise-displaced icosphere of 20k^2 triangles, GGX materials, glass
# sphere, mirror quad, 4 point lights, 2 emissive quads) and 4 of the
# instanced textured type (64 instances of 3 noise-displaced meshes, Perlin
# albedo and normal maps with mip chains, sampled with ray cones).
BATCH_PBR_K = [100, 200, 316, 500]                       # 200k, 800k, ~2.0M, 5.0M triangles
BATCH_INSTANCED = [[50_000, 100_000, 150_000], [100_000, 200_000, 500_000],
                   [200_000, 500_000, 1_000_000], [500_000, 1_000_000, 2_000_000]]
BATCH_TEXTURE = 1024
# The headline scenes are built once and stay resident: their meshes get the
# host binned SAH builder (SAH-quality trees with one triangle per leaf,
# about half a second per million triangles; the headline step takes 14% less
# time than over leaves of 4 and 25% less than over the device Morton LBVH).
# The served leg admits its scenes at run time and builds them on the device.
BATCH_MESH_BUILD = {"mesh_binned": True}
BATCH_MESH_LABEL = "host binned SAH (32 centroid bins per axis, one triangle per leaf)"
# The instanced frame's three meshes too (its scene BVH is the per-frame
# device LBVH); the 200k-triangle frame keeps the reference SAH builder.
INSTANCED_MESH_BUILD = BATCH_MESH_BUILD
BATCH_CONE = 0.001                                       # ray-cone angle of requests on the textured scenes
BATCH_SEED = 7
BATCH_REQUESTS = 256
CPU_SUBSET = 8
LEGS = ("c4", "c1", "c2", "c3", "c5", "cpu")
METRIC = "Mrays/s + p50/p99 frame latency, batched/served vs plain launches, hit-IDs exact"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    return len(os.sched_getaffinity(0))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def request_mix(rng, nscenes, count):
    """Mixed requests: log-uniform resolution 128x128..1920x1080, depth 1-6,
    mode uniform over Beauty/Depth/Normal/UV, camera on a sphere around the
    scene (azimuth uniform over +-90 degrees, elevation uniform over 8-50
    degrees, 14 units from the centre) looking at its centre."""
    views = []
    scene_ids = []
    target_y, dist = 1.5, 14.0
    for i in range(count):
        w = int(round(np.exp(rng.uniform(np.log(128), np.log(1920)))))
        h = int(round(np.exp(rng.uniform(np.log(128), np.log(1080)))))
        mode = int(rng.choice([MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS]))
        depth = int(rng.integers(1, 7))
        # rotations are given as rational amounts tan(angle/2)
        azim = float(rng.uniform(-np.pi / 2, np.pi / 2))
        elev = float(np.radians(rng.uniform(8.0, 50.0)))
        pos = (float(dist * np.cos(elev) * np.sin(azim)), float(target_y + dist * np.sin(elev)),
               float(-dist * np.cos(elev) * np.cos(azim)))
        views.append(make_view(width=w, height=h, mode=mode, depth=depth, position=pos, yaw=float(np.tan(azim / 2)),
                               pitch=float(np.tan(-elev / 2))))
        scene_ids.append(int(rng.integers(0, nscenes)))
    return np.concatenate(views), scene_ids


def batch_workload():
    """The headline workload: the 8 scene blobs (PBR type first, then the
    instanced textured type) and the seeded 256-request mix; requests on the
    textured scenes sample with ray cones."""
    blobs = [st.gen_pbr(100 + i, k) for i, k in enumerate(BATCH_PBR_K)]
    blobs += [st.gen_instanced(200 + i, 64, tris, BATCH_TEXTURE) for i, tris in enumerate(BATCH_INSTANCED)]
    views, sids = request_mix(np.random.default_rng(BATCH_SEED), len(blobs), BATCH_REQUESTS)
    views["cone_angle"] = [BATCH_CONE if s >= len(BATCH_PBR_K) else 0.0 for s in sids]
    return blobs, views, sids


def batch_scene_triangles():
    return [20 * k * k for k in BATCH_PBR_K] + [sum(t) for t in BATCH_INSTANCED]


def pct(a, q):
    return round(float(np.percentile(a, q)), 4)


# ---------------------------------------------------------------- headline

def plain_and_download_legs(rs, views, rays_per_step):
    """The headline requests as plain launches (one request per launch, the
    same kernels, frames left in HBM), and both ways with every frame
    downloaded into page-locked host memory."""
    # plain: one launch per request, the same kernels, frames left in HBM
    lat_plain, plain_rays = [], 0
    for i in range(len(views)):                     # untimed pass: per-request arenas
        st.render_timed([rs[i]], views[i:i + 1])
    tp = time.perf_counter()
    for i in range(len(views)):
        r0 = time.perf_counter()
        _, c = st.render_timed([rs[i]], views[i:i + 1])
        lat_plain.append((time.perf_counter() - r0) * 1e3)
        plain_rays += c[0].total()
    plain_s = time.perf_counter() - tp
    assert plain_rays == rays_per_step, (plain_rays, rays_per_step)

    # with downloads: every frame into page-locked memory, both ways
    grid = [(int(v["width"]) + 1) * (int(v["height"]) + 1) for v in views]
    fp_plain = st.pinned_split(grid, st.PIXEL_DTYPE)
    fp_batch = st.pinned_split(grid, st.PIXEL_DTYPE)
    for i in range(len(views)):
        st.render(rs[i], views[i:i + 1], hits=False, out=fp_plain[i])
    tq = time.perf_counter()
    for i in range(len(views)):
        st.render(rs[i], views[i:i + 1], hits=False, out=fp_plain[i])
    plain_dl_s = time.perf_counter() - tq
    st.render_batch(rs, views, persistent=True, outs=fp_batch)
    tb = time.perf_counter()
    st.render_batch(rs, views, persistent=True, outs=fp_batch)
    batch_dl_s = time.perf_counter() - tb
    same = all(a.tobytes() == b.tobytes() for a, b in zip(fp_plain, fp_batch))
    del fp_plain, fp_batch

    return {"plain_s": plain_s, "plain_rays": plain_rays, "lat_plain": lat_plain, "plain_dl_s": plain_dl_s,
            "batch_dl_s": batch_dl_s, "same": same, "grid": grid}


def c4_leg(args, blobs, views, sids):
    t0 = time.perf_counter()
    dscenes = [st.DeviceScene(b, **BATCH_MESH_BUILD) for b in blobs]
    build_s = time.perf_counter() - t0
    rs = [dscenes[i] for i in sids]
    st.render_timed(rs, views)                      # grows the launch arena to the whole batch
    if args.fixed_shape:
        tune = st.get_tuning()
    else:
        # median of 5 timings per candidate: shapes 2% apart (4 or 8 blocks
        # per CU) are told apart more reliably than with 3
        tune, _ = st.autotune(rs, views, reps=5)
    log(f"c4: {len(dscenes)} scenes built in {build_s:.1f} s; {'fixed' if args.fixed_shape else 'autotune'} "
        f"{tune.as_dict()}")
    for _ in range(args.warmup):
        st.render_timed(rs, views, stages=True)
    step_ms, stage_runs, rays = [], [], 0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        f0 = time.perf_counter()
        _, counts, sms = st.render_timed(rs, views, stages=True)
        step_ms.append((time.perf_counter() - f0) * 1e3)
        stage_runs.append(sms)
        rays += sum(c.total() for c in counts)
    elapsed = time.perf_counter() - t_start
    redone = st.last_redo_count()
    rays_per_step = rays // args.steps
    names = st.stage_names(max(int(v) for v in views["depth"]))
    stage_ms = {n: round(float(np.mean([r[i] for r in stage_runs])), 4) for i, n in enumerate(names)}
    log(f"c4: {rays_per_step} rays per step, {np.median(step_ms):.2f} ms per step, stages {stage_ms}")

    plain = None
    if not args.batched_only:
        plain = plain_and_download_legs(rs, views, rays_per_step)

    batched_mrays = rays / elapsed / 1e6
    grid = [(int(v["width"]) + 1) * (int(v["height"]) + 1) for v in views]
    detail = {
        "requests": len(views), "scenes": len(dscenes),
        "scene_triangles": [d.info().triangle_count for d in dscenes],
        "mesh_bvh": BATCH_MESH_LABEL, "scene_build_s": round(build_s, 2),
        "traced_pixels": int(sum(grid)), "rays": int(rays_per_step),
        "batched_mrays": round(batched_mrays, 2),
        "batched_ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "batched_latency_p50_ms": pct(step_ms, 50), "batched_latency_p99_ms": pct(step_ms, 99),
        "redo_pixels": redone,
        "launch_shape": tune.as_dict(),
    }
    if plain is not None:
        detail.update({
            "plain_mrays": round(plain["plain_rays"] / plain["plain_s"] / 1e6, 2),
            "speedup_vs_plain": round(plain["plain_s"] / (elapsed / args.steps), 3),
            "plain_wall_ms": round(plain["plain_s"] * 1e3, 2),
            "plain_request_latency_p50_ms": pct(plain["lat_plain"], 50),
            "plain_request_latency_p99_ms": pct(plain["lat_plain"], 99),
            "with_downloads": {
                "output": "FloatPixel grids into page-locked host memory, both legs",
                "plain_mrays": round(plain["plain_rays"] / plain["plain_dl_s"] / 1e6, 2),
                "batched_mrays": round(rays_per_step / plain["batch_dl_s"] / 1e6, 2),
                "speedup_vs_plain": round(plain["plain_dl_s"] / plain["batch_dl_s"], 3),
            },
            "batched_equals_plain": plain["same"],
        })
    out = {
        "value": batched_mrays,
        "elapsed": elapsed,
        "step_ms": step_ms,
        "rays_per_step": rays_per_step,
        "stage_ms": stage_ms,
        "tune": tune,
        "detail": detail,
    }
    out["_scenes"] = dscenes
    out["_rs"] = rs
    return out


def c4_parity_and_bytes(args, blobs, views, sids, rs):
    """Every request of the headline mix on the CPU oracle: hit records and
    ray counts against the GPU's, and the per-class event counts that price
    the launch's algorithmic bytes."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers as ck
    threads = host_cores()
    t0 = time.perf_counter()
    agg = ck.OracleStats()
    mism = {"prim_id": 0, "triangle_id": 0, "distance": 0, "counts": 0}
    for s in range(len(blobs)):
        idx = [i for i in range(len(views)) if sids[i] == s]
        if not idx:
            continue
        osc = ck.OracleScene(blobs[s], **BATCH_MESH_BUILD)
        _, ghits, gcounts = st.render_batch([rs[i] for i in idx], views[idx], hits=True, persistent=True)
        for k, i in enumerate(idx):
            _, oh, ost, _ = osc.render(views[i:i + 1], threads=threads, pixels=False)
            gh = ghits[k]
            mism["prim_id"] += int((gh["prim_id"] != oh["prim_id"]).sum())
            mism["triangle_id"] += int((gh["triangle_id"] != oh["triangle_id"]).sum())
            mism["distance"] += int((gh["distance"].view(np.uint32) != oh["distance"].view(np.uint32)).sum())
            c = gcounts[k]
            if (c.primary, c.shadow, c.bounce) != (ost.primary_rays, ost.shadow_rays, ost.bounce_rays):
                mism["counts"] += 1
            for f, _ in ck.OracleStats._fields_:
                if f == "class_events":
                    for a in range(4):
                        for b in range(4):
                            agg.class_events[a][b] += ost.class_events[a][b]
                else:
                    setattr(agg, f, getattr(agg, f) + getattr(ost, f))
        osc.close()
        del ghits
    log(f"c4 parity: {mism} ({time.perf_counter() - t0:.1f} s, {threads} threads)")
    return mism, agg


def cpu_subset_baseline(blobs, views, sids):
    """The reference CPU renderer on the first CPU_SUBSET requests of the
    headline mix: all host cores (rows interleaved, one trace state per
    thread), median of 3 frames, and one pinned core, one frame."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers as ck
    from slimtracin_amd.scene import blob_mesh_info, blob_with_prebuilt_meshes
    kind = "reference" if ck.ref_available() else "port"
    cores = host_cores()
    sub = list(range(CPU_SUBSET))
    rays, secs_all, secs_one = 0, 0.0, 0.0
    prebuilt = {}
    for i in sub:
        s = sids[i]
        v = views[i:i + 1]
        if kind == "reference":
            if s not in prebuilt:
                info = blob_mesh_info(blobs[s])
                prebuilt[s] = blob_with_prebuilt_meshes(
                    blobs[s], [ck.oracle_mesh_binned(blobs[s], m, mi["triangle_count"]) for m, mi in enumerate(info)])
            osc = ck.OracleScene(prebuilt[s])
            _, _, ost, _ = osc.render(v, threads=cores, hits=False, pixels=False)
            osc.close()
            _, sa = ck.ref_render_timed(prebuilt[s], v, cores, 3)
            saved = os.sched_getaffinity(0)
            os.sched_setaffinity(0, {min(saved)})
            try:
                _, s1 = ck.ref_render_timed(prebuilt[s], v, 1, 1)
            finally:
                os.sched_setaffinity(0, saved)
        else:
            osc = ck.OracleScene(blobs[s], **BATCH_MESH_BUILD)
            ts = []
            for _ in range(3):
                _, _, ost, t = osc.render(v, threads=cores, hits=False, pixels=False)
                ts.append(t)
            sa = ts
            saved = os.sched_getaffinity(0)
            os.sched_setaffinity(0, {min(saved)})
            try:
                _, _, _, t1 = osc.render(v, threads=1, hits=False, pixels=False)
            finally:
                os.sched_setaffinity(0, saved)
            s1 = [t1]
            osc.close()
        rays += ost.primary_rays + ost.shadow_rays + ost.bounce_rays
        secs_all += float(np.median(sa))
        secs_one += float(np.median(s1))
    return {
        "value": rays / secs_all / 1e6, "unit": "Mrays/s", "cores": cores, "kind": kind,
        "sample": f"the first {CPU_SUBSET} requests of the seeded 256-request mix ({rays} rays), "
                  f"{cores} threads, rows interleaved, median of 3 frames each: {secs_all:.2f} s",
        "single_core_mrays": round(rays / secs_one / 1e6, 3),
        "single_core_s": round(secs_one, 2),
        "cpu_model": cpu_model(),
    }


def frame_cpu(blob, view, frames=3):
    """The reference CPU renderer (or the oracle) on one frame: all cores and
    one pinned core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers as ck
    cores = host_cores()
    osc = ck.OracleScene(blob)
    _, _, ost, _ = osc.render(view, threads=cores, hits=False, pixels=False)
    osc.close()
    rays = ost.primary_rays + ost.shadow_rays + ost.bounce_rays
    if ck.ref_available():
        _, sa = ck.ref_render_timed(blob, view, cores, frames)
        kind = "reference"
    else:
        _, sa, _ = ck.oracle_render_timed(blob, view, cores, frames)
        kind = "port"
    saved = os.sched_getaffinity(0)
    os.sched_setaffinity(0, {min(saved)})
    try:
        if kind == "reference":
            _, s1 = ck.ref_render_timed(blob, view, 1, 1)
        else:
            _, s1, _ = ck.oracle_render_timed(blob, view, 1, 1)
    finally:
        os.sched_setaffinity(0, saved)
    return {"kind": kind, "cores": cores, "mrays": round(rays / float(np.median(sa)) / 1e6, 3),
            "single_core_mrays": round(rays / float(np.median(s1)) / 1e6, 3), "rays": int(rays)}, ost


# ---------------------------------------------------------------- frames

def frame_leg(scene, view, steps, warmup, label):
    for _ in range(warmup):
        st.render_timed([scene], view)
    kms, frame_ms, rays = [], [], 0
    t0 = time.perf_counter()
    for _ in range(steps):
        f0 = time.perf_counter()
        ms, counts = st.render_timed([scene], view)
        frame_ms.append((time.perf_counter() - f0) * 1e3)
        kms.append(ms)
        rays += counts[0].total()
    elapsed = time.perf_counter() - t0
    out = {"rays_per_frame": int(rays // steps), "mrays": round(rays / elapsed / 1e6, 2),
           "frame_p50_ms": pct(frame_ms, 50), "frame_p99_ms": pct(frame_ms, 99),
           "render_ms": round(float(np.mean(kms)), 4)}
    log(f"{label}: {out}")
    return out


def c1_leg(args):
    blob = st.gen_implicit(1)
    scene = st.DeviceScene(blob)
    view = make_view(width=640, height=480, mode=MODE_BEAUTY, depth=2, position=(0.0, 7.0, -11.0), pitch=-0.2)
    saved = st.get_tuning()
    tune, _ = st.autotune([scene], view, reps=7)    # this frame's own launch shape
    out = frame_leg(scene, view, max(args.steps, 20), 3, "c1")
    out["launch_shape"] = tune.as_dict()
    out["workload"] = "16 seeded implicit primitives (4 quads, 4 boxes, 4 spheres, 4 tetrahedra), 3 point lights, " \
                      "Blinn-Phong, 640x480 (641x481 traced), depth 2, Beauty"
    if "cpu" in args.legs:
        cpu, ost = frame_cpu(blob, view)
        assert ost.primary_rays + ost.shadow_rays + ost.bounce_rays == out["rays_per_frame"]
        out["cpu"] = cpu
        out["gpu_vs_cpu_all_cores"] = round(out["mrays"] / cpu["mrays"], 1)
    st.restore_tuning(saved)
    scene.close()
    return out


def c2_leg(args):
    t0 = time.time()
    blob = st.gen_pbr(PBR_SEED, PBR_K)
    scene = st.DeviceScene(blob)
    info = scene.info()
    log(f"c2: {info.triangle_count} triangles, mesh BVH height {info.max_mesh_height}, "
        f"{info.device_bytes / 2**20:.1f} MiB in HBM, built in {time.time() - t0:.1f} s")
    view = make_view(width=1920, height=1080, mode=MODE_BEAUTY, depth=4, **PBR_CAM)
    saved = st.get_tuning()
    tune, _ = st.autotune([scene], view, reps=7)
    out = frame_leg(scene, view, max(args.steps, 20), 3, "c2")
    stage_runs = [st.render_timed([scene], view, stages=True)[2] for _ in range(5)]
    names = st.stage_names(4)
    out["stage_ms"] = {n: round(float(np.median([r[i] for r in stage_runs])), 4) for i, n in enumerate(names)}
    c0, l0 = st.graph_stats()
    for _ in range(2):
        st.render_timed([scene], view, graph=True)
    g_ms = []
    for _ in range(10):
        f0 = time.perf_counter()
        st.render_timed([scene], view, graph=True)
        g_ms.append((time.perf_counter() - f0) * 1e3)
    c1, l1 = st.graph_stats()
    out["graph"] = {"frame_ms_graph_p50": pct(g_ms, 50), "captures": c1 - c0, "replays": l1 - l0 - (c1 - c0)}
    t1 = time.perf_counter()
    for _ in range(3):
        st.render(scene, view, hits=False)
    out["frame_ms_with_pageable_download"] = round((time.perf_counter() - t1) / 3 * 1e3, 3)
    out["launch_shape"] = tune.as_dict()
    out["workload"] = "200,000-triangle noise-displaced icosphere (reference SAH mesh BVH), glass sphere, mirror " \
                      "quad, 4 point + 2 emissive quad lights, 1920x1080 (1921x1081 traced), depth 4, Beauty"
    if "cpu" in args.legs:
        cpu, ost = frame_cpu(blob, view, frames=1)
        out["cpu"] = cpu
        out["gpu_vs_cpu_all_cores"] = round(out["mrays"] / cpu["mrays"], 1)
        _, gh, _ = st.render(scene, view)
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import checkers as ck
        _, oh, _ = ck.oracle_render(blob, view, threads=host_cores())
        out["hit_ids_exact"] = bool((gh["prim_id"] == oh["prim_id"]).all() and
                                    (gh["triangle_id"] == oh["triangle_id"]).all() and
                                    (gh["distance"].view(np.uint32) == oh["distance"].view(np.uint32)).all())
    st.restore_tuning(saved)
    scene.close()
    return out


def instanced_leg(args):
    from slimtracin_amd.scene import PRIM_DTYPE, blob_prims, quat_axis, quat_mul_norm
    t0 = time.perf_counter()
    blob = st.gen_instanced(2025, 64, [50_000, 200_000, 500_000], 2048)
    ds = st.DeviceScene(blob, **INSTANCED_MESH_BUILD)
    build_s = time.perf_counter() - t0
    prims = blob_prims(blob)
    rng = np.random.default_rng(3)
    view = make_view(width=1920, height=1080, mode=MODE_BEAUTY, depth=3, cone_angle=0.001,
                     position=(0.0, 14.0, -35.0), pitch=-0.3)

    def moved():
        p = prims.copy()
        for i in range(len(p) - 1):
            p["position"][i] = (rng.uniform(-30, 30), rng.uniform(1.5, 6), rng.uniform(-20, 40))
            q = quat_axis((1, 0, 0), float(rng.uniform(-3, 3)))
            q = quat_mul_norm(q, quat_axis((0, 1, 0), float(rng.uniform(-3, 3))))
            p["rotation"][i] = quat_mul_norm(q, quat_axis((0, 0, 1), float(rng.uniform(-3, 3))))
        return np.ascontiguousarray(p, dtype=PRIM_DTYPE)

    frames = [moved() for _ in range(args.instanced_frames + 2)]
    saved = st.get_tuning()
    ds.set_prims(frames[0])
    ds.rebuild(st.ST_REBUILD_LBVH)
    tune, _ = st.autotune([ds], view, reps=7)       # the frame's own launch shape, on its first layout
    for p in frames[:2]:                                            # warm-up
        ds.set_prims(p)
        ds.rebuild(st.ST_REBUILD_LBVH)
        st.render_timed([ds], view)
    frame_ms, rebuild_ms, kernel_ms, rays = [], [], [], 0
    tstart = time.perf_counter()
    for p in frames[2:]:
        f0 = time.perf_counter()
        ds.set_prims(p)
        rebuild_ms.append(ds.rebuild(st.ST_REBUILD_LBVH))
        k, c = st.render_timed([ds], view)
        frame_ms.append((time.perf_counter() - f0) * 1e3)
        kernel_ms.append(k)
        rays += c[0].total()
    wall = time.perf_counter() - tstart
    info = ds.info()
    ds.close()
    st.restore_tuning(saved)
    return {
        "instances": 64, "mesh_triangles": [50_000, 200_000, 500_000], "triangles": int(info.triangle_count),
        "texture": "2 x 2048^2 Perlin (albedo, normal), 10 mips", "cone_angle": 0.001, "depth": 3,
        "mesh_bvh": BATCH_MESH_LABEL, "scene_build_s": round(build_s, 2), "frames": len(frame_ms),
        "rays_per_frame": int(rays / len(frame_ms)), "mrays": round(rays / wall / 1e6, 2),
        "frame_p50_ms": pct(frame_ms, 50), "frame_p99_ms": pct(frame_ms, 99),
        "frame_max_ms": round(float(np.max(frame_ms)), 3),
        "lbvh_rebuild_ms": round(float(np.median(rebuild_ms)), 4),
        "render_ms": round(float(np.median(kernel_ms)), 3),
        "launch_shape": tune.as_dict(),
    }


def served_leg(args):
    import threading
    from slimtracin_amd.server import OUTPUT_BGRA8, Server
    rng = np.random.default_rng(11)
    tri_counts = np.exp(rng.uniform(np.log(1e6), np.log(20e6), args.serve_scenes)).astype(np.int64)
    srv = Server(hbm_budget=200 * 10**9, output=OUTPUT_BGRA8)
    ids, admitted_tris, rejected = [], [], 0
    t0 = time.perf_counter()
    for i, tri in enumerate(tri_counts):
        # the single-mesh PBR type: a 20k^2-triangle icosphere
        blob = st.gen_pbr(500 + i, max(1, int(round(np.sqrt(tri / 20.0)))))
        sid = srv.try_add_scene(blob, mesh_lbvh=True)
        if sid is None:
            rejected += 1
        else:
            ids.append(sid)
            admitted_tris.append(20 * int(round(np.sqrt(tri / 20.0))) ** 2)
        del blob
        if (i + 1) % 8 == 0:
            log(f"served: {i + 1}/{len(tri_counts)} scenes admitted-or-rejected, {time.perf_counter() - t0:.1f} s")
    build_s = time.perf_counter() - t0

    def burst(vs, ps):
        tb = time.perf_counter()
        ts = [srv.submit(ids[ps[k]], vs[k:k + 1]) for k in range(len(vs))]
        r = 0
        for t in ts:
            _, c, _ = srv.wait(t, copy=False)
            r += c.primary + c.shadow + c.bounce
            srv.release(t)
        return len(vs) / (time.perf_counter() - tb), r / (time.perf_counter() - tb) / 1e6

    wv, wp = request_mix(rng, len(ids), 64)
    burst(wv, wp)                                    # warm-up
    bv, bp = request_mix(rng, len(ids), 512)
    capacity, burst_mrays = burst(bv, bp)            # requests/s with a full queue
    # the same 512 requests as plain launches, one request per launch,
    # through the same server (max one request per launch, one in flight)
    plain = srv.plain_rate(bv, [ids[k] for k in bp])
    log(f"served: burst {capacity:.0f} req/s, {burst_mrays:.0f} Mrays/s; plain {plain}")
    # Poisson arrivals at 80% of the rate the plain loop sustains, served
    # both ways: through the server (continuous batching, several launches
    # in flight) and as plain launches (one request per launch, one launch
    # at a time), the same requests at the same arrival times
    rate = args.serve_rate if args.serve_rate > 0 else round(0.8 * plain["req_per_s"], 1)
    log(f"served: {len(ids)} scenes resident in {build_s:.1f} s; Poisson rate {rate} req/s for {args.serve_seconds} s")
    n = max(1, int(rate * args.serve_seconds))
    views, picks = request_mix(rng, len(ids), n)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))

    def poisson(label):
        stats0 = srv.stats()
        tickets = [None] * n
        lat = []
        rays = [0]

        def collector():
            last = time.perf_counter()
            for k in range(n):
                while tickets[k] is None:
                    time.sleep(0.0002)
                _, c, l_ms = srv.wait(tickets[k], copy=False)
                rays[0] += c.primary + c.shadow + c.bounce
                lat.append(l_ms)
                srv.release(tickets[k])
                if time.perf_counter() - last > 2.0:
                    last = time.perf_counter()
                    log(f"served ({label}): {k + 1}/{n} requests done")

        th = threading.Thread(target=collector)
        th.start()
        start = time.perf_counter()
        for k in range(n):
            while True:
                gap = arrivals[k] - (time.perf_counter() - start)
                if gap <= 0:
                    break
                time.sleep(gap if gap > 0.001 else 0)
            tickets[k] = srv.submit(ids[picks[k]], views[k:k + 1])
        th.join()
        wall = time.perf_counter() - start
        stats = srv.stats()
        launches = int(stats["launches"] - stats0["launches"])
        return {"wall_s": round(wall, 3), "mrays": round(rays[0] / wall / 1e6, 2),
                "latency_p50_ms": pct(lat, 50), "latency_p99_ms": pct(lat, 99),
                "latency_max_ms": round(float(np.max(lat)), 3), "launches": launches,
                "mean_batch": round((stats["completed"] - stats0["completed"]) / max(1, launches), 2),
                "max_in_flight": stats.get("max_in_flight")}

    served = poisson("server")
    log(f"served: {served}")
    srv.set_limits(1, 1)
    try:
        plain_poisson = poisson("plain")
    finally:
        srv.set_limits(0, 0)
    log(f"served: plain launches at the same arrivals {plain_poisson}")
    stats = srv.stats()
    srv.close()
    return {
        "scenes_requested": int(args.serve_scenes), "scenes_admitted": len(ids), "scenes_rejected": rejected,
        "scene_triangles_total": int(sum(admitted_tris)), "hbm_budget_gb": 200,
        "resident_gb": round(stats["resident_bytes"] / 1e9, 2), "scene_build_s": round(build_s, 1),
        "mesh_bvh": "device Morton LBVH (one triangle per leaf)", "output": "BGRA8 window image (device resolve)",
        "burst_requests": 512, "burst_capacity_req_per_s": round(capacity, 1), "burst_mrays": round(burst_mrays, 2),
        "plain_burst": plain,
        "arrival_rate_per_s": rate, "arrival_rate": "80% of the plain loop's request rate", "requests": n,
        **served,
        "plain_at_same_arrivals": plain_poisson,
    }


# ---------------------------------------------------------------- roofline

TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_primary4_batched256_c4.txt")


def measured_traffic():
    """HBM bytes per launch of the headline step's dominant kernel from the
    committed rocprofv3 PMC summary (line 'traffic_bytes_per_launch: N'),
    or None when absent."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def copy_probe(mib=512, reps=5):
    """A plain device copy of known size (elementwise negation, reads and
    writes `mib` MiB per launch) for calibrating rocprofv3's FETCH_SIZE and
    WRITE_SIZE counters."""
    import torch
    n = mib * 2**20 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for i in range(reps):
        torch.neg(a, out=b)
        torch.cuda.synchronize()
        print(f"copy probe {i}: {mib} MiB read, {mib} MiB written", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed batched launches of the headline workload")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--legs", default=",".join(LEGS),
                    help="comma list of " + ",".join(LEGS) + " (c4 is the headline; cpu adds the CPU baselines "
                         "and the oracle parity of the headline)")
    ap.add_argument("--copy-probe", action="store_true", help="only run the copy kernel used to calibrate PMC bytes")
    ap.add_argument("--fixed-shape", action="store_true",
                    help="headline leg at the library's default launch shape instead of the autotuner's choice "
                         "(so that several profiling runs time the same launches)")
    ap.add_argument("--batched-only", action="store_true",
                    help="headline leg without its plain-launch and download comparisons (for profiling runs)")
    ap.add_argument("--instanced-frames", type=int, default=10, help="instanced leg: timed frames")
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="served leg: Poisson arrival rate, requests/s (0: 80%% of the plain loop's rate)")
    ap.add_argument("--serve-seconds", type=float, default=6.0, help="served leg: arrival window, s")
    ap.add_argument("--serve-scenes", type=int, default=40, help="served leg: resident scenes requested")
    args = ap.parse_args()
    args.legs = [x for x in args.legs.split(",") if x]
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box in this setup has a single MI355X, so --gpus {args.gpus} "
              "is not supported (the multi-GPU design routes whole requests to the GPU that holds their scene; "
              "no per-frame collective exists to measure).", file=sys.stderr)
        sys.exit(2)
    st.init(0)
    blobs, views, sids = batch_workload()
    c4 = c4_leg(args, blobs, views, sids)
    result = {
        "metric": METRIC,
        "value": round(c4["value"], 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(c4["elapsed"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched_256_requests_8_scenes_pbr_and_instanced_textured_0.2-5M_tris_"
                               "mixed_res_depth1-6_4modes",
                   "requests": BATCH_REQUESTS, "scenes": len(BATCH_PBR_K) + len(BATCH_INSTANCED),
                   "scene_unique_triangles": batch_scene_triangles(), "rays_per_step": c4["rays_per_step"],
                   "seed": BATCH_SEED, "spp": 1},
        "p50_ms": pct(c4["step_ms"], 50),
        "p99_ms": pct(c4["step_ms"], 99),
        "stage_ms": c4["stage_ms"],
        "batched": c4["detail"],
    }
    if "cpu" in args.legs:
        mism, agg = c4_parity_and_bytes(args, blobs, views, sids, c4["_rs"])
        result["hit_ids_exact"] = mism["prim_id"] == 0 and mism["triangle_id"] == 0 and mism["distance"] == 0
        result["ray_counts_exact"] = mism["counts"] == 0
        result["parity_mismatches"] = mism
        cpu = cpu_subset_baseline(blobs, views, sids)
        result["cpu_baseline"] = cpu
        log("cpu baseline:", cpu)
        result["gpu_vs_cpu"] = round(result["value"] / cpu["value"], 1)
        result["gpu_vs_cpu_single_core"] = round(result["value"] / cpu["single_core_mrays"], 1)
        stage = c4["stage_ms"]
        dominant = max(("primary", "bounce"), key=lambda k: stage.get(k, 0.0))
        if dominant == "primary":
            nbytes = agg.traversal_bytes("primary") + agg.traversal_bytes("shadow0")
            # the 4-waves-per-SIMD build of the same stage when the launch shape asks for it
            kname = ("k_primary4" if c4["tune"].primary_waves == 4 else "k_primary") + \
                " (primary rays and the first level's shadow rays, all 256 requests)"
        else:
            nbytes = agg.traversal_bytes("bounce") + agg.traversal_bytes("shadow_deep")
            kname = "k_bounce (bounce rays and the deeper levels' shadow rays, all 256 requests)"
        achieved = nbytes / (stage[dominant] * 1e-3) / 1e9
        result["roofline"] = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(),
                              "kernel": kname, "algorithmic_bytes_per_launch": int(nbytes),
                              "launch_ms": stage[dominant]}
        total = agg.algorithmic_bytes()
        step_s = c4["elapsed"] / args.steps
        result["roofline_step"] = {"algorithmic_bytes": int(total), "bytes_per_ray": round(total / c4["rays_per_step"], 2),
                                   "achieved": round(total / step_s / 1e9, 2),
                                   "frac": round(total / step_s / 1e9 / HBM_PEAK_GBS, 5)}
    for d in c4["_scenes"]:
        d.close()
    del c4
    if "c1" in args.legs:
        result["frame_c1"] = c1_leg(args)
    if "c2" in args.legs:
        result["frame_c2"] = c2_leg(args)
    if "c3" in args.legs:
        result["instanced"] = instanced_leg(args)
        log("instanced:", result["instanced"])
    if "c5" in args.legs:
        result["served"] = served_leg(args)
        log("served:", result["served"])
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
