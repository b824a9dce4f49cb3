#!/usr/bin/env python3
"""Benchmark of the MI355X tracing path (one JSON line on stdout).

Workload (one step = one frame): the single-mesh PBR frame - a seeded
noise-displaced geodesic icosphere with 200,000 fp32 triangles (frequency
100), Cook-Torrance materials, a glass sphere, a mirror quad, 4 point lights
and 2 emissive quads; 1920x1080 window (1921x1081 traced QCAA grid), 1 spp,
trace depth 4, Beauty mode. The scene is resident in HBM before timing;
a step is the per-frame host set-up (camera basis, SSB rects, ~20 KB of
uploads), the primary-direction kernel and the render stages (primary,
which also runs the first depth-loop level's shading entry, shadow rays and
shading, then the bounce stage, which traces each bounce ray and runs its
pixel's further levels in the same lane); the frame stays in HBM
(the PCIe-inclusive rate is reported separately).

Also measured in the same run:
  - plain vs batched launches ("batched"): 256 requests over 8 resident
    scenes of 0.2-5M triangles (noise-displaced spheres, mesh BVHs built on
    the device), resolutions log-uniform 128x128..1920x1080, trace depth
    uniform 1-6, mode uniform over Beauty/Depth/Normal/UV, cameras on a
    sphere around each scene; first as 256 sequential per-request launches
    with a blocking FloatPixel download each (the reference's
    renderSceneOnGPU pattern), then as one persistent launch over one tile
    list, with FloatPixel downloads and with the window images resolved on
    the device (BGRA8); the batched frames are checked equal to the plain
    ones;
  - instanced ("instanced"): 64 instances of 3 noise-displaced meshes
    (50k/200k/500k triangles, reference SAH mesh BVHs), 2048^2 Perlin
    albedo and normal maps with full mip chains, ray-cone angle 0.001,
    1920x1080, depth 3, Beauty; every frame moves all instances (new random
    positions and rotations), rebuilds the scene BVH on the device (Morton
    LBVH) and renders: frame time includes upload, rebuild and render;
  - serving ("served"): 40 resident scenes of 1-20M triangles (log-uniform,
    admitted against a 200 GB HBM budget), requests of the same mix arriving
    as a Poisson process at a stated rate for a stated time, served by the
    continuous-batching loop; Mrays/s, p50/p99 submit-to-image latency;
  - roofline of the dominant stage launch (the primary stage, k_primary,
    which traces the primary rays and the first level's shadow rays): the
    algorithmic bytes of those rays' traversals from the CPU oracle's
    per-ray-class event counters on the same 1920x1080 frame (bytes per
    event at the reference's record sizes), over the launch's duration from
    HIP events recorded around each stage on the launch stream; the same
    for the whole frame ("roofline_frame"); the measured HBM traffic per
    launch comes from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary kept in profiles/ (calibrated
    against a copy kernel of known size, --copy-probe);
  - the reference CPU renderer (compiled from the reference's own sources)
    on the same full frame, all threads used stated;
  - hit-id parity of the GPU frame against the CPU oracle (every pixel).
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
from slimtracin_amd.scene import MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS, VIEW_DTYPE, make_view  # noqa: E402

HBM_PEAK_GBS = 8000.0
PBR_K = 100            # 20 * 100^2 = 200,000 triangles
PBR_SEED = 2024
PBR_CAM = dict(position=(0.0, 6.0, -14.0), pitch=-0.15)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pbr_view(w, h):
    return make_view(width=w, height=h, mode=MODE_BEAUTY, depth=4, **PBR_CAM)


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


def cpu_baseline(blob, view, threads, frames=1):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import checkers as ck
    pixels, hits, stats = ck.oracle_render(blob, view, threads=threads)
    rays = stats.primary_rays + stats.shadow_rays + stats.bounce_rays
    kind = "reference" if ck.ref_available() else "port"
    if kind == "reference":
        _, secs = ck.ref_render_timed(blob, view, threads, frames)
    else:
        _, secs, _ = ck.oracle_render_timed(blob, view, threads, frames)
    t = float(np.median(secs))
    return dict(value=rays / t / 1e6, unit="Mrays/s", cores=threads, kind=kind,
                sample=f"the same {int(view['width'][0])}x{int(view['height'][0])} Beauty frame, "
                       f"{threads} threads, median of {frames} frame(s), {rays} rays, {t:.2f} s/frame"), stats, hits


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_primary_pbr200k_1920x1080.txt")


def measured_traffic():
    """HBM bytes per primary-stage launch from the committed rocprofv3 PMC
    summary (line 'traffic_bytes_per_launch: N'), or None when absent."""
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


def instanced_leg(args):
    from slimtracin_amd.scene import PRIM_DTYPE, blob_prims, quat_axis, quat_mul_norm
    t0 = time.perf_counter()
    blob = st.gen_instanced(2025, 64, [50_000, 200_000, 500_000], 2048)
    ds = st.DeviceScene(blob)
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
    return {
        "instances": 64, "mesh_triangles": [50_000, 200_000, 500_000], "triangles": int(info.triangle_count),
        "texture": "2 x 2048^2 Perlin (albedo, normal), 10 mips", "cone_angle": 0.001, "depth": 3,
        "scene_build_s": round(build_s, 2), "frames": len(frame_ms),
        "rays_per_frame": int(rays / len(frame_ms)), "mrays": round(rays / wall / 1e6, 2),
        "frame_p50_ms": round(float(np.percentile(frame_ms, 50)), 3),
        "frame_p99_ms": round(float(np.percentile(frame_ms, 99)), 3),
        "lbvh_rebuild_ms": round(float(np.median(rebuild_ms)), 4),
        "render_ms": round(float(np.median(kernel_ms)), 3),
    }


BATCH_TRIANGLES = [200_000, 320_000, 500_000, 800_000, 1_250_000, 2_000_000, 3_200_000, 5_000_000]


def batched_leg(args):
    rng = np.random.default_rng(7)
    dscenes = []
    t0 = time.perf_counter()
    for i, tri in enumerate(BATCH_TRIANGLES):
        dscenes.append(st.DeviceScene(st.gen_mesh_scene(100 + i, tri), mesh_lbvh=True))
    build_s = time.perf_counter() - t0
    views, sids = request_mix(rng, len(dscenes), 256)
    req_scenes = [dscenes[i] for i in sids]
    st.render_timed(req_scenes, views)          # warm-up: grows the launch arena to the whole batch
    tune, _ = st.autotune(req_scenes, views, reps=1)
    lat_plain, plain_rays, plain_pix = [], 0, []
    tp = time.perf_counter()
    for i in range(len(views)):
        r0 = time.perf_counter()
        px, _, c = st.render(req_scenes[i], views[i:i + 1], hits=False)
        lat_plain.append((time.perf_counter() - r0) * 1e3)
        plain_rays += c.total()
        plain_pix.append(px)
    plain_s = time.perf_counter() - tp
    # page-locked outputs, allocated once outside the timed regions (as the
    # serving loop's buffer pool does)
    grid = [(int(v["width"]) + 1) * (int(v["height"]) + 1) for v in views]
    win = [int(v["width"]) * int(v["height"]) for v in views]
    fp_out = st.pinned_split(grid, st.PIXEL_DTYPE)
    bgra_out = st.pinned_split(win, np.uint32)
    tq = time.perf_counter()
    for i in range(len(views)):
        st.render(req_scenes[i], views[i:i + 1], hits=False, out=fp_out[i])
    plain_pinned_s = time.perf_counter() - tq
    tb = time.perf_counter()
    bpix, _, cb = st.render_batch(req_scenes, views, persistent=True, outs=fp_out)
    batch_s = time.perf_counter() - tb
    tr = time.perf_counter()
    _, cr = st.render_batch_bgra8(req_scenes, views, persistent=True, outs=bgra_out)
    served_s = time.perf_counter() - tr
    kernel_ms, _ = st.render_timed(req_scenes, views)
    batch_rays = sum(c.total() for c in cb)
    assert batch_rays == plain_rays == sum(c.total() for c in cr), (batch_rays, plain_rays)
    same = all(a.tobytes() == b.tobytes() for a, b in zip(plain_pix, bpix))
    out = {
        "requests": len(views), "scenes": len(dscenes),
        "scene_triangles": [d.info().triangle_count for d in dscenes],
        "mesh_bvh": "device Morton LBVH (leaves <= 4)", "scene_build_s": round(build_s, 2),
        "traced_pixels": int(sum((int(v["width"]) + 1) * (int(v["height"]) + 1) for v in views)),
        "rays": int(batch_rays),
        "plain_mrays": round(plain_rays / plain_s / 1e6, 2),
        "plain_pinned_mrays": round(plain_rays / plain_pinned_s / 1e6, 2),
        "batched_mrays": round(batch_rays / batch_s / 1e6, 2),
        "batched_bgra8_mrays": round(batch_rays / served_s / 1e6, 2),
        "speedup": round(plain_s / batch_s, 3),
        "speedup_bgra8": round(plain_s / served_s, 3),
        "plain_wall_ms": round(plain_s * 1e3, 2),
        "batched_wall_ms": round(batch_s * 1e3, 2),
        "batched_bgra8_wall_ms": round(served_s * 1e3, 2),
        "plain_latency_p50_ms": round(float(np.percentile(lat_plain, 50)), 3),
        "plain_latency_p99_ms": round(float(np.percentile(lat_plain, 99)), 3),
        "batched_equals_plain": same,
        "batched_render_ms": round(kernel_ms, 3),
        "launch_shape": tune.as_dict(),
    }
    for d in dscenes:
        d.close()
    return out


def served_leg(args):
    import threading
    from slimtracin_amd.server import OUTPUT_BGRA8, Server
    rng = np.random.default_rng(11)
    tri_counts = np.exp(rng.uniform(np.log(1e6), np.log(20e6), args.serve_scenes)).astype(np.int64)
    srv = Server(hbm_budget=200 * 10**9, output=OUTPUT_BGRA8)
    ids, admitted_tris, rejected = [], [], 0
    t0 = time.perf_counter()
    for i, tri in enumerate(tri_counts):
        blob = st.gen_mesh_scene(500 + i, int(tri))
        sid = srv.try_add_scene(blob, mesh_lbvh=True)
        if sid is None:
            rejected += 1
        else:
            ids.append(sid)
            admitted_tris.append(int(tri))
        del blob
        if (i + 1) % 8 == 0:
            log(f"served: {i + 1}/{len(tri_counts)} scenes admitted-or-rejected, {time.perf_counter() - t0:.1f} s")
    build_s = time.perf_counter() - t0

    def burst(n):
        vs, ps = request_mix(rng, len(ids), n)
        tb = time.perf_counter()
        ts = [srv.submit(ids[ps[k]], vs[k:k + 1]) for k in range(n)]
        for t in ts:
            srv.wait(t, copy=False)
            srv.release(t)
        return n / (time.perf_counter() - tb)

    burst(64)                                       # warm-up
    capacity = burst(512)                           # requests/s with a full queue
    log(f"served: burst stats {srv.stats()}")
    # under Poisson arrivals the loop runs smaller batches than a full queue
    # gives, so the stated rate is kept well below the burst capacity
    rate = args.serve_rate if args.serve_rate > 0 else round(0.35 * capacity, 1)
    log(f"served: {len(ids)} scenes resident in {build_s:.1f} s; burst capacity {capacity:.0f} req/s; "
        f"Poisson rate {rate} req/s for {args.serve_seconds} s")
    stats0 = srv.stats()
    n = max(1, int(rate * args.serve_seconds))
    views, picks = request_mix(rng, len(ids), n)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
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
                st_ = srv.stats()
                log(f"served: {k + 1}/{n} requests done; launches {st_['launches']}, kernel {st_['kernel_ms']:.0f} ms, "
                    f"launch wall {st_['launch_ms']:.0f} ms, prep {st_['prep_ms']:.0f} ms, alloc {st_['alloc_ms']:.0f} ms")

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
    srv.close()
    return {
        "scenes_requested": int(args.serve_scenes), "scenes_admitted": len(ids), "scenes_rejected": rejected,
        "scene_triangles_total": int(sum(admitted_tris)), "hbm_budget_gb": 200,
        "resident_gb": round(stats["resident_bytes"] / 1e9, 2), "scene_build_s": round(build_s, 1),
        "mesh_bvh": "device Morton LBVH (leaves <= 4)", "output": "BGRA8 window image (device resolve)",
        "burst_capacity_req_per_s": round(capacity, 1), "arrival_rate_per_s": rate, "requests": n,
        "wall_s": round(wall, 3), "mrays": round(rays[0] / wall / 1e6, 2),
        "latency_p50_ms": round(float(np.percentile(lat, 50)), 3),
        "latency_p99_ms": round(float(np.percentile(lat, 99)), 3),
        "latency_max_ms": round(float(np.max(lat)), 3),
        "launches": int(stats["launches"] - stats0["launches"]),
        "server_ms": {k: round(stats[k] - stats0[k], 1) for k in ("prep_ms", "launch_ms", "kernel_ms", "alloc_ms")},
        "mean_batch": round((stats["completed"] - stats0["completed"]) / max(1, stats["launches"] - stats0["launches"]), 2),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and oracle legs")
    ap.add_argument("--no-batch", action="store_true", help="skip the plain-vs-batched leg")
    ap.add_argument("--cpu-frames", type=int, default=3, help="frames of the CPU baseline (median)")
    ap.add_argument("--copy-probe", action="store_true", help="only run the copy kernel used to calibrate PMC bytes")
    ap.add_argument("--no-serve", action="store_true", help="skip the served (Poisson arrivals) leg")
    ap.add_argument("--no-instanced", action="store_true", help="skip the instanced (per-frame LBVH) leg")
    ap.add_argument("--instanced-frames", type=int, default=10, help="instanced leg: timed frames")
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="served leg: Poisson arrival rate, requests/s (0: 35%% of the measured burst capacity)")
    ap.add_argument("--serve-seconds", type=float, default=8.0, help="served leg: arrival window, s")
    ap.add_argument("--serve-scenes", type=int, default=40, help="served leg: resident scenes requested")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box in this setup has a single MI355X, so --gpus {args.gpus} "
              "is not supported (the multi-GPU design routes whole requests to the GPU that holds their scene; "
              "no per-frame collective exists to measure).", file=sys.stderr)
        sys.exit(2)

    st.init(0)
    t0 = time.time()
    blob = st.gen_pbr(PBR_SEED, PBR_K)
    scene = st.DeviceScene(blob)
    info = scene.info()
    log(f"scene: {info.triangle_count} triangles, mesh BVH height {info.max_mesh_height}, "
        f"{info.device_bytes / 2**20:.1f} MiB in HBM, built in {time.time() - t0:.1f} s")
    view = pbr_view(1920, 1080)
    tune, tune_ms = st.autotune([scene], view)
    log(f"autotune: {tune.as_dict()} ({tune_ms:.3f} ms)")

    for _ in range(args.warmup):
        st.render_timed([scene], view)
    kms, frame_ms, rays = [], [], 0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        f0 = time.perf_counter()
        ms, counts = st.render_timed([scene], view)
        frame_ms.append((time.perf_counter() - f0) * 1e3)
        kms.append(ms)
        rays += counts[0].total()
    elapsed = time.perf_counter() - t_start
    rays_per_frame = rays / args.steps
    mrays = rays / elapsed / 1e6
    kernel_ms = float(np.mean(kms))
    log(f"frame: {rays_per_frame / 1e6:.2f} Mrays, render stages {kernel_ms:.3f} ms, "
        f"frame {np.median(frame_ms):.3f} ms, {mrays:.1f} Mrays/s")

    # per-stage durations (events between the stage launches), after the
    # timed region so that the extra events do not touch the headline figure
    stage_runs = []
    for _ in range(max(3, min(args.steps, 10))):
        _, _, sms = st.render_timed([scene], view, stages=True)
        stage_runs.append(sms)
    stage_ms = [float(np.median([r[i] for r in stage_runs])) for i in range(len(stage_runs[0]))]
    names = st.stage_names(int(view["depth"][0]))
    stages = {n: round(m, 4) for n, m in zip(names, stage_ms)}
    log("stages (ms):", stages)

    # the same frames as HIP-graph replays (the launch captured once)
    c0, l0 = st.graph_stats()
    for _ in range(2):
        st.render_timed([scene], view, graph=True)
    g_ms = []
    for _ in range(args.steps):
        f0 = time.perf_counter()
        st.render_timed([scene], view, graph=True)
        g_ms.append((time.perf_counter() - f0) * 1e3)
    c1, l1 = st.graph_stats()
    graph = {"frame_ms_eager_p50": round(float(np.median(frame_ms)), 4),
             "frame_ms_graph_p50": round(float(np.median(g_ms)), 4),
             "captures": c1 - c0, "replays": l1 - l0 - (c1 - c0)}
    log("graph:", graph)

    # PCIe-inclusive: a plain frame with the blocking download (renderSceneOnGPU)
    t1 = time.perf_counter()
    for _ in range(3):
        st.render(scene, view, hits=False)
    pcie_frame_ms = (time.perf_counter() - t1) / 3 * 1e3

    result = {
        "metric": "Mrays/s + p50/p99 frame latency, batched/served vs plain launches, hit-IDs exact",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "pbr_icosphere_200k_tris_1920x1080_depth4_beauty", "triangles": info.triangle_count,
                   "traced_pixels": 1921 * 1081, "rays_per_frame": int(rays_per_frame), "spp": 1},
        "p50_ms": round(float(np.percentile(frame_ms, 50)), 4),
        "p99_ms": round(float(np.percentile(frame_ms, 99)), 4),
        "frame_ms_with_download": round(pcie_frame_ms, 4),
        "render_ms": round(kernel_ms, 4),
        "stage_ms": stages,
        "graph": graph,
        "launch_shape": tune.as_dict(),
    }

    if not args.no_instanced:
        result["instanced"] = instanced_leg(args)
        log("instanced:", result["instanced"])
    if not args.no_batch:
        result["batched"] = batched_leg(args)
        log("batched:", result["batched"])
    if not args.no_serve:
        result["served"] = served_leg(args)
        log("served:", result["served"])

    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        cpu, stats, ohits = cpu_baseline(blob, view, threads, frames=args.cpu_frames)
        result["cpu_baseline"] = cpu
        log("cpu baseline:", cpu)
        oracle_rays = stats.primary_rays + stats.shadow_rays + stats.bounce_rays
        assert oracle_rays == int(rays_per_frame), (oracle_rays, rays_per_frame)
        frame_bytes = stats.algorithmic_bytes()

        def line(nbytes, ms):
            a = nbytes / (ms * 1e-3) / 1e9
            return {"achieved": round(a, 2), "frac": round(a / HBM_PEAK_GBS, 5),
                    "algorithmic_bytes_per_launch": int(nbytes), "ms": round(ms, 4)}

        primary = line(stats.traversal_bytes("primary") + stats.traversal_bytes("shadow0"), stages["primary"])
        result["roofline"] = {"bound": "hbm", "achieved": primary["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": primary["frac"], "traffic": measured_traffic(),
                              "kernel": "k_primary (primary rays and the first level's shadow rays)",
                              "algorithmic_bytes_per_launch": primary["algorithmic_bytes_per_launch"],
                              "launch_ms": primary["ms"]}
        result["roofline_frame"] = dict(line(frame_bytes, kernel_ms), bytes_per_ray=round(frame_bytes / oracle_rays, 2))
        gp, gh, gc = st.render(scene, view)
        same = bool((gh["prim_id"] == ohits["prim_id"]).all() and (gh["triangle_id"] == ohits["triangle_id"]).all()
                    and (gh["distance"].view(np.uint32) == ohits["distance"].view(np.uint32)).all())
        result["hit_ids_exact"] = same
        result["gpu_vs_cpu"] = round(mrays / cpu["value"], 1)
    scene.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
