"""GPU parity at the benchmark configurations' full sizes.

The other parity tests run reduced scenes; these run each configuration the
benchmark measures at its real size against the CPU oracle (the reference
CPU renderer restated; render/raytracer.h:136-171 and below), with every
host core:
  C2  the 200,000-triangle PBR frame, 1920x1080, depth 4: every hit record
      bit for bit, every pixel, the ray counts;
  C3  the instanced textured frame: 64 instances of 50k/200k/500k-triangle
      meshes (binned SAH trees, as the benchmark builds them) moved and the
      scene BVH rebuilt on the device (Morton LBVH), 2048^2 albedo and
      normal maps with mip chains, cone angle 0.001, 1920x1080, depth 3,
      against the oracle over the same trees;
  C4  the 256-request batch over 8 scenes of 0.2-5M triangles in one launch
      (meshes built as the benchmark builds them, binned SAH): equal to the
      256 plain launches byte for byte, and every request's hit records,
      pixels and ray counts against the oracle over the same trees;
  C5  served requests over three scenes of 1, 5 and 20M triangles: each
      frame the same alone and in the mix, and small requests against the
      oracle.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import checkers as ck
import slimtracin_amd as st
from parity import compare_hits, compare_pixels
from slimtracin_amd.scene import (MODE_BEAUTY, MODE_UVS, PRIM_DTYPE, blob_prims, blob_with_prims,
                                  blob_with_scene_bvh, make_view, quat_axis, quat_mul_norm)

pytestmark = pytest.mark.gpu
THREADS = len(os.sched_getaffinity(0))


def _assert_frame(gp, gh, gc, op, oh, oc, mode, label):
    hs = compare_hits(gh, oh)
    ps = compare_pixels(gp, op, mode)
    print(label, "hits", hs, "pixels", ps, "rays", gc.total(), flush=True)
    assert hs["id_mismatch"] == 0 and hs["t_mismatch"] == 0, (label, hs)
    assert ps["ok"], (label, ps)
    assert (gc.primary, gc.shadow, gc.bounce) == (oc.primary_rays, oc.shadow_rays, oc.bounce_rays), label


def test_c2_pbr_200k_full_frame():
    blob = st.gen_pbr(2024, 100)
    view = make_view(width=1920, height=1080, mode=MODE_BEAUTY, depth=4, position=(0.0, 6.0, -14.0), pitch=-0.15)
    ds = st.DeviceScene(blob)
    try:
        assert ds.info().triangle_count == 200_000
        gp, gh, gc = st.render(ds, view)
    finally:
        ds.close()
    op, oh, oc = ck.oracle_render(blob, view, threads=THREADS)
    _assert_frame(gp, gh, gc, op, oh, oc, MODE_BEAUTY, "c2")


def test_c3_instanced_full_frame():
    blob = st.gen_instanced(2025, 64, [50_000, 200_000, 500_000], 2048)
    prims = blob_prims(blob)
    rng = np.random.default_rng(3)
    for i in range(len(prims) - 1):                          # the last primitive is the floor
        prims["position"][i] = (rng.uniform(-30, 30), rng.uniform(1.5, 6), rng.uniform(-20, 40))
        q = quat_axis((1, 0, 0), float(rng.uniform(-3, 3)))
        q = quat_mul_norm(q, quat_axis((0, 1, 0), float(rng.uniform(-3, 3))))
        prims["rotation"][i] = quat_mul_norm(q, quat_axis((0, 0, 1), float(rng.uniform(-3, 3))))
    prims = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    view = make_view(width=1920, height=1080, mode=MODE_BEAUTY, depth=3, cone_angle=0.001,
                     position=(0.0, 14.0, -35.0), pitch=-0.3)
    import bench
    ds = st.DeviceScene(blob, **bench.INSTANCED_MESH_BUILD)
    try:
        ds.set_prims(prims)
        ds.rebuild(st.ST_REBUILD_LBVH)
        nodes, leaf, h = ds.bvh()
        gp, gh, gc = st.render(ds, view)
    finally:
        ds.close()
    ref_blob = blob_with_scene_bvh(blob_with_prims(blob, prims), nodes, leaf, h)
    osc = ck.OracleScene(ref_blob, **bench.INSTANCED_MESH_BUILD)    # the same mesh trees
    try:
        op, oh, oc, _ = osc.render(view, threads=THREADS)
    finally:
        osc.close()
    assert oc.texel_fetches > 100_000
    _assert_frame(gp, gh, gc, op, oh, oc, MODE_BEAUTY, "c3")


def _batch_workload():
    import bench
    return bench.batch_workload()


def test_c4_batch_of_256_full_size():
    import bench
    blobs, views, sids = _batch_workload()
    ds = [st.DeviceScene(b, **bench.BATCH_MESH_BUILD) for b in blobs]
    try:
        rs = [ds[i] for i in sids]
        bp, bh, bc = st.render_batch(rs, views, hits=True)
        for s in range(len(blobs)):
            idx = [i for i in range(len(views)) if sids[i] == s]
            if not idx:
                continue
            osc = ck.OracleScene(blobs[s], **bench.BATCH_MESH_BUILD)
            for i in idx:
                v = views[i:i + 1]
                pp, ph, pc = st.render(rs[i], v)
                assert pp.tobytes() == bp[i].tobytes() and ph.tobytes() == bh[i].tobytes(), i
                assert (pc.primary, pc.shadow, pc.bounce) == (bc[i].primary, bc[i].shadow, bc[i].bounce), i
                op, oh, oc, _ = osc.render(v, threads=THREADS)
                _assert_frame(bp[i], bh[i], bc[i], op, oh, oc, int(v["render_mode"][0]), f"c4 request {i}")
            osc.close()
    finally:
        for d in ds:
            d.close()


def test_c5_served_large_scenes_alone_equals_mix():
    import bench
    from slimtracin_amd.server import OUTPUT_FLOAT_PIXELS, Server
    sizes = [1_000_000, 5_000_000, 20_000_000]
    blobs = [st.gen_mesh_scene(700 + i, t) for i, t in enumerate(sizes)]
    srv = Server(output=OUTPUT_FLOAT_PIXELS, max_in_flight=3)
    try:
        ids = [srv.add_scene(b, mesh_lbvh=True) for b in blobs]
        views, picks = bench.request_mix(np.random.default_rng(17), len(ids), 24)
        small = np.where(views["width"].astype(np.int64) * views["height"] <= 250_000)[0][:4]
        tickets = [srv.submit(ids[picks[k]], views[k:k + 1]) for k in range(len(views))]
        mixed = []
        for t in tickets:
            img, c, _ = srv.wait(t)
            srv.release(t)
            mixed.append((img, c))
        for k in range(len(views)):
            t = srv.submit(ids[picks[k]], views[k:k + 1])
            img, c, _ = srv.wait(t)
            srv.release(t)
            assert img.tobytes() == mixed[k][0].tobytes(), k
            assert c.primary + c.shadow + c.bounce == mixed[k][1].primary + mixed[k][1].shadow + mixed[k][1].bounce
        stats = srv.stats()
        print("c5 served", {k: stats[k] for k in ("launches", "overlapped_launches", "max_in_flight")}, flush=True)
    finally:
        srv.close()
    assert len(small) >= 2, small
    for k in small:
        osc = ck.OracleScene(blobs[picks[k]], mesh_lbvh=True)
        op, oh, oc, _ = osc.render(views[k:k + 1], threads=THREADS)
        osc.close()
        img, c = mixed[k]
        ps = compare_pixels(img, op, int(views["render_mode"][k]))
        print("c5 request", int(k), ps, flush=True)
        assert ps["ok"], (int(k), ps)
        assert (c.primary, c.shadow, c.bounce) == (oc.primary_rays, oc.shadow_rays, oc.bounce_rays)
