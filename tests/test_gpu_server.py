"""GPU: the serving loop (include/slimtracin_server.h).

A served request's frame equals the plain per-request launch's frame for the
same scene and view (FloatPixels bit for bit; the window image equals the
resolve of those FloatPixels), whether it ran alone or batched with other
requests of other scenes; admission respects the HBM budget; many threads
can submit at once."""
from __future__ import annotations

import threading

import numpy as np
import pytest

import checkers as ck
import scenes
import slimtracin_amd as st
from slimtracin_amd.scene import MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS, make_view
from slimtracin_amd.server import OUTPUT_BGRA8, OUTPUT_FLOAT_PIXELS, Server

pytestmark = pytest.mark.gpu
MODES = [MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS]


def _scene_set():
    return [scenes.ex04_glass_mirror(), scenes.ex07_monkey(prebuilt=True), scenes.ex09_textures(),
            scenes.crowd(2, 80)]


def _views(rng, cams, n):
    out = []
    for i in range(n):
        sid = int(rng.integers(len(cams)))
        w, h = int(rng.integers(16, 200)), int(rng.integers(16, 150))
        out.append((sid, make_view(width=w, height=h, mode=MODES[int(rng.integers(4))],
                                   **dict(cams[sid], depth=int(rng.integers(1, 5))))))
    return out


@pytest.mark.parametrize("output", [OUTPUT_FLOAT_PIXELS, OUTPUT_BGRA8])
def test_served_frames_equal_plain_frames(output):
    sc = _scene_set()
    blobs = [s.to_blob() for s, _ in sc]
    cams = [{k: v for k, v in c.items() if k != "depth"} for _, c in sc]
    plain = [st.DeviceScene(b) for b in blobs]
    srv = Server(output=output)
    try:
        ids = [srv.add_scene(b) for b in blobs]
        reqs = _views(np.random.default_rng(11), cams, 40)
        tickets = [srv.submit(ids[sid], v) for sid, v in reqs]
        for (sid, v), t in zip(reqs, tickets):
            img, counts, lat = srv.wait(t)
            srv.release(t)
            pix, _, c = st.render(plain[sid], v, hits=False)
            if output == OUTPUT_FLOAT_PIXELS:
                assert img.tobytes() == pix.tobytes()
            else:
                w1, h1 = int(v["width"][0]) + 1, int(v["height"][0]) + 1
                assert (img == ck.oracle_resolve(pix, w1, h1)).all()
            assert counts.primary + counts.shadow + counts.bounce == c.total()
            assert lat >= 0
        stats = srv.stats()
        print(stats, flush=True)
        assert stats["completed"] == 40 and stats["launches"] >= 1 and stats["resident_scenes"] == 4
    finally:
        srv.close()
        for d in plain:
            d.close()


def test_alone_equals_in_mix():
    sc = _scene_set()
    blobs = [s.to_blob() for s, _ in sc]
    cams = [{k: v for k, v in c.items() if k != "depth"} for _, c in sc]
    srv = Server(output=OUTPUT_FLOAT_PIXELS)
    try:
        ids = [srv.add_scene(b) for b in blobs]
        reqs = _views(np.random.default_rng(5), cams, 24)
        mixed = [srv.submit(ids[sid], v) for sid, v in reqs]     # queued together -> batched
        mixed_imgs = []
        for t in mixed:
            mixed_imgs.append(srv.wait(t)[0])
            srv.release(t)
        assert srv.stats()["launches"] < 24
        for (sid, v), img in zip(reqs, mixed_imgs):
            t = srv.submit(ids[sid], v)                           # alone
            alone = srv.wait(t)[0]
            srv.release(t)
            assert alone.tobytes() == img.tobytes()
    finally:
        srv.close()


def test_hbm_budget_admission():
    small = scenes.ex02_geometry()[0].to_blob()
    big = st.gen_mesh_scene(3, 60000)
    probe = st.DeviceScene(big)
    big_bytes = probe.info().device_bytes
    probe.close()
    srv = Server(hbm_budget=big_bytes + big_bytes // 2)
    try:
        a = srv.add_scene(big)
        assert srv.try_add_scene(big) is None                    # second copy does not fit
        b = srv.add_scene(small)                                  # a small one does
        st_ = srv.stats()
        assert st_["resident_scenes"] == 2 and st_["resident_bytes"] <= st_["budget_bytes"]
        t = srv.submit(a, make_view(width=64, height=48, position=(0.0, 3.0, -8.0)))
        srv.remove_scene(a)                                       # waits for the queued request
        img, counts, _ = srv.wait(t)
        srv.release(t)
        assert counts.primary == 65 * 49
        c = srv.add_scene(big)                                    # fits again after removal
        assert c == a
        with pytest.raises(st.SlimTracinError):
            srv.submit(99, make_view(width=8, height=8))
        with pytest.raises(st.SlimTracinError):
            srv.submit(b, make_view(width=8, height=8, mode=9))
    finally:
        srv.close()


def test_many_submitting_threads():
    blob = scenes.ex06_area_lights()[0].to_blob()
    cam = {k: v for k, v in scenes.ex06_area_lights()[1].items() if k != "depth"}
    srv = Server(output=OUTPUT_BGRA8)
    ref = st.DeviceScene(blob)
    try:
        sid = srv.add_scene(blob)
        results = {}
        errors = []

        def client(k):
            try:
                v = make_view(width=40 + k, height=30 + k, **cam)
                t = srv.submit(sid, v)
                img, _, _ = srv.wait(t, timeout_ms=60000)
                srv.release(t)
                results[k] = (v, img)
            except Exception as e:  # noqa: BLE001
                errors.append(e)

        th = [threading.Thread(target=client, args=(k,)) for k in range(16)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors, errors
        for k, (v, img) in results.items():
            pix, _, _ = st.render(ref, v, hits=False)
            assert (img == ck.oracle_resolve(pix, int(v["width"][0]) + 1, int(v["height"][0]) + 1)).all(), k
    finally:
        srv.close()
        ref.close()


def test_pinned_outputs_equal_pageable():
    s, cam = scenes.ex09_textures()
    ds = st.DeviceScene(s.to_blob())
    try:
        views = np.concatenate([make_view(width=50 + 7 * i, height=30 + 5 * i, mode=MODES[i % 4], **cam)
                                for i in range(5)])
        grid = [(int(v["width"]) + 1) * (int(v["height"]) + 1) for v in views]
        fp = st.pinned_split(grid, st.PIXEL_DTYPE)
        bg = st.pinned_split([int(v["width"]) * int(v["height"]) for v in views], np.uint32)
        a, _, _ = st.render_batch([ds] * 5, views, outs=fp)
        b, _, _ = st.render_batch([ds] * 5, views)
        c, _ = st.render_batch_bgra8([ds] * 5, views, outs=bg)
        d, _ = st.render_batch_bgra8([ds] * 5, views)
        for i in range(5):
            one, _, _ = st.render(ds, views[i:i + 1], hits=False, out=st.pinned_empty(grid[i], st.PIXEL_DTYPE))
            assert a[i].tobytes() == b[i].tobytes() == one.tobytes()
            assert (c[i] == d[i]).all()
    finally:
        ds.close()


def test_streamed_arrivals_overlap_and_equal_plain():
    """Requests arriving one by one while launches run: free workers launch
    them at once on their own streams (launches overlap on the device) and
    every frame still equals the plain launch's."""
    import time
    sc = _scene_set()
    blobs = [s.to_blob() for s, _ in sc]
    cams = [{k: v for k, v in c.items() if k != "depth"} for _, c in sc]
    plain = [st.DeviceScene(b) for b in blobs]
    srv = Server(output=OUTPUT_FLOAT_PIXELS, max_in_flight=3)
    try:
        ids = [srv.add_scene(b) for b in blobs]
        reqs = _views(np.random.default_rng(23), cams, 60)
        tickets = []
        for sid, v in reqs:
            tickets.append(srv.submit(ids[sid], v))
            time.sleep(0.0005)
        for (sid, v), t in zip(reqs, tickets):
            img, counts, _ = srv.wait(t)
            srv.release(t)
            pix, _, c = st.render(plain[sid], v, hits=False)
            assert img.tobytes() == pix.tobytes()
            assert counts.primary + counts.shadow + counts.bounce == c.total()
        stats = srv.stats()
        print("streamed:", {k: stats[k] for k in ("launches", "overlapped_launches", "max_in_flight", "mean_batch")},
              flush=True)
        assert stats["completed"] == 60
        assert 1 <= stats["max_in_flight"] <= 3
    finally:
        srv.close()
        for d in plain:
            d.close()


def test_plain_limits_one_request_per_launch():
    """set_limits(1, 1): one request per launch, one launch at a time (the
    plain baseline over the same resident scenes), then back to batching."""
    sc = _scene_set()[:2]
    blobs = [s.to_blob() for s, _ in sc]
    cams = [{k: v for k, v in c.items() if k != "depth"} for _, c in sc]
    srv = Server(output=OUTPUT_BGRA8)
    try:
        ids = [srv.add_scene(b) for b in blobs]
        reqs = _views(np.random.default_rng(4), cams, 12)
        views = np.concatenate([v for _, v in reqs])
        r = srv.plain_rate(views, [ids[sid] for sid, _ in reqs])
        stats = srv.stats()
        assert r["requests"] == 12 and stats["launches"] == 12 and stats["max_in_flight"] == 1
        assert stats["overlapped_launches"] == 0
        ts = [srv.submit(ids[sid], v) for sid, v in reqs]
        for t in ts:
            srv.wait(t, copy=False)
            srv.release(t)
        assert srv.stats()["completed"] == 24
    finally:
        srv.close()


def test_limited_workers_take_requests_submitted_while_idle():
    """With set_limits(1, 1) only worker 0 of the server's 3 may take work.
    Each request here is submitted while every worker is idle and waited for
    before the next: the submission must wake worker 0 (a single wake-up of
    a worker that may not take work left such a request queued until the
    next submission, and the last one for good). Waits are bounded, so a
    request left queued fails the test instead of holding it."""
    sc = _scene_set()[:1]
    blob = sc[0][0].to_blob()
    cam = {k: v for k, v in sc[0][1].items() if k != "depth"}
    srv = Server(output=OUTPUT_BGRA8)
    try:
        sid = srv.add_scene(blob)
        reqs = _views(np.random.default_rng(5), [cam], 24)
        for limits in ((1, 1), (0, 2), (0, 0)):
            srv.set_limits(*limits)
            for _, v in reqs:
                t = srv.submit(sid, v)
                srv.wait(t, timeout_ms=20000, copy=False)
                srv.release(t)
        assert srv.stats()["completed"] == 3 * len(reqs)
    finally:
        srv.set_limits(0, 0)
        srv.close()
