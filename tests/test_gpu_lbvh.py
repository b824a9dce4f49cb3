"""GPU: scene-BVH rebuild after primitives move.

The device Morton LBVH (st_scene_rebuild, ST_REBUILD_LBVH) must equal the
oracle's LBVH over the oracle's primitive boxes bit for bit (nodes, leaf
order, height), and frames rendered over it must equal the oracle's frames of
the same scene carrying that BVH. The host SAH rebuild must equal the
reference builder's tree for the moved primitives.
"""
from __future__ import annotations

import numpy as np
import pytest

import checkers as ck
import scenes
import slimtracin_amd as st
from parity import compare_hits, compare_pixels
from slimtracin_amd.scene import (
    MODE_BEAUTY, MODE_NORMALS, PRIM_SPHERE, Light, Material, Primitive, Scene, blob_prims, blob_with_prims,
    blob_with_scene_bvh, make_view,
)

pytestmark = pytest.mark.gpu


def _frame_parity(ds, blob, view, label):
    gp, gh, gc = st.render(ds, view)
    op, oh, oc = ck.oracle_render(blob, view)
    hs = compare_hits(gh, oh)
    ps = compare_pixels(gp, op, int(view["render_mode"][0]))
    print(label, "hits", hs, "pixels", ps, flush=True)
    assert hs["id_mismatch"] == 0 and hs["t_mismatch"] == 0, (label, hs)
    assert ps["ok"], (label, ps)
    assert (gc.primary, gc.shadow, gc.bounce) == (oc.primary_rays, oc.shadow_rays, oc.bounce_rays), label


def _check_lbvh(ds, moved, label):
    P = len(blob_prims(moved))
    nodes, leaf, h = ds.bvh()
    want_nodes, want_leaf, want_h = ck.oracle_lbvh_build(ck.oracle_prim_boxes(moved, P))
    print(label, "nodes", len(nodes), "height", h, flush=True)
    assert len(nodes) == len(want_nodes) == 2 * P - 1, label
    assert (leaf == want_leaf).all(), label
    assert h == want_h, label
    bad = np.nonzero(nodes.view(np.uint8).reshape(len(nodes), -1) != want_nodes.view(np.uint8).reshape(len(nodes), -1))
    assert nodes.tobytes() == want_nodes.tobytes(), (label, np.unique(bad[0])[:8])
    return nodes, leaf, h


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_device_lbvh_rebuild_and_render(seed):
    s, cam = scenes.crowd(seed, 300)
    blob = s.to_blob()
    prims = blob_prims(blob)
    rng = np.random.default_rng(100 + seed)
    ds = st.DeviceScene(blob)
    try:
        for frame in range(3):
            prims = scenes.move_prims(prims, rng)
            ds.set_prims(prims)
            ms = ds.rebuild(st.ST_REBUILD_LBVH)
            moved = blob_with_prims(blob, prims)
            label = f"crowd seed={seed} frame={frame} build_ms={ms:.3f}"
            nodes, leaf, h = _check_lbvh(ds, moved, label)
            with_bvh = blob_with_scene_bvh(moved, nodes, leaf, h)
            for mode in (MODE_BEAUTY, MODE_NORMALS):
                view = make_view(width=160, height=120, mode=mode, **cam)
                _frame_parity(ds, with_bvh, view, f"{label} mode={mode}")
    finally:
        ds.close()


def _spheres(n, seed, duplicates=False):
    rng = np.random.default_rng(seed)
    s = Scene(ambient=(0.01, 0.01, 0.01))
    s.lights = [Light((0, 30, -10), (1, 1, 1), 200.0)]
    s.materials = [Material(albedo=(0.8, 0.6, 0.4))]
    for i in range(n):
        if duplicates:
            pos = (float(rng.integers(-3, 4)), float(rng.integers(0, 3)), float(rng.integers(0, 6)))
        else:
            pos = scenes.random_position(rng)
        s.primitives.append(Primitive(PRIM_SPHERE, position=pos, scale=tuple(rng.uniform(0.1, 0.6, 3))))
    return s.to_blob()


@pytest.mark.parametrize("n,dup", [(1, False), (2, False), (3, True), (17, False), (64, True), (1000, False),
                                   (4096, False), (4097, True), (20000, False), (150000, True)])
def test_device_lbvh_sizes(n, dup):
    """Single primitive, tiny trees, equal Morton codes (duplicate centres),
    and scenes of many sort tiles and radix-tree blocks (the build is
    multi-block at every step)."""
    blob = _spheres(n, n, dup)
    ds = st.DeviceScene(blob)
    try:
        ds.rebuild(st.ST_REBUILD_LBVH)
        nodes, leaf, h = _check_lbvh(ds, blob, f"spheres n={n} dup={dup}")
        if n <= 1000 or n == 20000:
            view = make_view(position=(0.0, 12.0, -25.0), pitch=-0.3, width=96, height=64, mode=MODE_NORMALS)
            _frame_parity(ds, blob_with_scene_bvh(blob, nodes, leaf, h), view, f"spheres n={n}")
    finally:
        ds.close()


def test_host_sah_rebuild_after_move():
    s, cam = scenes.crowd(7, 200)
    blob = s.to_blob()
    prims = scenes.move_prims(blob_prims(blob), np.random.default_rng(3))
    moved = blob_with_prims(blob, prims)
    ds = st.DeviceScene(blob)
    try:
        ds.set_prims(prims)
        ds.rebuild(st.ST_REBUILD_SAH)
        nodes, leaf, h = ds.bvh()
        want_nodes, want_leaf, want_h = ck.oracle_scene_bvh(moved, len(prims))
        assert nodes.tobytes() == want_nodes.tobytes() and (leaf == want_leaf).all() and h == want_h
        _frame_parity(ds, moved, make_view(width=128, height=96, mode=MODE_BEAUTY, **cam), "crowd SAH rebuild")
        # back to the LBVH and again to SAH on the same handle
        ds.rebuild(st.ST_REBUILD_LBVH)
        _check_lbvh(ds, moved, "crowd LBVH after SAH")
        ds.rebuild(st.ST_REBUILD_SAH)
        assert ds.bvh()[0].tobytes() == want_nodes.tobytes()
    finally:
        ds.close()


def test_moved_prims_without_rebuild_keep_stale_bvh():
    """uploadPrimitives without updateSceneBVH (the reference's 02_Geometry
    animation): the device traverses the old tree with the new transforms,
    exactly as the oracle does with the old tree carried in the blob."""
    s, cam = scenes.crowd(4, 150)
    blob = s.to_blob()
    P = len(s.primitives)
    old_nodes, old_leaf, old_h = ck.oracle_scene_bvh(blob, P)
    prims = blob_prims(blob)
    prims["position"][1:, 1] += np.float32(0.25)          # small lift of everything but the floor
    moved = blob_with_prims(blob, prims)
    ds = st.DeviceScene(blob)
    try:
        ds.set_prims(prims)
        _frame_parity(ds, blob_with_scene_bvh(moved, old_nodes, old_leaf, old_h),
                      make_view(width=128, height=96, mode=MODE_NORMALS, **{k: v for k, v in cam.items()
                                                                            if k != "depth"}), "stale BVH")
    finally:
        ds.close()
