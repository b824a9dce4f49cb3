"""GPU: mesh BVHs built on the device (st_scene_create_ex with
ST_SCENE_MESH_LBVH) against the oracle's restatement of the same Morton
LBVH: node slots reachable from the root, leaf order, triangle records and
height bit for bit; frames over the device-built trees equal the oracle's
frames over the same trees. A 1M-triangle build is compared in full."""
from __future__ import annotations

import time

import numpy as np
import pytest

import checkers as ck
import slimtracin_amd as st
from parity import compare_hits, compare_pixels
from slimtracin_amd.scene import MODE_BEAUTY, MODE_NORMALS, blob_mesh_info, blob_with_prebuilt_meshes, make_view

pytestmark = pytest.mark.gpu


def _compare_build(blob, label):
    T = blob_mesh_info(blob)[0]["triangle_count"]
    t0 = time.perf_counter()
    ds = st.DeviceScene(blob, mesh_lbvh=True)
    create_s = time.perf_counter() - t0
    try:
        nodes, tris, h = ds.mesh_bvh(0, T)
    finally:
        ds.close()
    onodes, otris, oleaf, oh = ck.oracle_mesh_lbvh(blob, 0, T)
    reach = ck.reachable_nodes(onodes)
    print(label, "T", T, "slots", len(nodes), "height", h, "create_s", round(create_s, 3), flush=True)
    assert len(nodes) == len(onodes) and h == oh, (label, len(nodes), len(onodes), h, oh)
    bad = np.nonzero(nodes[reach].view(np.uint8).reshape(len(reach), -1)
                     != onodes[reach].view(np.uint8).reshape(len(reach), -1))[0]
    assert nodes[reach].tobytes() == onodes[reach].tobytes(), (label, reach[np.unique(bad)[:5]])
    assert tris.tobytes() == otris.tobytes(), label
    return (onodes, otris, oleaf, oh)


def _tiny_mesh_blob(T, seed=0):
    """One primitive holding a random T-triangle mesh (some triangles flat
    along an axis, so their boxes are padded)."""
    from slimtracin_amd.scene import PRIM_MESH, Light, Material, Mesh, Primitive, Scene
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, (3 * T, 3)).astype(np.float32)
    pos[0:3, 1] = np.float32(0.25)                       # first triangle flat in y
    idx = np.arange(3 * T, dtype=np.uint32).reshape(T, 3)
    s = Scene(ambient=(0.01, 0.01, 0.01))
    s.lights = [Light((0, 5, -5), (1, 1, 1), 50.0)]
    s.materials = [Material()]
    s.meshes = [Mesh(pos, idx, normals=pos.copy(), normal_indices=idx.copy(),
                     uvs=pos[:, :2].copy(), uv_indices=idx.copy())]
    s.primitives = [Primitive(PRIM_MESH, mesh_id=0)]
    return s.to_blob()


@pytest.mark.parametrize("tris", [1, 2, 3, 4, 5, 9])
def test_device_mesh_lbvh_tiny(tris):
    _compare_build(_tiny_mesh_blob(tris, tris), f"tiny {tris}")


@pytest.mark.parametrize("tris", [17, 2000, 30000, 200000])
def test_device_mesh_lbvh_equals_oracle(tris):
    _compare_build(st.gen_mesh_scene(31, tris), f"mesh {tris}")


def test_device_mesh_lbvh_one_million():
    _compare_build(st.gen_mesh_scene(32, 1_000_000), "mesh 1M")


@pytest.mark.parametrize("mode", [MODE_BEAUTY, MODE_NORMALS])
def test_frames_over_device_built_meshes(mode):
    blob = st.gen_mesh_scene(33, 60000)
    built = _compare_build(blob, "frame mesh")
    pre = blob_with_prebuilt_meshes(blob, [built])
    view = make_view(position=(0.0, 5.0, -13.0), pitch=-0.12, width=160, height=120, mode=mode, depth=3)
    ds = st.DeviceScene(blob, mesh_lbvh=True)
    try:
        gp, gh, gc = st.render(ds, view)
    finally:
        ds.close()
    op, oh, oc = ck.oracle_render(pre, view)
    hs = compare_hits(gh, oh)
    ps = compare_pixels(gp, op, mode)
    print("hits", hs, "pixels", ps, flush=True)
    assert hs["id_mismatch"] == 0 and hs["t_mismatch"] == 0, hs
    assert ps["ok"], ps
    assert (gc.primary, gc.shadow, gc.bounce) == (oc.primary_rays, oc.shadow_rays, oc.bounce_rays)
