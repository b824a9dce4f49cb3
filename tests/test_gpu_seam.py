"""GPU: the reference seam (include/slimtracin_seam.h) driven the way the
reference's host drives its device mirror.

The scene, its SAH BVHs, the projection plane and the SSB rects are what the
reference's own host code computes (here: the oracle's restatement of it,
pinned against the reference build elsewhere); the frames must equal the CPU
renderer's: hit records bit-exact, pixels to the Beauty tolerance.
"""
from __future__ import annotations

import numpy as np
import pytest

import checkers as ck
import scenes
import slimtracin_amd as st
from parity import compare_hits, compare_pixels
from slimtracin_amd.scene import (
    MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS, blob_prims, blob_records, blob_with_prims, blob_with_scene_bvh,
    make_view, mesh_aabb,
)
from slimtracin_amd.seam import SeamScene, frame_from_camera

pytestmark = pytest.mark.gpu


def _mesh_uploads(s, blob):
    out = []
    for m, mesh in enumerate(s.meshes):
        nodes, tris, leaf, h = ck.oracle_mesh_bvh(blob, m, len(mesh.indices))
        aabb = mesh.aabb if mesh.aabb is not None else mesh_aabb(mesh.positions)
        out.append(dict(aabb=aabb, nodes=nodes, triangles=tris, height=h,
                        normals_count=0 if mesh.normals is None else len(mesh.normals),
                        uvs_count=0 if mesh.uvs is None else len(mesh.uvs)))
    return out


def _textures(s):
    return [(t.width, t.height, t.wrap, t.mipmap, t.mips) for t in s.textures]


def _frame(blob, view):
    rec = blob_records(blob)
    return frame_from_camera(ck.oracle_camera(view), view["position"][0], view["cone_angle"][0], rec["ambient"],
                             int(view["width"][0]) + 1, int(view["height"][0]) + 1, int(view["render_mode"][0]),
                             int(view["depth"][0]))


def _ssb(blob, view):
    rec = blob_records(blob)
    rects, flags = ck.oracle_ssb(blob, view, rec["prim_count"])
    prims = rec["prims"].copy()
    prims["flags"] = flags
    return prims, rects


def _upload_all(seam, s, blob, view):
    rec = blob_records(blob)
    prims, rects = _ssb(blob, view)
    seam.upload_scene(rec["lights"], prims, rects, rec["materials"], _textures(s))
    if s.meshes:
        seam.upload_mesh_bvhs(_mesh_uploads(s, blob))
    nodes, leaf, h = ck.oracle_scene_bvh(blob, rec["prim_count"])
    seam.upload_scene_bvh(nodes, leaf, h)


def _compare(seam, blob, view, label):
    gp, gh, gc = seam.render(_frame(blob, view))
    op, oh, oc = ck.oracle_render(blob, view)
    hs = compare_hits(gh, oh)
    ps = compare_pixels(gp, op, int(view["render_mode"][0]))
    print(label, "hits", hs, "pixels", ps, flush=True)
    assert hs["id_mismatch"] == 0 and hs["t_mismatch"] == 0, (label, hs)
    assert ps["ok"], (label, ps)
    assert (gc.primary, gc.shadow, gc.bounce) == (oc.primary_rays, oc.shadow_rays, oc.bounce_rays), label


def _new(s):
    return SeamScene(len(s.primitives), len(s.materials), len(s.lights), len(s.meshes), len(s.textures))


CASES = {
    "ex02": lambda: scenes.ex02_geometry(),
    "ex04": lambda: scenes.ex04_glass_mirror(),
    "ex06": lambda: scenes.ex06_area_lights(),
    "ex07": lambda: scenes.ex07_monkey(prebuilt=True),
    "ex09": lambda: scenes.ex09_textures(),
    "crowd": lambda: scenes.crowd(3, 120),
}


@pytest.mark.parametrize("name", list(CASES))
def test_seam_frame_matches_cpu_renderer(name):
    s, cam = CASES[name]()
    blob = s.to_blob()
    seam = _new(s)
    try:
        modes = [MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS] if name in ("ex07", "ex09") else [MODE_BEAUTY]
        for mode in modes:
            view = make_view(width=144, height=96, mode=mode, **cam)
            _upload_all(seam, s, blob, view)
            _compare(seam, blob, view, f"{name} mode={mode}")
    finally:
        seam.close()


def test_seam_animation_uploads():
    """Per-frame uploadPrimitives with a stale scene BVH, then an
    uploadSceneBVH with a Morton LBVH, then new materials."""
    s, cam = scenes.crowd(5, 100)
    blob = s.to_blob()
    view = make_view(width=128, height=96, mode=MODE_BEAUTY, **cam)
    seam = _new(s)
    try:
        _upload_all(seam, s, blob, view)
        _compare(seam, blob, view, "initial")
        old_nodes, old_leaf, old_h = ck.oracle_scene_bvh(blob, len(s.primitives))
        rng = np.random.default_rng(9)
        prims = blob_prims(blob)
        for frame in range(2):
            prims = prims.copy()
            prims["position"][1:, 1] += np.float32(0.2)
            moved = blob_with_prims(blob, prims)
            p2, rects = _ssb(moved, view)
            seam.upload_primitives(p2, rects)
            _compare(seam, blob_with_scene_bvh(moved, old_nodes, old_leaf, old_h), view, f"stale BVH frame {frame}")
        moved = blob_with_prims(blob, scenes.move_prims(prims, rng))
        p2, rects = _ssb(moved, view)
        seam.upload_primitives(p2, rects)
        lb = ck.oracle_lbvh_build(ck.oracle_prim_boxes(moved, len(s.primitives)))
        seam.upload_scene_bvh(*lb)
        _compare(seam, blob_with_scene_bvh(moved, *lb), view, "LBVH upload")
        # uploadMaterials: swap two materials' records
        s2, _ = scenes.crowd(5, 100)
        s2.materials[0], s2.materials[3] = s2.materials[3], s2.materials[0]
        s2.primitives = s.primitives
        blob2 = blob_with_prims(s2.to_blob(), blob_prims(moved))
        seam.upload_materials(blob_records(blob2)["materials"])
        _compare(seam, blob_with_scene_bvh(blob2, *lb), view, "material upload")
    finally:
        seam.close()


def test_seam_errors():
    s, cam = scenes.ex02_geometry()
    blob = s.to_blob()
    view = make_view(width=32, height=24, **cam)
    seam = _new(s)
    try:
        with pytest.raises(st.SlimTracinError):
            seam.render(_frame(blob, view))                       # nothing uploaded yet
        rec = blob_records(blob)
        bad = rec["prims"].copy()
        bad["material_id"][1] = 7
        with pytest.raises(st.SlimTracinError):
            seam.upload_primitives(bad, np.zeros((len(bad), 4), np.int32))
        _upload_all(seam, s, blob, view)
        f = _frame(blob, view)
        f.render_mode = 9
        with pytest.raises(st.SlimTracinError):
            seam.render(f)
        _compare(seam, blob, view, "after errors")
    finally:
        seam.close()
