"""Rays with an exactly zero direction component.

With a zero component the reference's slab test (render/AABB.h:6-23, the
origin pre-scaled by the reciprocal direction) computes 0 * inf or
inf - inf = NaN on that axis, and its ternary min/max chains then drop or
keep that axis depending on where the NaN sits (x, y or z). The traversal
visits many more boxes than the ray crosses, and which triangles it tests,
in which order, decides the hit state that later rays read. These views put
a whole pixel row (y = 0) and a pixel column (x = 0), or with the camera
turned a quarter, a column with z = 0, through an unrotated mesh whose
centre is level with the camera (so the object-space origin's y is 0 too).

CPU: the views really contain such rays (the projection plane and the
device's row and column sums restated in float32), and the oracle's primary
hits equal the reference renderer's own (compiled from its sources) on them.
GPU: the device's hits and frames equal the oracle's, over the host SAH mesh
BVH and over the device-built Morton LBVH.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import checkers as ck
from parity import compare_hits, compare_pixels
import slimtracin_amd as st
from slimtracin_amd.scene import (IS_ROTATED, MODE_BEAUTY, MODE_NORMALS, blob_mesh_info, blob_prims,
                                  blob_with_prebuilt_meshes, blob_with_prims, make_view)

TRIANGLES = 20000


def _blob():
    """The mesh scene with its mesh unrotated (centre at y = 3)."""
    blob = st.gen_mesh_scene(5, TRIANGLES)
    prims = blob_prims(blob)
    prims[0]["rotation"] = (0.0, 0.0, 0.0, 1.0)
    prims[0]["flags"] &= ~np.uint32(IS_ROTATED)
    return blob_with_prims(blob, prims)


# odd window sizes put the centre row and column of the traced grid (w+1 x
# h+1) on the camera axis; yaw 1.0 is a quarter turn (rational parameter)
VIEWS = [
    dict(position=(0.0, 3.0, -10.0), width=33, height=21),
    dict(position=(0.0, 3.0, -10.0), width=65, height=49),
    dict(position=(-10.0, 3.0, 0.0), yaw=1.0, width=45, height=27),
]


def _zero_components(blob, view):
    """(x, y, axis) of every traced pixel whose unnormalised primary direction
    has an exactly zero component: start + y*down, then + x*right, summed in
    float32 in that order as the device's row and column sums do."""
    cam = np.zeros(17, np.float32)
    n = int(np.frombuffer(blob[8:12], np.uint32)[0]) + 1
    rects = np.zeros(4 * n, np.int32)
    flags = np.zeros(n, np.uint8)
    assert st.lib().st_host_view_info(blob, C.c_size_t(len(blob)), view.ctypes.data_as(C.c_void_p),
                                      rects.ctypes.data_as(C.c_void_p), flags.ctypes.data_as(C.c_void_p),
                                      cam.ctypes.data_as(C.c_void_p)) == 0
    start, right, down = cam[0:3], cam[3:6], cam[6:9]
    w1, h1 = int(view["width"][0]) + 1, int(view["height"][0]) + 1
    out = []
    row = start.copy()
    for y in range(h1):
        cur = row.copy()
        for x in range(w1):
            out += [(x, y, c) for c in range(3) if cur[c] == 0]
            cur = (cur + right).astype(np.float32)
        row = (row + down).astype(np.float32)
    return out


@pytest.mark.parametrize("k", range(len(VIEWS)))
def test_views_hold_zero_direction_components(k):
    blob = _blob()
    view = make_view(mode=MODE_NORMALS, **VIEWS[k])
    zeros = _zero_components(blob, view)
    axes = {c for _, _, c in zeros}
    assert zeros, "no degenerate ray in the view"
    assert axes == ({0, 1} if k < 2 else {1, 2}), axes


@pytest.mark.skipif(not ck.ref_available(), reason="reference build missing")
@pytest.mark.parametrize("k", range(len(VIEWS)))
def test_oracle_equals_reference_on_degenerate_rays(k):
    blob = _blob()
    view = make_view(mode=MODE_NORMALS, **VIEWS[k])
    ref = ck.ref_primary_hits(blob, view)
    _, ora, _ = ck.oracle_render(blob, view)
    hs = compare_hits(ora, ref)
    assert hs["id_mismatch"] == 0 and hs["t_mismatch"] == 0, hs


def _render_device(blob, view, lbvh):
    ds = st.DeviceScene(blob, mesh_lbvh=lbvh)
    try:
        return st.render(ds, view)
    finally:
        ds.close()


@pytest.mark.gpu
@pytest.mark.parametrize("lbvh", [False, True])
@pytest.mark.parametrize("mode", [MODE_NORMALS, MODE_BEAUTY])
@pytest.mark.parametrize("k", range(len(VIEWS)))
def test_device_equals_oracle_on_degenerate_rays(k, mode, lbvh):
    blob = _blob()
    view = make_view(mode=mode, depth=3, **VIEWS[k])
    want_blob = blob
    if lbvh:
        T = blob_mesh_info(blob)[0]["triangle_count"]
        want_blob = blob_with_prebuilt_meshes(blob, [ck.oracle_mesh_lbvh(blob, 0, T)])
    gp, gh, gc = _render_device(blob, view, lbvh)
    op, oh, oc = ck.oracle_render(want_blob, view)
    hs = compare_hits(gh, oh)
    ps = compare_pixels(gp, op, mode)
    print("view", k, "mode", mode, "lbvh", lbvh, "hits", hs, "pixels", ps, flush=True)
    assert hs["id_mismatch"] == 0 and hs["t_mismatch"] == 0, hs
    assert ps["ok"], ps
    assert (gc.primary, gc.shadow, gc.bounce) == (oc.primary_rays, oc.shadow_rays, oc.bounce_rays)
