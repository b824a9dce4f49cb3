"""The CPU oracle against the reference CPU renderer's golden fixtures.

The fixtures (tests/golden/reference_cases.npz, written by
tests/golden/make_golden.py from the reference compiled from its own
sources) hold, per case: the FloatPixel frame, primary hit records, SSB
rects/flags, camera set-up, scene BVH and (mesh scenes) the mesh BVH and
triangle records. Everything must match bit for bit.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import checkers as ck

FIX = os.path.join(os.path.dirname(__file__), "golden", "reference_cases.npz")

pytestmark = pytest.mark.skipif(not ck.oracle_available(), reason="oracle not built (make -C oracle)")


def _cases():
    with np.load(FIX, allow_pickle=False) as z:
        return [str(n) for n in z["names"]]


@pytest.fixture(scope="module")
def fixtures():
    with np.load(FIX, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", _cases())
def test_frame_and_hits(fixtures, name):
    blob = fixtures[f"{name}/blob"].tobytes()
    view = fixtures[f"{name}/view"]
    pixels, hits, _ = ck.oracle_render(blob, view)
    want = fixtures[f"{name}/pixels"]
    bad, nbad = ck.first_pixel_mismatches(pixels, want)
    assert nbad == 0, (name, nbad, bad)
    wh = fixtures[f"{name}/hits"]
    assert (hits["prim_id"] == wh["prim_id"]).all()
    assert (hits["object_type"] == wh["object_type"]).all()
    assert (hits["triangle_id"] == wh["triangle_id"]).all()
    assert (hits["distance"].view(np.uint32) == wh["distance"].view(np.uint32)).all()


@pytest.mark.parametrize("name", _cases())
def test_ssb_camera_bvh(fixtures, name):
    blob = fixtures[f"{name}/blob"].tobytes()
    view = fixtures[f"{name}/view"]
    rects = fixtures[f"{name}/rects"]
    P = len(rects)
    r, f = ck.oracle_ssb(blob, view, P)
    assert (f == fixtures[f"{name}/flags"]).all()
    assert (r == rects).all()
    assert ck.oracle_camera(view).tobytes() == fixtures[f"{name}/camera"].tobytes()
    nodes, leaf, height = ck.oracle_scene_bvh(blob, P)
    assert nodes.tobytes() == fixtures[f"{name}/scene_nodes"].tobytes()
    assert (leaf == fixtures[f"{name}/scene_leaf"]).all()
    assert height == int(fixtures[f"{name}/scene_height"][0])


def test_mesh_bvh_build(fixtures):
    """updateMeshBVH on monkey.mesh's own geometry reproduces the reference build."""
    name = "monkey_built"
    blob = fixtures[f"{name}/blob"].tobytes()
    tris = fixtures[f"{name}/mesh_tris"]
    n, t, l, h = ck.oracle_mesh_bvh(blob, 0, len(tris))
    assert n.tobytes() == fixtures[f"{name}/mesh_nodes"].tobytes()
    assert t.tobytes() == tris.tobytes()
    assert (l == fixtures[f"{name}/mesh_leaf"]).all()
    assert h == int(fixtures[f"{name}/mesh_height"][0])
