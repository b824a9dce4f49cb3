"""GPU: picking (st_pick) against the oracle's restatement of
manipulateSelection's pick branch (viewport/manipulation.h:55-124), which
tests/test_pick_oracle.py pins to the reference's own code: object type
and id, hit point and view-space distance bit for bit, on the example
scenes, a mesh scene and points on light glow spheres."""
from __future__ import annotations

import numpy as np
import pytest

import checkers as ck
import scenes
import slimtracin_amd as st
from slimtracin_amd.scene import PRIM_LIGHT, make_view

pytestmark = pytest.mark.gpu

CASES = ["ex01_lights", "ex02_geometry", "ex04_glass_mirror", "ex06_area_lights", "ex07_monkey", "ex09_textures"]


def _grid(w, h, step):
    ys, xs = np.mgrid[0:h + 1:step, 0:w + 1:step]
    return np.stack([xs.ravel(), ys.ravel()], axis=1).astype(np.uint32)


@pytest.mark.parametrize("name", CASES)
def test_pick_matches_oracle(name):
    st.init(0)
    s, cam = getattr(scenes, name)()
    blob = s.to_blob()
    view = make_view(width=200, height=150, **cam)
    xy = _grid(200, 150, 2)
    ds = st.DeviceScene(blob)
    try:
        got = st.pick(ds, view, xy)
    finally:
        ds.close()
    want = ck.oracle_pick(blob, view, xy)
    same = got.view(np.uint8).reshape(len(xy), -1) == want.view(np.uint8).reshape(len(xy), -1)
    rows = np.nonzero(~same.all(axis=1))[0]
    print(name, "points", len(xy), "hits", int((got["object_type"] > 0).sum()), "mismatches", len(rows), flush=True)
    assert len(rows) == 0, (name, xy[rows[:4]], got[rows[:2]], want[rows[:2]])


def test_pick_lights_and_mesh_scene():
    st.init(0)
    s, cam = scenes.ex01_lights()
    blob = s.to_blob()
    view = make_view(width=96, height=72, **cam)
    ds = st.DeviceScene(blob)
    try:
        xy = _grid(96, 72, 1)
        got = st.pick(ds, view, xy)
    finally:
        ds.close()
    want = ck.oracle_pick(blob, view, xy)
    assert got.tobytes() == want.tobytes()
    assert (got["object_type"] == PRIM_LIGHT).any()
    blob = st.gen_pbr(3, 30)
    view = make_view(width=128, height=96, position=(0.0, 6.0, -14.0), pitch=-0.15)
    ds = st.DeviceScene(blob)
    try:
        xy = _grid(128, 96, 2)
        got = st.pick(ds, view, xy)
    finally:
        ds.close()
    want = ck.oracle_pick(blob, view, xy)
    assert got.tobytes() == want.tobytes()
