"""Picking (viewport/manipulation.h:55-124): the oracle's restatement
against the reference's own manipulateSelection compiled from its sources
(oracle/_ref), bit for bit: object type and id (primitives and light glow
spheres), world hit point and view-space distance, over a grid of window
points on the example scenes and random views."""
from __future__ import annotations

import numpy as np
import pytest

import checkers as ck
import scenes
from slimtracin_amd.scene import PRIM_LIGHT, make_view

pytestmark = pytest.mark.skipif(not (ck.ref_available() and ck.oracle_available()),
                                reason="reference build (oracle/_ref) or oracle missing")

CASES = ["ex01_lights", "ex02_geometry", "ex04_glass_mirror", "ex06_area_lights", "ex07_monkey", "ex09_textures"]


def _grid(w, h, step):
    ys, xs = np.mgrid[0:h + 1:step, 0:w + 1:step]
    return np.stack([xs.ravel(), ys.ravel()], axis=1).astype(np.uint32)


def _same(a, b):
    return a.tobytes() == b.tobytes()


@pytest.mark.parametrize("name", CASES)
def test_pick_matches_reference(name):
    s, cam = getattr(scenes, name)()
    blob = s.to_blob()
    view = make_view(width=160, height=120, **cam)
    xy = _grid(160, 120, 3)
    got, want = ck.oracle_pick(blob, view, xy), ck.ref_pick(blob, view, xy)
    bad = np.nonzero(got.view(np.uint8).reshape(len(xy), -1).any(axis=1)
                     != want.view(np.uint8).reshape(len(xy), -1).any(axis=1))[0]
    assert _same(got, want), (name, bad[:5], got[bad[:3]] if len(bad) else None)
    assert (got["object_type"] > 0).mean() > 0.3, name       # the grid hits things


def test_pick_lights_and_random_views():
    s, cam = scenes.ex01_lights()
    blob = s.to_blob()
    rng = np.random.default_rng(12)
    n_light = 0
    for k in range(6):
        pos = (cam["position"][0] + rng.uniform(-2, 2), cam["position"][1] + rng.uniform(-1, 1),
               cam["position"][2] + rng.uniform(-2, 2))
        view = make_view(width=96, height=72, position=pos, yaw=float(rng.uniform(-0.2, 0.2)),
                         pitch=float(rng.uniform(-0.2, 0.1)))
        xy = _grid(96, 72, 2)
        got, want = ck.oracle_pick(blob, view, xy), ck.ref_pick(blob, view, xy)
        assert _same(got, want), k
        n_light += int((got["object_type"] == PRIM_LIGHT).sum())
    assert n_light > 0            # some points land on a light's glow sphere
