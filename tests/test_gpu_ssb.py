"""GPU: the device SSB pass (st_ssb, and the k_ssb stage every render launch
runs) against the oracle's SSB restatement, which the golden fixtures and
the reference build pin (test_oracle_golden.py, test_oracle_reference.py):
rects as computeSSB leaves them and IS_VISIBLE flags, bit for bit, for the
example scenes, a 300-primitive crowd, and cameras placed inside primitive
bounds (the weak-projection branch) and behind them."""
from __future__ import annotations

import numpy as np
import pytest

import checkers as ck
import scenes
import slimtracin_amd as st
from slimtracin_amd.scene import blob_prims, make_view

pytestmark = pytest.mark.gpu

CASES = ["ex01_lights", "ex02_geometry", "ex04_glass_mirror", "ex06_area_lights", "ex07_monkey", "ex09_textures"]


def _check(ds, blob, view, label):
    P = len(blob_prims(blob))
    got_r, got_f = st.device_ssb(ds, view)
    want_r, want_f = ck.oracle_ssb(blob, view, P)
    bad = np.nonzero((got_r != want_r).any(axis=1) | (got_f != want_f))[0]
    assert len(bad) == 0, (label, bad[:5], got_r[bad[:2]], want_r[bad[:2]], got_f[bad[:2]], want_f[bad[:2]])
    return int((got_f & 1).sum())


@pytest.mark.parametrize("name", CASES)
def test_device_ssb_example_scenes(name):
    st.init(0)
    s, cam = getattr(scenes, name)()
    blob = s.to_blob()
    ds = st.DeviceScene(blob)
    try:
        for w, h in ((160, 120), (640, 360), (33, 517)):
            _check(ds, blob, make_view(width=w, height=h, **cam), f"{name} {w}x{h}")
    finally:
        ds.close()


def test_device_ssb_crowd_and_inside_views():
    st.init(0)
    s, cam = scenes.crowd(3, 300)
    blob = s.to_blob()
    prims = blob_prims(blob)
    rng = np.random.default_rng(21)
    ds = st.DeviceScene(blob)
    try:
        visible = 0
        for k in range(24):
            if k % 3 == 0:    # inside or right next to a primitive's bound
                c = prims["position"][rng.integers(0, len(prims))] + rng.uniform(-0.5, 0.5, 3)
            else:
                c = np.array(cam["position"]) + rng.uniform(-8, 8, 3)
            view = make_view(width=int(rng.integers(64, 700)), height=int(rng.integers(48, 500)),
                             position=tuple(float(x) for x in c), yaw=float(rng.uniform(-1, 1)),
                             pitch=float(rng.uniform(-0.5, 0.5)), roll=float(rng.uniform(-0.2, 0.2)),
                             focal_length=float(rng.uniform(0.5, 3.0)))
            visible += _check(ds, blob, view, f"crowd view {k}")
        assert visible > 0
    finally:
        ds.close()
