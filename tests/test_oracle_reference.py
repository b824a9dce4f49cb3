"""The CPU oracle against the live reference build on fresh seeded inputs.

Runs where oracle/_ref/libslimref.so exists (it is compiled from the
reference's own sources by oracle/Makefile). Complements the fixed golden
cases with randomised scenes, views and multi-threaded row splits.
"""
from __future__ import annotations

import numpy as np
import pytest

import checkers as ck
import scenes
import slimtracin_amd as st
from slimtracin_amd.scene import (MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS, f32, make_view, quat_axis,
                                  quat_mul_norm)

pytestmark = pytest.mark.skipif(not (ck.ref_available() and ck.oracle_available()),
                                reason="reference build (oracle/_ref) or oracle missing")


@pytest.mark.parametrize("seed", range(6))
def test_implicit_scenes_random_views(seed):
    rng = np.random.default_rng(seed)
    blob = st.gen_implicit(100 + seed)
    for _ in range(3):
        pos = tuple(float(x) for x in rng.uniform([-12, 2, -16], [12, 12, -6]))
        view = make_view(width=int(rng.integers(20, 90)), height=int(rng.integers(20, 70)),
                         mode=int(rng.integers(0, 4)), depth=int(rng.integers(1, 5)), position=pos,
                         yaw=float(rng.uniform(-0.4, 0.4)), pitch=float(rng.uniform(-0.4, 0.1)),
                         roll=float(rng.uniform(-0.1, 0.1)))
        ref = ck.ref_render(blob, view)
        ora, oh, _ = ck.oracle_render(blob, view)
        bad, nbad = ck.first_pixel_mismatches(ref, ora)
        assert nbad == 0, (seed, bad)
        rh = ck.ref_primary_hits(blob, view)
        assert (rh["prim_id"] == oh["prim_id"]).all()
        assert (rh["distance"].view(np.uint32) == oh["distance"].view(np.uint32)).all()


@pytest.mark.parametrize("mode", [MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS])
def test_pbr_mesh_scene(mode):
    blob = st.gen_pbr(9, 10)
    view = make_view(width=64, height=40, mode=mode, depth=4, position=(0.0, 6.0, -14.0), pitch=-0.15)
    ref = ck.ref_render(blob, view)
    ora, oh, _ = ck.oracle_render(blob, view)
    assert ck.pixels_equal(ref, ora)
    rh = ck.ref_primary_hits(blob, view)
    assert (rh["triangle_id"] == oh["triangle_id"]).all()


def test_threads_match_single_thread():
    """Row-interleaved threads with one trace state each give the single-thread frame."""
    s, cam = scenes.ex04_glass_mirror()
    blob = s.to_blob()
    view = make_view(width=80, height=60, **cam)
    a = ck.ref_render(blob, view, threads=1)
    b = ck.ref_render(blob, view, threads=4)
    c, _, _ = ck.oracle_render(blob, view, threads=3)
    assert ck.pixels_equal(a, b) and ck.pixels_equal(a, c)


def test_quaternion_helpers_match_reference():
    for axis, amt in [((1, 0, 0), 0.02), ((0, 1, 0), 0.04), ((0, 0, 1), -0.7), ((0.6, 0.8, 0), 0.33)]:
        q = np.array(quat_axis(axis, amt), dtype=np.float32)
        assert q.tobytes() == ck.ref_quat_axis(axis, amt).tobytes()
    a = quat_axis((1, 0, 0), f32(0.3))
    b = quat_axis((0, 1, 0), f32(-0.2))
    assert np.array(quat_mul_norm(a, b), dtype=np.float32).tobytes() == ck.ref_quat_mul_norm(a, b).tobytes()

