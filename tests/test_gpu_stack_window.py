"""GPU: traversal stacks kept in an LDS window with the older entries in the
launch's spill buffer (StTuning.stack_entries), with the primary stage at 3
and 4 waves per SIMD (StTuning.primary_waves). With windows of 4, 8 and 16
slots every deep traversal spills and reloads entries, and frames, hit
records and ray counts equal the whole-stack launch's and the oracle's: over
host SAH and device LBVH mesh BVHs, in a batch of several requests, and for
pixels the redo stage traces again."""
from __future__ import annotations

import numpy as np
import pytest

import checkers as ck
import scenes
import slimtracin_amd as st
from parity import compare_hits, compare_pixels
from slimtracin_amd.scene import (MODE_BEAUTY, MODE_NORMALS, blob_mesh_info, blob_with_prebuilt_meshes, make_view)

pytestmark = pytest.mark.gpu


def _with_window(win, fn, waves=3):
    saved = st.get_tuning()
    try:
        st.set_tuning(3, 4, True, False, False, win, waves)
        return fn()
    finally:
        st.restore_tuning(saved)


@pytest.mark.parametrize("lbvh", [False, True])
def test_windows_equal_whole_stacks_and_oracle(lbvh):
    blob = st.gen_mesh_scene(11, 60000)
    want = blob
    if lbvh:
        T = blob_mesh_info(blob)[0]["triangle_count"]
        want = blob_with_prebuilt_meshes(blob, [ck.oracle_mesh_lbvh(blob, 0, T)])
    views = np.concatenate([
        make_view(width=96, height=72, mode=MODE_BEAUTY, depth=4, position=(4.0, 7.0, -12.0), yaw=0.15, pitch=-0.25),
        make_view(width=80, height=60, mode=MODE_NORMALS, depth=2, position=(-3.0, 5.0, -11.0), yaw=-0.1,
                  pitch=-0.2)])
    ds = st.DeviceScene(blob, mesh_lbvh=lbvh)
    try:
        ref = _with_window(0, lambda: st.render_batch([ds, ds], views, hits=True))
        for win, waves in ((4, 3), (8, 4), (16, 4), (16, 3)):
            got = _with_window(win, lambda: st.render_batch([ds, ds], views, hits=True), waves)
            for i in range(len(views)):
                assert got[0][i].tobytes() == ref[0][i].tobytes(), (win, waves, i)
                assert got[1][i].tobytes() == ref[1][i].tobytes(), (win, waves, i)
                assert (got[2][i].shadow, got[2][i].bounce) == (ref[2][i].shadow, ref[2][i].bounce), (win, waves, i)
    finally:
        ds.close()
    for i in range(len(views)):
        op, oh, oc = ck.oracle_render(want, views[i:i + 1])
        hs = compare_hits(ref[1][i], oh)
        ps = compare_pixels(ref[0][i], op, int(views["render_mode"][i]))
        assert hs["id_mismatch"] == 0 and hs["t_mismatch"] == 0, (i, hs)
        assert ps["ok"], (i, ps)
        assert (ref[2][i].shadow, ref[2][i].bounce) == (oc.shadow_rays, oc.bounce_rays), i


def test_windows_in_the_redo_stage():
    """Every pixel through the redo stage (give-up after one step), with a
    window of 4 slots: the redo stage's sequential traversals spill too."""
    s, cam = scenes.ex07_monkey()
    view = make_view(width=64, height=48, mode=MODE_BEAUTY, **dict(cam, depth=3))
    ds = st.DeviceScene(s.to_blob())
    st.debug_redo(1, 16)
    try:
        ref = _with_window(0, lambda: st.render(ds, view))
        got = _with_window(4, lambda: st.render(ds, view))
        assert st.last_redo_count() > 100
    finally:
        st.debug_redo(0, 0)
        ds.close()
    assert got[0].tobytes() == ref[0].tobytes() and got[1].tobytes() == ref[1].tobytes()
    op, oh, oc = ck.oracle_render(s.to_blob(), view)
    assert compare_hits(got[1], oh)["id_mismatch"] == 0
    assert compare_pixels(got[0], op, MODE_BEAUTY)["ok"]
