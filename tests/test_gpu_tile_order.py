"""GPU: the primary stage claims its tiles in an order taken from each tile's
time in the library's previous launch (slowest first). The order must be a
permutation of the tiles whatever that launch was, so a frame is the same
after a launch of the same shape, after one of another shape (whose tile
times index other tiles) and in a persistent or one-wave-per-tile launch."""
from __future__ import annotations

import numpy as np
import pytest

import scenes
import slimtracin_amd as st
from slimtracin_amd.scene import MODE_BEAUTY, make_view

pytestmark = pytest.mark.gpu


def _frame(ds, view):
    pixels, hits, counts = st.render(ds, view)
    return np.array(pixels, copy=True), np.array(hits, copy=True), counts


def test_frames_do_not_depend_on_the_previous_launch():
    s, cam = scenes.ex04_glass_mirror()
    cam = {k: v for k, v in cam.items() if k != "depth"}
    ds = st.DeviceScene(s.to_blob())
    try:
        a = make_view(width=200, height=120, mode=MODE_BEAUTY, depth=4, **cam)
        b = make_view(width=96, height=300, mode=MODE_BEAUTY, depth=2, **cam)
        p0, h0, c0 = _frame(ds, a)
        runs = [_frame(ds, a)]                   # order from a launch of the same shape
        _frame(ds, b)
        runs.append(_frame(ds, a))               # order from another shape's tile times
        saved = st.get_tuning()
        try:
            for persistent in (True, False):
                st.set_tuning(tile_shift=3, blocks_per_cu=4 if persistent else 8, persistent=persistent,
                              slowest_first=True)
                runs.append(_frame(ds, a))
        finally:
            st.restore_tuning(saved)
        for k, (p, h, c) in enumerate(runs):
            assert p.tobytes() == p0.tobytes(), k
            assert h.tobytes() == h0.tobytes(), k
            assert (c.primary, c.shadow, c.bounce) == (c0.primary, c0.shadow, c0.bounce), k
    finally:
        ds.close()
