"""GPU: every launch shape, tile claim (one counter or XCD chunks), claim
order and bounce-ray order the autotuner may pick renders the same frames
bit for bit (tiles only change which wavefront traces a pixel, the claims,
the claim order and the bounce-ray sort only when), and the autotuner keeps
one of its candidates."""
from __future__ import annotations

import numpy as np
import pytest

import scenes
import slimtracin_amd as st
from slimtracin_amd.scene import MODE_BEAUTY, MODE_NORMALS, make_view

pytestmark = pytest.mark.gpu


def test_all_launch_shapes_give_identical_frames():
    s, cam = scenes.ex09_textures()
    ds = st.DeviceScene(s.to_blob())
    saved = st.get_tuning()
    try:
        views = np.concatenate([make_view(width=131, height=77, mode=MODE_BEAUTY, **cam),
                                make_view(width=64, height=200, mode=MODE_NORMALS, **cam)])
        ref = None
        for shift in (3, 4, 5):
            for bpc in (2, 3, 8):
                for persistent, xcd in ((True, False), (True, True), (False, False)):
                    for order in ("tiles", "slowest_first", "super_tiles"):
                        for ray_sort in (False, True):
                            # the bounce queue's parts in either order (the sort replaces them)
                            st.set_tuning(shift, bpc, persistent, order == "slowest_first", ray_sort, xcd_claim=xcd,
                                          super_tiles=order == "super_tiles", refracting_first=(shift + bpc) % 2 == 0)
                            for _ in range(2):   # the second launch has its predecessor's tile times
                                pix, hits, counts = st.render_batch([ds, ds], views, hits=True)
                                got = b"".join(p.tobytes() for p in pix) + b"".join(h.tobytes() for h in hits)
                                got += b"".join(np.array([c.shadow, c.bounce], np.uint64).tobytes() for c in counts)
                                if ref is None:
                                    ref = got
                                assert got == ref, (shift, bpc, persistent, xcd, order, ray_sort)
        best, ms = st.autotune([ds, ds], views, reps=3)
        assert best.tile_shift in (3, 4, 5) and ms > 0
        print("autotune:", best.as_dict(), f"{ms:.3f} ms", flush=True)
    finally:
        st.restore_tuning(saved)
        ds.close()


def test_xcd_claim_gives_identical_frames():
    """A batch over three scenes with enough tiles per wave (one block per
    CU) that the XCD-aware claim is used: the frames equal those of the
    one-counter claim, in tile and super-tile order, with the bounce stage
    starting on either part of its queue."""
    built = [scenes.ex04_glass_mirror(), scenes.ex09_textures(), scenes.ex07_monkey()]
    ds = [st.DeviceScene(s.to_blob()) for s, _ in built]
    saved = st.get_tuning()
    try:
        rs, views = [], []
        for k in range(6):   # scenes interleaved in batch order: 0 1 2 0 1 2
            s, cam = built[k % 3]
            rs.append(ds[k % 3])
            views.append(make_view(width=1024 - 64 * k, height=640 + 16 * k, mode=MODE_BEAUTY, **cam))
        views = np.concatenate(views)
        # 8x8 tiles over the traced (width + 1) x (height + 1) grids: at least
        # 32 per wave of a one-block-per-CU grid on a 256-CU MI355X
        tiles = sum((int(v["width"]) + 8) // 8 * ((int(v["height"]) + 8) // 8) for v in views)
        assert tiles >= 32 * 4 * 256, tiles
        ref = None
        for xcd in (False, True):
            for super_tiles in (False, True):
                # the glass-and-mirror scene fills both parts of the bounce queue
                st.set_tuning(3, 1, True, False, False, 16, 4, xcd_claim=xcd, super_tiles=super_tiles,
                              refracting_first=super_tiles)
                pix, hits, counts = st.render_batch(rs, views, hits=True)
                got = b"".join(p.tobytes() for p in pix) + b"".join(h.tobytes() for h in hits)
                got += b"".join(np.array([c.shadow, c.bounce], np.uint64).tobytes() for c in counts)
                if ref is None:
                    ref = got
                assert got == ref, (xcd, super_tiles)
    finally:
        st.restore_tuning(saved)
        for d in ds:
            d.close()
