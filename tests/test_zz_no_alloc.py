"""st_render_noalloc (include/slimtracin.h): the plain launch with every
device buffer carved from the library's fixed __device__ scratch array. Its
frames, hit records and ray counts must equal st_render's bit for bit, on
calls in a row whose second is smaller (the array keeps the first call's
contents), and a request too large for the array fails with an error before
anything runs."""
from __future__ import annotations

import pytest

import scenes
import slimtracin_amd as st
from slimtracin_amd.scene import MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS, make_view

pytestmark = pytest.mark.gpu


def _same(ds, view, hits=True):
    a_pix, a_hits, a_c = st.render(ds, view, hits=hits)
    b_pix, b_hits, b_c = st.render_noalloc(ds, view, hits=hits)
    assert a_pix.tobytes() == b_pix.tobytes()
    if hits:
        assert a_hits.tobytes() == b_hits.tobytes()
    assert (a_c.primary, a_c.shadow, a_c.bounce) == (b_c.primary, b_c.shadow, b_c.bounce)
    return b_c


def test_noalloc_equals_render_on_calls_in_a_row():
    st.init(0)
    implicit = st.DeviceScene(st.gen_implicit(1))
    s, cam = scenes.ex07_monkey()
    monkey = st.DeviceScene(s.to_blob())
    try:
        c = _same(implicit, make_view(width=640, height=480, mode=MODE_BEAUTY, depth=2, position=(0.0, 7.0, -11.0),
                                      pitch=-0.2))
        assert c.shadow > 0
        c = _same(monkey, make_view(width=200, height=150, mode=MODE_BEAUTY, depth=4, **{k: v for k, v in cam.items()
                                                                                           if k != "depth"}))
        assert c.shadow > 0
        _same(monkey, make_view(width=96, height=64, mode=MODE_NORMALS, **cam), hits=False)
        _same(implicit, make_view(width=33, height=17, mode=MODE_UVS, position=(0.0, 7.0, -11.0), pitch=-0.2))
    finally:
        implicit.close()
        monkey.close()


def test_noalloc_rejects_a_request_beyond_the_scratch_array():
    s, cam = scenes.ex07_monkey()
    ds = st.DeviceScene(s.to_blob())
    try:
        with pytest.raises(st.SlimTracinError, match="bytes of scratch"):
            st.render_noalloc(ds, make_view(width=1920, height=1080, mode=MODE_BEAUTY, depth=4,
                                            **{k: v for k, v in cam.items() if k != "depth"}))
        _same(ds, make_view(width=64, height=48, mode=MODE_BEAUTY, **cam))
    finally:
        ds.close()


def test_noalloc_with_the_xcd_claim():
    """A 1080p Depth frame at one block per CU has 32 tiles per wave, so
    both launches claim tiles per XCD: the frames still agree."""
    s, cam = scenes.ex07_monkey()
    ds = st.DeviceScene(s.to_blob())
    saved = st.get_tuning()
    try:
        st.set_tuning(3, 1, True, False, False, 16, 4, xcd_claim=True)
        view = make_view(width=1920, height=1080, mode=MODE_DEPTH, **{k: v for k, v in cam.items() if k != "depth"})
        assert ((1920 + 8) // 8) * ((1080 + 8) // 8) >= 32 * 4 * 256
        _same(ds, view)
    finally:
        st.restore_tuning(saved)
        ds.close()
