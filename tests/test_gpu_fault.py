"""A traversal that runs past its bound fails its launch with an error
instead of looping (the reference's traversal loops, render/shaders/trace.h:8-76
and intersection/mesh.h:38-111, have no bound; on a GPU an endless loop
would hold the device). A valid tree never reaches the bound (each node pair
is loaded at most once, the stack holds at most the tree's height), so the
test forces it from the host with st_debug_step_cap and checks that every
kernel still exits, the launch reports ST_ERR_SCENE, and the next launch
without the cap renders the same frame as before."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import scenes
import slimtracin_amd as st
from slimtracin_amd.scene import MODE_BEAUTY, MODE_NORMALS, make_view

pytestmark = pytest.mark.gpu


def _cap(n):
    assert st.lib().st_debug_step_cap(C.c_uint32(n)) == 0


@pytest.mark.parametrize("mode", [MODE_NORMALS, MODE_BEAUTY])
def test_step_cap_fails_the_launch_and_clears(mode):
    s, cam = scenes.ex07_monkey()
    ds = st.DeviceScene(s.to_blob())
    view = make_view(width=96, height=72, mode=mode, **cam)
    try:
        before, bh, _ = st.render(ds, view)
        _cap(2)
        try:
            with pytest.raises(st.SlimTracinError, match="ran past its bound"):
                st.render(ds, view)
            with pytest.raises(st.SlimTracinError, match="ran past its bound"):
                ys, xs = np.mgrid[0:72:3, 0:96:3]
                st.pick(ds, view, np.stack([xs.ravel(), ys.ravel()], axis=1))
        finally:
            _cap(0)
        after, ah, _ = st.render(ds, view)
    finally:
        ds.close()
    assert before.tobytes() == after.tobytes()
    assert bh.tobytes() == ah.tobytes()
