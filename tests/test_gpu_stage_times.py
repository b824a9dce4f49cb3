"""GPU: per-stage timing (st_last_stage_ms): one duration per stage launch
in the order stage_names lists them, each positive, together no longer than
the whole timed launch."""
from __future__ import annotations

import pytest

import scenes
import slimtracin_amd as st
from slimtracin_amd.scene import MODE_BEAUTY, MODE_DEPTH, make_view

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,depth", [(MODE_BEAUTY, 4), (MODE_BEAUTY, 1), (MODE_DEPTH, 1)])
def test_stage_times_match_stage_names(mode, depth):
    s, cam = scenes.ex04_glass_mirror()
    cam = {k: v for k, v in cam.items() if k != "depth"}
    ds = st.DeviceScene(s.to_blob())
    try:
        view = make_view(width=160, height=90, mode=mode, depth=depth, **cam)
        total, counts, stages = st.render_timed([ds], view, stages=True)
        levels = depth if mode == MODE_BEAUTY else 0
        names = st.stage_names(levels)
        assert len(stages) == len(names), (stages, names)
        assert all(t > 0 for t in stages), stages
        assert sum(stages) <= total * 1.01 + 0.01, (sum(stages), total)
    finally:
        ds.close()
