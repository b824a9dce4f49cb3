"""GPU: the device the library reports, and a workload tuned once through
the store starts at the stored shape in a later call (no autotuning)."""
from __future__ import annotations

import numpy as np
import pytest

import scenes
import slimtracin_amd as st
from slimtracin_amd import tuning_store as ts
from slimtracin_amd.scene import MODE_BEAUTY, MODE_NORMALS, make_view

pytestmark = pytest.mark.gpu


def test_device_info():
    d = st.device_info()
    assert d["arch"].startswith("gfx950"), d
    assert d["compute_units"] > 0 and d["hbm_bytes"] > 2**34 and d["name"], d


def test_autotune_cached_round_trip(tmp_path):
    path = str(tmp_path / "tuning.json")
    s, cam = scenes.ex04_glass_mirror()
    ds = st.DeviceScene(s.to_blob())
    saved = st.get_tuning()
    try:
        views = np.concatenate([make_view(width=160, height=90, mode=MODE_BEAUTY, **cam),
                                make_view(width=96, height=96, mode=MODE_NORMALS, **cam)])
        t1, ms1, hit1 = ts.autotune_cached([ds, ds], views, path=path, reps=1)
        assert not hit1 and ms1 > 0
        assert ts.tuning_to_dict(st.get_tuning()) == ts.tuning_to_dict(t1)
        st.restore_tuning(saved)
        t2, ms2, hit2 = ts.autotune_cached([ds, ds], views, path=path, reps=1)
        assert hit2 and ms2 == pytest.approx(ms1, abs=1e-4)
        assert ts.tuning_to_dict(t2) == ts.tuning_to_dict(t1) == ts.tuning_to_dict(st.get_tuning())
        data = ts.load(path)
        assert list(data["gpus"]) == [ts.gpu_key()]
        # another workload is tuned and stored beside it
        _, _, hit3 = ts.autotune_cached([ds], views[:1], path=path, reps=1)
        assert not hit3 and len(ts.load(path)["gpus"][ts.gpu_key()]) == 2
    finally:
        st.restore_tuning(saved)
        ds.close()
