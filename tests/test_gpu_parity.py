"""GPU parity: the HIP library against the CPU oracle on identical scenes.

Scenes: the ported reference examples (every render mode), the seeded
implicit-primitive workload, and a small noise-displaced icosphere PBR scene.
Ray counters of the device must equal the oracle's exactly.
"""
from __future__ import annotations

import numpy as np
import pytest

import checkers as ck
import scenes
import slimtracin_amd as st
from parity import compare_hits, compare_pixels
from slimtracin_amd.scene import MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS, make_view

pytestmark = pytest.mark.gpu

MODES = [MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS]
CAM = dict(position=(0.0, 7.0, -11.0), pitch=-0.2)


def _check_frame(blob, view, label):
    ds = st.DeviceScene(blob)
    try:
        gp, gh, gc = st.render(ds, view)
    finally:
        ds.close()
    op, oh, oc = ck.oracle_render(blob, view)
    mode = int(view["render_mode"][0])
    hs = compare_hits(gh, oh)
    ps = compare_pixels(gp, op, mode)
    print(label, "hits", hs, "pixels", ps, flush=True)
    assert hs["id_mismatch"] == 0 and hs["t_mismatch"] == 0, (label, hs)
    assert ps["ok"], (label, ps)
    assert (gc.primary, gc.shadow, gc.bounce) == (oc.primary_rays, oc.shadow_rays, oc.bounce_rays), \
        (label, (gc.primary, gc.shadow, gc.bounce), (oc.primary_rays, oc.shadow_rays, oc.bounce_rays))


EXAMPLES = ["ex01", "ex02", "ex03", "ex04", "ex05", "ex06"]


@pytest.mark.parametrize("name", EXAMPLES)
@pytest.mark.parametrize("mode", MODES)
def test_examples(name, mode):
    s, cam = scenes.ALL_EXAMPLES[name]()
    view = make_view(width=160, height=120, mode=mode, **cam)
    _check_frame(s.to_blob(), view, f"{name}/m{mode}")


@pytest.mark.parametrize("variant", ["ex02t", "ex06t"])
def test_transparency(variant):
    s, cam = (scenes.ex02_geometry(True) if variant == "ex02t" else scenes.ex06_area_lights(True))
    view = make_view(width=160, height=120, mode=MODE_BEAUTY, **cam)
    _check_frame(s.to_blob(), view, variant)


@pytest.mark.parametrize("mode", MODES)
def test_monkey_mesh(mode):
    s, cam = scenes.ex07_monkey()
    view = make_view(width=160, height=120, mode=mode, **cam)
    _check_frame(s.to_blob(), view, f"monkey/m{mode}")


@pytest.mark.parametrize("mips", [False, True])
@pytest.mark.parametrize("mode", [MODE_BEAUTY, MODE_NORMALS])
def test_textures(mips, mode):
    s, cam = scenes.ex09_textures(mips)
    view = make_view(width=160, height=120, mode=mode, cone_angle=0.002, **cam)
    _check_frame(s.to_blob(), view, f"tex/mips{mips}/m{mode}")


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_implicit_seeded(seed):
    blob = st.gen_implicit(seed)
    view = make_view(width=160, height=120, mode=MODE_BEAUTY, depth=2, **CAM)
    _check_frame(blob, view, f"implicit/{seed}")


def test_implicit_full_frame():
    blob = st.gen_implicit(7)
    view = make_view(width=640, height=480, mode=MODE_BEAUTY, depth=2, **CAM)
    _check_frame(blob, view, "implicit/640x480")


@pytest.mark.parametrize("mode", MODES)
def test_pbr_small(mode):
    blob = st.gen_pbr(5, 12)
    view = make_view(width=192, height=108, mode=mode, depth=4, position=(0.0, 6.0, -14.0), pitch=-0.15)
    _check_frame(blob, view, f"pbr/m{mode}")


def test_batch_equals_plain_and_oracle():
    blobs = [st.gen_implicit(11), st.gen_pbr(12, 8), scenes.ex04_glass_mirror()[0].to_blob()]
    dss = [st.DeviceScene(b) for b in blobs]
    try:
        views = np.concatenate([
            make_view(width=97, height=61, mode=MODE_BEAUTY, depth=2, **CAM),
            make_view(width=64, height=64, mode=MODE_BEAUTY, depth=4, position=(0.0, 6.0, -14.0), pitch=-0.15),
            make_view(width=33, height=150, mode=MODE_NORMALS, depth=4, **CAM),
            make_view(width=120, height=40, mode=MODE_DEPTH, depth=1, **CAM),
        ])
        sc = [dss[0], dss[1], dss[2], dss[0]]
        bl = [blobs[0], blobs[1], blobs[2], blobs[0]]
        for persistent in (True, False):
            pix, hits, counts = st.render_batch(sc, views, hits=True, persistent=persistent)
            for i in range(len(views)):
                v = views[i:i + 1]
                alone, ahits, acounts = st.render(sc[i], v)
                assert pix[i].tobytes() == alone.tobytes(), ("batch vs plain pixels", i, persistent)
                assert hits[i].tobytes() == ahits.tobytes(), ("batch vs plain hits", i, persistent)
                assert (counts[i].primary, counts[i].shadow, counts[i].bounce) == \
                    (acounts.primary, acounts.shadow, acounts.bounce)
                op, oh, oc = ck.oracle_render(bl[i], v)
                assert compare_hits(hits[i], oh)["id_mismatch"] == 0
                assert compare_pixels(pix[i], op, int(v["render_mode"][0]))["ok"]
    finally:
        for d in dss:
            d.close()
