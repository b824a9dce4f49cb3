"""GPU: the instanced textured workload - mesh instances moving every frame,
the scene BVH rebuilt on the device by the Morton LBVH each frame, Perlin
albedo and normal maps with full mip chains sampled through ray cones - 
against the oracle rendering the same moved scene over the same BVH."""
from __future__ import annotations

import numpy as np
import pytest

import checkers as ck
import scenes
import slimtracin_amd as st
from parity import compare_hits, compare_pixels
from slimtracin_amd.scene import (
    MODE_BEAUTY, MODE_UVS, blob_prims, blob_with_prims, blob_with_scene_bvh, make_view,
)

pytestmark = pytest.mark.gpu
CAM = dict(position=(0.0, 14.0, -35.0), pitch=-0.3)


def _move(prims, rng):
    out = prims.copy()
    for i in range(len(out) - 1):                       # the last primitive is the floor
        out["position"][i] = (rng.uniform(-30, 30), rng.uniform(1.5, 6), rng.uniform(-20, 40))
        out["rotation"][i] = scenes.random_rotation(rng)
    return out


@pytest.mark.parametrize("cone", [0.0, 0.002])
def test_instanced_frames_with_per_frame_lbvh(cone):
    blob = st.gen_instanced(9, 64, [3000, 8000, 20000], 256)
    prims = blob_prims(blob)
    rng = np.random.default_rng(12)
    ds = st.DeviceScene(blob)
    try:
        for frame in range(3):
            prims = _move(prims, rng)
            ds.set_prims(prims)
            ms = ds.rebuild(st.ST_REBUILD_LBVH)
            nodes, leaf, h = ds.bvh()
            ref_blob = blob_with_scene_bvh(blob_with_prims(blob, prims), nodes, leaf, h)
            for mode in (MODE_BEAUTY, MODE_UVS):
                view = make_view(width=192, height=108, mode=mode, depth=3, cone_angle=cone, **CAM)
                gp, gh, gc = st.render(ds, view)
                op, oh, oc = ck.oracle_render(ref_blob, view)
                hs = compare_hits(gh, oh)
                ps = compare_pixels(gp, op, mode)
                print(f"frame {frame} mode {mode} cone {cone} lbvh {ms:.3f} ms", hs, ps, oc.texel_fetches, flush=True)
                assert hs["id_mismatch"] == 0 and hs["t_mismatch"] == 0, hs
                assert ps["ok"], ps
                assert (gc.primary, gc.shadow, gc.bounce) == (oc.primary_rays, oc.shadow_rays, oc.bounce_rays)
                if mode == MODE_BEAUTY:
                    assert oc.texel_fetches > 1000
    finally:
        ds.close()
