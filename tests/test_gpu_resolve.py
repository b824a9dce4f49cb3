"""GPU: the device display resolve (QCAA 2x2 average + sqrt -> BGRA8) against
the oracle's restatement of drawViewportToWindowContent (itself pinned to
the reference's compiled function in test_resolve_oracle.py): bit-exact on
edge-case grids and on rendered frames, single and batched."""
from __future__ import annotations

import numpy as np
import pytest

import checkers as ck
import scenes
import slimtracin_amd as st
from slimtracin_amd.scene import MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS, make_view

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", range(6))
def test_device_resolve_edge_cases(seed):
    rng = np.random.default_rng(100 + seed)
    w1, h1 = int(rng.integers(2, 300)), int(rng.integers(2, 200))
    g = ck.resolve_test_grid(rng, w1, h1)
    got = st.resolve_bgra8(g, w1, h1)
    want = ck.oracle_resolve(g, w1, h1)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, (w1, h1, bad[:5], got[bad[:5]], want[bad[:5]])


def test_device_resolve_every_channel_value():
    """All 65026 integer channel values and their float neighbours."""
    c = np.arange(0, 65026, dtype=np.float32)
    c = np.concatenate([c, np.nextafter(c, np.float32(1e9)), np.nextafter(c[1:], np.float32(-1))])
    m = 1000
    k = -(-len(c) // m)
    vals = np.zeros(k * m, dtype=np.float32)
    vals[:len(c)] = c
    V = vals.reshape(k, m)
    G = np.repeat(np.repeat(V, 2, axis=0), 2, axis=1)      # every value fills a 2x2 block of corners
    h1, w1 = G.shape
    g = np.zeros(h1 * w1, dtype=ck.PIXEL_DTYPE)
    g["opacity"], g["depth"] = 1, 1
    g["g"] = G.reshape(-1)
    got = st.resolve_bgra8(g, w1, h1)
    want = ck.oracle_resolve(g, w1, h1)
    assert (got == want).all()
    chan = ((want.reshape(h1 - 1, w1 - 1)[0::2, 0::2] >> 8) & 0xFF).reshape(-1)[:len(c)]
    assert (chan == np.floor(np.sqrt(c.astype(np.float64))).astype(np.uint32)).all()


def test_batched_resolve_matches_host_resolve_of_same_frames():
    sc = [scenes.ex04_glass_mirror(), scenes.ex06_area_lights(), scenes.ex09_textures()]
    dscenes = [st.DeviceScene(s.to_blob()) for s, _ in sc]
    try:
        views, handles = [], []
        modes = [MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS]
        for i in range(8):
            s, cam = sc[i % 3]
            views.append(make_view(width=64 + 16 * i, height=40 + 8 * i, mode=modes[i % 4], **cam))
            handles.append(dscenes[i % 3])
        views = np.concatenate(views)
        outs, counts = st.render_batch_bgra8(handles, views)
        for i in range(len(views)):
            pix, _, c = st.render(handles[i], views[i:i + 1], hits=False)
            w1, h1 = int(views["width"][i]) + 1, int(views["height"][i]) + 1
            assert (outs[i] == ck.oracle_resolve(pix, w1, h1)).all(), i
            assert counts[i].total() == c.total()
    finally:
        for d in dscenes:
            d.close()
