"""GPU: HIP-graph frames (ST_BATCH_GRAPH).

A launch captured once and replayed while its shape repeats must give the
frames and ray counts of eager launches of the same views, on every replay,
with eager work of other shapes in between (which must not disturb the
replayed graph's buffers), and a change of shape must capture a new graph.
"""
from __future__ import annotations

import numpy as np
import pytest

import slimtracin_amd as st
from slimtracin_amd.scene import MODE_BEAUTY, MODE_NORMALS, make_view

pytestmark = pytest.mark.gpu


def _views(k, w=200, h=120, mode=MODE_BEAUTY, depth=4):
    return make_view(width=w, height=h, mode=mode, depth=depth, position=(0.5 * k - 1.0, 6.0 + 0.25 * k, -14.0),
                     yaw=0.05 * k, pitch=-0.15)


def _same(a, b):
    return a.tobytes() == b.tobytes()


def test_graph_replays_equal_eager_launches():
    st.init(0)
    ds = st.DeviceScene(st.gen_pbr(7, 16))
    other = st.DeviceScene(st.gen_implicit(3))
    try:
        c0, l0 = st.graph_stats()
        for k in range(5):
            v = _views(k)
            ep, eh, ec = st.render_batch([ds], v, hits=True)
            print(f"frame {k}: eager done", flush=True)
            gp, gh, gc = st.render_batch([ds], v, hits=True, graph=True)
            print(f"frame {k}: graph done", flush=True)
            assert _same(ep[0], gp[0]) and _same(eh[0], gh[0]), k
            assert (ec[0].primary, ec[0].shadow, ec[0].bounce) == (gc[0].primary, gc[0].shadow, gc[0].bounce), k
            # eager work of another shape (and scene) between the replays
            st.render_batch([other], _views(k, 96, 64, MODE_NORMALS, 2))
        c1, l1 = st.graph_stats()
        assert c1 - c0 == 1, (c0, c1)          # one capture, four replays
        assert l1 - l0 == 5, (l0, l1)
        # a new shape captures a new graph; the old shape again, another
        v2 = _views(1, 128, 96)
        ep, _, _ = st.render_batch([ds], v2)
        gp, _, _ = st.render_batch([ds], v2, graph=True)
        assert _same(ep[0], gp[0])
        c2, _ = st.graph_stats()
        assert c2 - c1 == 1
        gp, _, _ = st.render_batch([ds], _views(2), graph=True)
        ep, _, _ = st.render_batch([ds], _views(2))
        assert _same(ep[0], gp[0])
        assert st.graph_stats()[0] - c2 == 1
    finally:
        ds.close()
        other.close()


def test_graph_batched_requests():
    st.init(0)
    a = st.DeviceScene(st.gen_pbr(8, 12))
    b = st.DeviceScene(st.gen_implicit(4))
    try:
        for k in range(3):
            views = np.concatenate([_views(k, 160, 90), _views(k + 1, 120, 80, MODE_NORMALS, 3), _views(k, 96, 96)])
            sc = [a, b, a]
            ep, _, ec = st.render_batch(sc, views)
            gp, _, gc = st.render_batch(sc, views, graph=True)
            for i in range(3):
                assert _same(ep[i], gp[i]), (k, i)
                assert (ec[i].shadow, ec[i].bounce) == (gc[i].shadow, gc[i].bounce), (k, i)
    finally:
        a.close()
        b.close()
