"""The display resolve (QCAA 2x2 average, sqrt, BGRA8) of the oracle against
the reference's own drawViewportToWindowContent (viewport/viewport.h:12-66),
compiled from its sources: rendered frames and edge-case grids. Alpha is
compared only where the reference defines it (the reference leaves it
uninitialised until the first background pixel; the oracle defines 255)."""
from __future__ import annotations

import numpy as np
import pytest

import checkers as ck
import scenes
from slimtracin_amd.scene import MODE_BEAUTY, MODE_NORMALS, make_view

pytestmark = pytest.mark.skipif(not (ck.oracle_available() and ck.ref_available()),
                                reason="oracle or reference build missing")
RGB = np.uint32(0x00FFFFFF)


def _same(a, b):
    assert ((a & RGB) == (b & RGB)).all(), np.nonzero((a & RGB) != (b & RGB))[0][:10]
    assert (a >> 24 == 255).all()


@pytest.mark.parametrize("name,mode", [("ex04", MODE_BEAUTY), ("ex06", MODE_BEAUTY), ("ex02", MODE_NORMALS)])
def test_resolve_of_rendered_frames(name, mode):
    s, cam = {"ex04": scenes.ex04_glass_mirror, "ex06": scenes.ex06_area_lights,
              "ex02": scenes.ex02_geometry}[name]()
    view = make_view(width=120, height=80, mode=mode, **cam)
    pix, _, _ = ck.oracle_render(s.to_blob(), view)
    _same(ck.oracle_resolve(pix, 121, 81), ck.ref_resolve(pix, 121, 81))


@pytest.mark.parametrize("seed", range(4))
def test_resolve_edge_cases(seed):
    rng = np.random.default_rng(seed)
    w1, h1 = int(rng.integers(2, 90)), int(rng.integers(2, 70))
    g = ck.resolve_test_grid(rng, w1, h1)
    a = ck.oracle_resolve(g, w1, h1)
    b = ck.ref_resolve(g, w1, h1)
    _same(a, b)
    # background pixels: the reference sets alpha there too
    bg = (g["depth"].reshape(h1, w1) == np.inf)
    allbg = (bg[:-1, :-1] & bg[:-1, 1:] & bg[1:, :-1] & bg[1:, 1:]).reshape(-1)
    assert (a[allbg] == 0xFF000000).all() and (b[allbg] == 0xFF000000).all()


def test_resolve_channel_is_floor_sqrt():
    """Closed form: a channel is floor(sqrt(c)) for 0 <= c <= 255^2, 255 above."""
    c = np.concatenate([np.arange(0, 65026, 7, dtype=np.float32),
                        np.nextafter(np.arange(1, 256, dtype=np.float32) ** 2, np.float32(0))])
    n = len(c)
    w1 = 2 * n                                   # columns 2i and 2i+1 both hold c[i]
    g = np.zeros(2 * w1, dtype=ck.PIXEL_DTYPE)
    g["opacity"] = 1
    g["depth"] = 1
    g["r"] = np.tile(np.repeat(c, 2), 2)
    out = ck.oracle_resolve(g, w1, 2)
    got = (out[0::2] >> 16) & 0xFF               # output pixel 2i averages four copies of c[i]
    want = np.floor(np.sqrt(c.astype(np.float64))).astype(np.uint32)
    assert (got == want).all()
    assert ((ck.ref_resolve(g, w1, 2)[0::2] >> 16) & 0xFF == want).all()
