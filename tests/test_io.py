"""Asset formats (include/slimtracin_io.h) on the host.

Pins: the reference's own .mesh file (src/examples/monkey.mesh, the only
asset it ships, kept as tests/golden/monkey.mesh) parsed field for field
and written back byte for byte; obj2mesh run on an OBJ carrying that mesh's
vertices, uvs, normals and faces reproduces the file byte for byte (bound,
edges, the reference SAH BVH, triangle records, leaf ids); texel quads and
mips of bmp2texture against their definitions (the four texels around each
grid corner, clamped or wrapped at the borders, with the tool's order of
border writes; each coarser texel the mean of a 2x2 block, truncated), and
the BMP path against the texel path. No .texture file ships with the
reference, so the texture side is pinned by these definitions only.
"""
from __future__ import annotations

import os
import struct

import numpy as np
import pytest

import scenes
import slimtracin_amd as st
from slimtracin_amd import io

MONKEY = os.path.join(os.path.dirname(__file__), "golden", "monkey.mesh")


def test_monkey_mesh_parse_and_write_back():
    raw = open(MONKEY, "rb").read()
    d = io.parse_mesh(raw)
    ref = scenes.read_reference_mesh(MONKEY)
    assert np.array_equal(d["positions"], ref.positions)
    assert np.array_equal(d["position_indices"], ref.indices)
    assert np.array_equal(d["normals"], ref.normals) and np.array_equal(d["normal_indices"], ref.normal_indices)
    assert np.array_equal(d["uvs"], ref.uvs) and np.array_equal(d["uv_indices"], ref.uv_indices)
    nodes, tris, leaf, height = ref.prebuilt
    assert d["nodes"].tobytes() == nodes.tobytes() and d["triangles"].tobytes() == tris.tobytes()
    assert np.array_equal(d["leaf_ids"], leaf) and d["height"] == height
    assert io.write_mesh(d) == raw
    m = io.read_mesh(MONKEY)
    assert m.prebuilt[0].tobytes() == nodes.tobytes()


def _obj_text(d):
    f9 = lambda x: repr(float(np.float32(x)))  # shortest text that round-trips the float32
    lines = ["# monkey.mesh written back as OBJ"]
    lines += [f"v {f9(x)} {f9(y)} {f9(z)}" for x, y, z in d["positions"]]
    lines += [f"vt {f9(u)} {f9(v)}" for u, v in d["uvs"]]
    lines += [f"vn {f9(x)} {f9(y)} {f9(z)}" for x, y, z in d["normals"]]
    for p, t, n in zip(d["position_indices"], d["uv_indices"], d["normal_indices"]):
        lines.append("f " + " ".join(f"{p[k] + 1}/{t[k] + 1}/{n[k] + 1}" for k in range(3)))
    return "\n".join(lines) + "\n"


def test_obj2mesh_reproduces_monkey_mesh():
    raw = open(MONKEY, "rb").read()
    d = io.parse_mesh(raw)
    out = io.obj_to_mesh(_obj_text(d))
    got = io.parse_mesh(out)
    for k in ("positions", "position_indices", "edges", "uvs", "uv_indices", "normals", "normal_indices",
              "leaf_ids", "aabb"):
        assert np.array_equal(got[k], d[k]), k
    assert got["nodes"].tobytes() == d["nodes"].tobytes()
    assert got["triangles"].tobytes() == d["triangles"].tobytes()
    assert out == raw


def test_obj2mesh_small_cases():
    # positions only, off-centre: the bound includes the origin, then recentres
    obj = "v 1 1 1\nv 3 1 1\nv 1 3 1\nv 1 1 3\nf 1 2 3\nf 1 3 4\nf 1 4 2\nf 2 4 3\n"
    d = io.parse_mesh(io.obj_to_mesh(obj))
    assert np.array_equal(d["aabb"], np.array([-1.5, -1.5, -1.5, 1.5, 1.5, 1.5], np.float32))
    assert np.array_equal(d["positions"][0], np.array([-0.5, -0.5, -0.5], np.float32))
    assert d["uvs"] is None and d["normals"] is None
    # 4 faces of a tetrahedron: 6 unique undirected edges, first-seen order
    assert d["edges"].tolist() == [[0, 1], [1, 2], [0, 2], [2, 3], [0, 3], [1, 3]]
    assert len(d["nodes"]) == 1 and d["nodes"]["child_count"][0] == 4
    inv = io.parse_mesh(io.obj_to_mesh(obj, invert_winding=True))
    assert np.array_equal(inv["position_indices"][:, [0, 2, 1]], d["position_indices"])
    with pytest.raises(st.SlimTracinError):
        io.obj_to_mesh("v 0 0 0\nvn 0 1 0\nf 1 1 1\n")      # normals without the v/t/n face form
    with pytest.raises(st.SlimTracinError):
        io.obj_to_mesh("v 0 0 0\nf 1 2 3\n")                 # index out of range


def test_mesh_parse_rejects_bad_files():
    raw = open(MONKEY, "rb").read()
    for bad in (raw[:40], raw[:-4], raw + b"\0\0\0\0"):
        with pytest.raises(st.SlimTracinError):
            io.parse_mesh(bad)


def _quads(q):
    return q.reshape(-1, 3, 4)   # [quad][R,G,B][TL,TR,BL,BR]


@pytest.mark.parametrize("wrap", [False, True])
def test_texture_quads_and_mips_by_definition(wrap):
    rng = np.random.default_rng(4)
    H, W = 32, 64
    img = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    t = io.read_texture(io.texture_from_texels(img, wrap=wrap, mipmap=True))
    assert (t.width, t.height, t.wrap, t.mipmap) == (W, H, int(wrap), 1)
    assert [(w, h) for w, h, _ in t.mips] == [(64, 32), (32, 16), (16, 8), (8, 4)]
    level = img.astype(np.float64)
    for k, (w, h, q) in enumerate(t.mips):
        if k:
            level = (level[0::2, 0::2] + level[0::2, 1::2] + level[1::2, 0::2] + level[1::2, 1::2]) * 0.25
        tex = np.floor(level).astype(np.uint8)      # exact: the float32 chain carries <= 8+2k bits
        qd = _quads(q).reshape(h + 1, w + 1, 3, 4)
        ix = np.arange(w + 1)
        iy = np.arange(h + 1)
        fix = (lambda i, n: np.mod(i, n)) if wrap else (lambda i, n: np.clip(i, 0, n - 1))
        for c, (dx, dy) in enumerate([(-1, -1), (0, -1), (-1, 0), (0, 0)]):   # TL, TR, BL, BR
            want = tex[fix(iy + dy, h)][:, fix(ix + dx, w)].copy()               # (h+1, w+1, 3)
            if wrap and c in (1, 3):
                # the tool writes the right border column's clamped corners
                # (TR, BR) after the wrapped ones from the left edge, so they
                # win except on the wrapped top (TR) / bottom (BR) row
                rows = np.arange(1, h + 1) if c == 1 else np.arange(0, h)
                want[rows, w] = tex[rows + dy, w - 1]
            assert np.array_equal(qd[:, :, :, c], want), (k, c)


def test_texture_single_level_and_constant():
    img = np.full((8, 8, 4), 77, np.uint8)
    t = io.read_texture(io.texture_from_texels(img, wrap=False, mipmap=False))
    assert len(t.mips) == 1 and (t.mips[0][2] == 77).all()
    t = io.read_texture(io.texture_from_texels(img, wrap=True, mipmap=True))
    assert [(w, h) for w, h, _ in t.mips] == [(8, 8), (4, 4)]
    assert all((q == 77).all() for _, _, q in t.mips)


def _bmp24(img_rows_bgr):
    h, w, _ = img_rows_bgr.shape
    data = img_rows_bgr.tobytes()
    hdr = struct.pack("<2sIHHI", b"BM", 54 + len(data), 0, 0, 54)
    info = struct.pack("<IiiHHIIiiII", 40, w, h, 1, 24, 0, len(data), 2835, 2835, 0, 0)
    return hdr + info + data


def test_bmp_path_matches_texel_path():
    rng = np.random.default_rng(9)
    stored = rng.integers(0, 256, (16, 32, 3), dtype=np.uint8)     # rows as stored, B G R
    a = io.bmp_to_texture(_bmp24(stored), wrap=True, mipmap=True)
    b = io.texture_from_texels(stored[:, :, ::-1], wrap=True, mipmap=True)
    assert a == b
    with pytest.raises(st.SlimTracinError):
        io.bmp_to_texture(b"not a bitmap")
