"""ctypes access to the two CPU checkers used by the tests.

- the oracle (oracle/lib/libstoracle.so): the C restatement of the reference
  CPU renderer;
- the reference build (oracle/_ref/libslimref.so): the reference's own CPU
  renderer compiled from its headers. Present wherever oracle/Makefile's
  `ref` target ran (it needs the reference checkout).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from slimtracin_amd.scene import HIT_DTYPE, NODE_DTYPE, PIXEL_DTYPE, TRIANGLE_DTYPE, grid_size

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "lib", "libstoracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libslimref.so")


RAY_CLASSES = ("primary", "shadow0", "shadow_deep", "bounce")
TRAVERSAL_EVENTS = ("node_pairs", "root_tests", "triangle_tests", "primitive_visits")


class OracleStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "primary_rays", "shadow_rays", "bounce_rays", "node_pairs", "root_tests", "triangle_tests",
        "primitive_visits", "mesh_attr_fetches", "texel_fetches", "material_reads", "light_reads")] + [
        ("class_events", (C.c_uint64 * 4) * 4)]

    def as_dict(self):
        d = {n: int(getattr(self, n)) for n, _ in self._fields_ if n != "class_events"}
        for c, cname in enumerate(RAY_CLASSES):
            for e, ename in enumerate(TRAVERSAL_EVENTS):
                d[f"{cname}_{ename}"] = int(self.class_events[c][e])
        return d

    def traversal_bytes(self, ray_class):
        """Traversal bytes of one ray class (primary, shadow0, shadow_deep,
        bounce) at the reference's record sizes: node pairs 64, root AABB
        tests 24, triangle tests 60, primitive records 72."""
        ev = self.class_events[RAY_CLASSES.index(ray_class)]
        return 64 * ev[0] + 24 * ev[1] + 60 * ev[2] + 72 * ev[3]

    def algorithmic_bytes(self):
        """Bytes the reference algorithm reads, at the reference's record sizes."""
        return (64 * self.node_pairs + 24 * self.root_tests + 60 * self.triangle_tests + 72 * self.primitive_visits
                + 68 * self.mesh_attr_fetches + 12 * self.texel_fetches + 84 * self.material_reads
                + 44 * self.light_reads)


_libs = {}


def _load(path):
    if path not in _libs:
        _libs[path] = C.CDLL(path)
    return _libs[path]


def oracle_available():
    return os.path.exists(ORACLE_SO)


def ref_available():
    return os.path.exists(REF_SO)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _blob(blob: bytes):
    buf = np.frombuffer(blob, dtype=np.uint8)
    return _ptr(buf), C.c_size_t(len(blob)), buf


def oracle_render(blob, view, threads=1, hits=True):
    lib = _load(ORACLE_SO)
    n = grid_size(view)
    pixels = np.zeros(n, dtype=PIXEL_DTYPE)
    hrec = np.zeros(n, dtype=HIT_DTYPE) if hits else None
    st = OracleStats()
    bp, bs, keep = _blob(blob)
    rc = lib.oracle_render(bp, bs, _ptr(view), _ptr(pixels), _ptr(hrec) if hits else None, C.c_int(threads),
                           C.byref(st))
    assert rc == 0, "oracle_render failed"
    return pixels, hrec, st


def oracle_render_steps(blob, view, threads=1):
    """oracle_render plus each pixel's traversal events (node pairs +
    triangle tests + primitive visits over all its rays)."""
    lib = _load(ORACLE_SO)
    n = grid_size(view)
    pixels = np.zeros(n, dtype=PIXEL_DTYPE)
    hrec = np.zeros(n, dtype=HIT_DTYPE)
    steps = np.zeros(n, dtype=np.uint64)
    st = OracleStats()
    bp, bs, keep = _blob(blob)
    rc = lib.oracle_render_steps(bp, bs, _ptr(view), _ptr(pixels), _ptr(hrec), _ptr(steps), C.c_int(threads),
                                 C.byref(st))
    assert rc == 0, "oracle_render_steps failed"
    return pixels, hrec, steps, st


class OracleScene:
    """A scene blob loaded into the oracle once, rendered from many views.
    mesh_lbvh: meshes without a prebuilt BVH get the device's Morton LBVH
    (one triangle per leaf), as DeviceScene(mesh_lbvh=True) builds them; mesh_binned:
    the binned SAH tree, as DeviceScene(mesh_binned=True) builds them."""

    def __init__(self, blob, mesh_lbvh=False, mesh_binned=False):
        assert not (mesh_lbvh and mesh_binned)
        lib = _load(ORACLE_SO)
        lib.oracle_scene_open.restype = C.c_void_p
        self._keep = np.frombuffer(blob, dtype=np.uint8)
        flags = (1 if mesh_lbvh else 0) | (2 if mesh_binned else 0)
        self.handle = lib.oracle_scene_open(_ptr(self._keep), C.c_size_t(len(blob)), C.c_uint32(flags))
        assert self.handle, "oracle_scene_open failed"
        self._keep = None

    def render(self, view, threads=1, hits=True, pixels=True):
        """(pixels or None, hits or None, OracleStats, seconds of the render call)."""
        n = grid_size(view)
        px = np.zeros(n, dtype=PIXEL_DTYPE)
        hr = np.zeros(n, dtype=HIT_DTYPE) if hits else None
        st = OracleStats()
        secs = C.c_double()
        rc = _load(ORACLE_SO).oracle_scene_render(C.c_void_p(self.handle), _ptr(view), _ptr(px),
                                                  _ptr(hr) if hits else None, C.c_int(threads), C.byref(st),
                                                  C.byref(secs))
        assert rc == 0
        return (px if pixels else None), hr, st, secs.value

    def close(self):
        if self.handle:
            _load(ORACLE_SO).oracle_scene_close(C.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def oracle_render_timed(blob, view, threads, frames):
    lib = _load(ORACLE_SO)
    pixels = np.zeros(grid_size(view), dtype=PIXEL_DTYPE)
    secs = np.zeros(frames, dtype=np.float64)
    st = OracleStats()
    bp, bs, keep = _blob(blob)
    rc = lib.oracle_render_timed(bp, bs, _ptr(view), _ptr(pixels), C.c_int(threads), C.c_int(frames), _ptr(secs),
                                 C.byref(st))
    assert rc == 0
    return pixels, secs, st


def ref_render(blob, view, threads=1):
    lib = _load(REF_SO)
    pixels = np.zeros(grid_size(view), dtype=PIXEL_DTYPE)
    bp, bs, keep = _blob(blob)
    rc = lib.ref_render(bp, bs, _ptr(view), _ptr(pixels), C.c_int(threads))
    assert rc == 0, "ref_render failed"
    return pixels


def ref_render_timed(blob, view, threads, frames):
    lib = _load(REF_SO)
    pixels = np.zeros(grid_size(view), dtype=PIXEL_DTYPE)
    secs = np.zeros(frames, dtype=np.float64)
    bp, bs, keep = _blob(blob)
    rc = lib.ref_render_timed(bp, bs, _ptr(view), _ptr(pixels), C.c_int(threads), C.c_int(frames), _ptr(secs))
    assert rc == 0
    return pixels, secs


def ref_primary_hits(blob, view):
    lib = _load(REF_SO)
    hits = np.zeros(grid_size(view), dtype=HIT_DTYPE)
    bp, bs, keep = _blob(blob)
    assert lib.ref_primary_hits(bp, bs, _ptr(view), _ptr(hits)) == 0
    return hits


def _bvh(lib, fn, blob, max_nodes):
    nodes = np.zeros(max_nodes, dtype=NODE_DTYPE)
    leaf = np.zeros(max_nodes, dtype=np.uint32)
    cnt = C.c_uint32()
    h = C.c_uint32()
    bp, bs, keep = _blob(blob)
    assert getattr(lib, fn)(bp, bs, _ptr(nodes), _ptr(leaf), C.byref(cnt), C.byref(h)) == 0
    return nodes[:cnt.value], leaf, h.value


def oracle_scene_bvh(blob, prim_count):
    n, l, h = _bvh(_load(ORACLE_SO), "oracle_scene_bvh", blob, 2 * prim_count + 2)
    return n, l[:prim_count], h


def ref_scene_bvh(blob, prim_count):
    n, l, h = _bvh(_load(REF_SO), "ref_scene_bvh", blob, 2 * prim_count + 2)
    return n, l[:prim_count], h


def _mesh_bvh(lib, fn, blob, m, tri_count):
    nodes = np.zeros(2 * tri_count + 2, dtype=NODE_DTYPE)
    tris = np.zeros(tri_count, dtype=TRIANGLE_DTYPE)
    leaf = np.zeros(tri_count, dtype=np.uint32)
    cnt = C.c_uint32()
    h = C.c_uint32()
    bp, bs, keep = _blob(blob)
    assert getattr(lib, fn)(bp, bs, C.c_uint32(m), _ptr(nodes), _ptr(tris), _ptr(leaf), C.byref(cnt), C.byref(h)) == 0
    return nodes[:cnt.value], tris, leaf, h.value


def oracle_mesh_bvh(blob, m, tri_count):
    return _mesh_bvh(_load(ORACLE_SO), "oracle_mesh_bvh", blob, m, tri_count)


def ref_mesh_bvh(blob, m, tri_count):
    return _mesh_bvh(_load(REF_SO), "ref_mesh_bvh", blob, m, tri_count)


def _ssb(lib, fn, blob, view, prim_count):
    rects = np.zeros((prim_count, 4), dtype=np.int32)
    flags = np.zeros(prim_count, dtype=np.uint8)
    bp, bs, keep = _blob(blob)
    assert getattr(lib, fn)(bp, bs, _ptr(view), _ptr(rects), _ptr(flags)) == 0
    return rects, flags


def oracle_ssb(blob, view, prim_count):
    return _ssb(_load(ORACLE_SO), "oracle_ssb", blob, view, prim_count)


def ref_ssb(blob, view, prim_count):
    return _ssb(_load(REF_SO), "ref_ssb", blob, view, prim_count)


def _pick(lib, fn, blob, view, xy):
    from slimtracin_amd.scene import PICK_DTYPE
    xy = np.ascontiguousarray(xy, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros(len(xy), dtype=PICK_DTYPE)
    bp, bs, keep = _blob(blob)
    assert getattr(lib, fn)(bp, bs, _ptr(view), C.c_uint32(len(xy)), _ptr(xy), _ptr(out)) == 0
    return out


def oracle_pick(blob, view, xy):
    return _pick(_load(ORACLE_SO), "oracle_pick", blob, view, xy)


def ref_pick(blob, view, xy):
    return _pick(_load(REF_SO), "ref_pick", blob, view, xy)


def oracle_camera(view):
    out = np.zeros(17, dtype=np.float32)
    _load(ORACLE_SO).oracle_camera(_ptr(view), _ptr(out))
    return out


def ref_camera(view):
    out = np.zeros(17, dtype=np.float32)
    _load(REF_SO).ref_camera(_ptr(view), _ptr(out))
    return out


def oracle_sah_build(boxes, max_leaf):
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = len(boxes)
    nodes = np.zeros(2 * n + 2, dtype=NODE_DTYPE)
    leaf = np.zeros(n + 1, dtype=np.uint32)
    h = C.c_uint32()
    lib = _load(ORACLE_SO)
    lib.oracle_sah_build.restype = C.c_uint32
    cnt = lib.oracle_sah_build(_ptr(boxes), C.c_uint32(n), C.c_uint32(max_leaf), _ptr(nodes), _ptr(leaf), C.byref(h))
    return nodes[:cnt], leaf[:n], h.value


def ref_quat_axis(axis, amount):
    out = np.zeros(4, dtype=np.float32)
    a = np.asarray(axis, dtype=np.float32)
    _load(REF_SO).ref_quat_axis(_ptr(a), C.c_float(amount), _ptr(out))
    return out


def ref_quat_mul_norm(a, b):
    out = np.zeros(4, dtype=np.float32)
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    _load(REF_SO).ref_quat_mul_norm(_ptr(a), _ptr(b), _ptr(out))
    return out


def pixels_equal(a, b):
    """Bit-exact FloatPixel comparison (NaN-safe through the raw bytes)."""
    return a.tobytes() == b.tobytes()


def first_pixel_mismatches(a, b, k=5):
    av = a.view(np.uint8).reshape(len(a), -1)
    bv = b.view(np.uint8).reshape(len(b), -1)
    bad = np.nonzero((av != bv).any(axis=1))[0]
    return bad[:k], len(bad)


def oracle_lbvh_build(boxes):
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = len(boxes)
    nodes = np.zeros(2 * n + 2, dtype=NODE_DTYPE)
    leaf = np.zeros(n + 1, dtype=np.uint32)
    h = C.c_uint32()
    lib = _load(ORACLE_SO)
    lib.oracle_lbvh_build.restype = C.c_uint32
    cnt = lib.oracle_lbvh_build(_ptr(boxes), C.c_uint32(n), _ptr(nodes), _ptr(leaf), C.byref(h))
    return nodes[:cnt], leaf[:n], h.value


def oracle_binned_build(boxes, max_leaf):
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = len(boxes)
    nodes = np.zeros(2 * n + 2, dtype=NODE_DTYPE)
    leaf = np.zeros(n + 1, dtype=np.uint32)
    h = C.c_uint32()
    lib = _load(ORACLE_SO)
    lib.oracle_binned_build.restype = C.c_uint32
    cnt = lib.oracle_binned_build(_ptr(boxes), C.c_uint32(n), C.c_uint32(max_leaf), _ptr(nodes), _ptr(leaf),
                                  C.byref(h))
    return nodes[:cnt], leaf[:n], h.value


def sah_cost(nodes):
    """SAH cost of a reference-layout BVH: sum over reachable nodes of the
    node's half area over the root's, times 1 for an inner node and times its
    box count for a leaf."""
    root = nodes[0]
    def area(nd):
        e = (nd["max"] - nd["min"]).astype(np.float64)
        return e[0] * e[1] + e[1] * e[2] + e[2] * e[0]
    a0 = area(root)
    cost = 0.0
    for o in reachable_nodes(nodes):
        nd = nodes[o]
        cost += area(nd) / a0 * (int(nd["child_count"]) if nd["child_count"] else 1)
    return cost


def oracle_prim_boxes(blob, prim_count):
    out = np.zeros((prim_count, 6), dtype=np.float32)
    bp, bs, keep = _blob(blob)
    assert _load(ORACLE_SO).oracle_prim_boxes(bp, bs, _ptr(out)) == 0
    return out


def oracle_resolve(grid, grid_w, grid_h):
    out = np.zeros((grid_h - 1) * (grid_w - 1), dtype=np.uint32)
    g = np.ascontiguousarray(grid, dtype=PIXEL_DTYPE)
    _load(ORACLE_SO).oracle_resolve(_ptr(g), C.c_uint32(grid_w), C.c_uint32(grid_h), _ptr(out))
    return out


def ref_resolve(grid, grid_w, grid_h):
    out = np.zeros((grid_h - 1) * (grid_w - 1), dtype=np.uint32)
    g = np.ascontiguousarray(grid, dtype=PIXEL_DTYPE)
    assert _load(REF_SO).ref_resolve(_ptr(g), C.c_uint32(grid_w), C.c_uint32(grid_h), _ptr(out)) == 0
    return out


def resolve_test_grid(rng, w1, h1):
    """A traced grid exercising the resolve's edges: background corners
    (depth inf), channels at exact squares and one ulp either side, above
    255^2, zero, negative zero, fractional opacities."""
    n = w1 * h1
    g = np.zeros(n, dtype=PIXEL_DTYPE)
    k = rng.integers(0, 256, (n, 3)).astype(np.float32) ** 2
    bump = rng.integers(-1, 2, (n, 3))
    vals = np.where(bump > 0, np.nextafter(k, np.float32(np.inf)), np.where(bump < 0, np.nextafter(k, np.float32(-1)), k))
    vals = vals.astype(np.float32)
    big = rng.random((n, 3)) < 0.05
    vals[big] = np.float32(70000.0)
    vals[rng.random((n, 3)) < 0.03] = np.float32(-0.0)
    g["r"], g["g"], g["b"] = vals[:, 0], vals[:, 1], vals[:, 2]
    g["opacity"] = np.where(rng.random(n) < 0.8, np.float32(1.0), rng.random(n).astype(np.float32))
    g["depth"] = np.where(rng.random(n) < 0.3, np.inf, rng.random(n) * 10)
    return g


def oracle_mesh_lbvh(blob, m, tri_count):
    """Morton LBVH (one triangle per leaf) of mesh m and its triangles in leaf order."""
    return _mesh_bvh(_load(ORACLE_SO), "oracle_mesh_lbvh", blob, m, tri_count)


def oracle_mesh_binned(blob, m, tri_count):
    """Binned SAH tree (one triangle per leaf) of mesh m and its triangles in leaf order."""
    return _mesh_bvh(_load(ORACLE_SO), "oracle_mesh_binned", blob, m, tri_count)


def reachable_nodes(nodes):
    """Slots reachable from the root (node 0) of a reference-layout BVH."""
    seen, stack = [], [0]
    while stack:
        o = stack.pop()
        seen.append(o)
        if nodes["child_count"][o] == 0:
            stack += [int(nodes["first_child_id"][o]), int(nodes["first_child_id"][o]) + 1]
    return np.array(sorted(seen), dtype=np.int64)
