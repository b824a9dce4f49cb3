"""MI355X-native tracing path of SlimTracin: Python binding of libslimtracin.so.

The library (slimtracin_amd/lib/libslimtracin.so, built by
`make -C slimtracin_amd/csrc`) is the product: host C + gfx950 HIP kernels
behind a C ABI (include/slimtracin.h). This module only marshals arguments.
It raises if the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .scene import HIT_DTYPE, NODE_DTYPE, PIXEL_DTYPE, PRIM_DTYPE, VIEW_DTYPE, grid_size  # noqa: F401

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libslimtracin.so")

ST_BATCH_PERSISTENT = 1
ST_BATCH_STAGE_TIMES = 2
ST_BATCH_GRAPH = 4
ST_REBUILD_SAH = 0
ST_REBUILD_LBVH = 1


class SlimTracinError(RuntimeError):
    pass


class RayCounts(C.Structure):
    _fields_ = [("primary", C.c_uint64), ("shadow", C.c_uint64), ("bounce", C.c_uint64)]

    def total(self):
        return int(self.primary + self.shadow + self.bounce)


class SceneInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("prim_count", "mesh_count", "triangle_count", "texture_count",
                                          "scene_node_count", "scene_height", "mesh_node_count",
                                          "max_mesh_height")] + [("device_bytes", C.c_uint64)]


_lib = None


def lib():
    """Load the HIP library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SlimTracinError(f"{LIB_PATH} is missing: build it with `make -C slimtracin_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        L.st_last_error.restype = C.c_char_p
        for name in ("st_gen_implicit", "st_gen_pbr", "st_gen_mesh_scene", "st_gen_instanced"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise SlimTracinError(f"status {rc}: {lib().st_last_error().decode(errors='replace')}")


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def init(device: int = 0):
    _check(lib().st_init(C.c_int(device)))


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("device", C.c_int32),
                ("compute_units", C.c_uint32), ("hbm_bytes", C.c_uint64)]


def device_info() -> dict:
    """The GPU the library runs on (st_device_info): name, architecture,
    device index, compute units, HBM bytes."""
    d = DeviceInfo()
    _check(lib().st_device_info(C.byref(d)))
    return {"name": d.name.decode(errors="replace"), "arch": d.arch.decode(errors="replace"), "device": d.device,
            "compute_units": d.compute_units, "hbm_bytes": d.hbm_bytes}


def _take_blob(rc, out, size):
    _check(rc)
    try:
        return C.string_at(out, size.value)
    finally:
        lib().st_free(out)


def gen_implicit(seed: int) -> bytes:
    out, size = C.c_void_p(), C.c_size_t()
    return _take_blob(lib().st_gen_implicit(C.c_uint64(seed), C.byref(out), C.byref(size)), out, size)


def gen_pbr(seed: int, k: int) -> bytes:
    out, size = C.c_void_p(), C.c_size_t()
    return _take_blob(lib().st_gen_pbr(C.c_uint64(seed), C.c_uint32(k), C.byref(out), C.byref(size)), out, size)


def gen_mesh_scene(seed: int, triangles: int) -> bytes:
    out, size = C.c_void_p(), C.c_size_t()
    return _take_blob(lib().st_gen_mesh_scene(C.c_uint64(seed), C.c_uint32(triangles), C.byref(out), C.byref(size)),
                      out, size)


ST_SCENE_MESH_LBVH = 1
ST_SCENE_MESH_BINNED = 2


def gen_instanced(seed: int, instances: int, mesh_triangles, tex_size: int) -> bytes:
    out, size = C.c_void_p(), C.c_size_t()
    tri = (C.c_uint32 * len(mesh_triangles))(*mesh_triangles)
    return _take_blob(lib().st_gen_instanced(C.c_uint64(seed), C.c_uint32(instances), tri,
                                             C.c_uint32(len(mesh_triangles)), C.c_uint32(tex_size),
                                             C.byref(out), C.byref(size)), out, size)


def scene_flags(mesh_lbvh: bool = False, mesh_binned: bool = False) -> int:
    """st_scene_create_ex flags of a mesh builder choice (at most one)."""
    if mesh_lbvh and mesh_binned:
        raise ValueError("mesh_lbvh and mesh_binned together")
    return (ST_SCENE_MESH_LBVH if mesh_lbvh else 0) | (ST_SCENE_MESH_BINNED if mesh_binned else 0)


class DeviceScene:
    """A scene resident in HBM (st_scene_create_ex). Meshes without a
    prebuilt BVH get the reference's host SAH build, or with mesh_lbvh the
    device Morton LBVH, or with mesh_binned the host binned SAH builder."""

    def __init__(self, blob: bytes, mesh_lbvh: bool = False, mesh_binned: bool = False):
        flags = scene_flags(mesh_lbvh, mesh_binned)
        self._blob = np.frombuffer(blob, dtype=np.uint8)
        h = C.c_void_p()
        _check(lib().st_scene_create_ex(_ptr(self._blob), C.c_size_t(len(blob)), C.c_uint32(flags), C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().st_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> SceneInfo:
        inf = SceneInfo()
        _check(lib().st_scene_info(self.handle, C.byref(inf)))
        return inf

    def bvh(self):
        inf = self.info()
        nodes = np.zeros(2 * inf.prim_count + 2, dtype=NODE_DTYPE)
        leaf = np.zeros(inf.prim_count + 1, dtype=np.uint32)
        cnt, h = C.c_uint32(), C.c_uint32()
        _check(lib().st_scene_bvh(self.handle, _ptr(nodes), _ptr(leaf), C.byref(cnt), C.byref(h)))
        return nodes[:cnt.value], leaf[:inf.prim_count], h.value

    def mesh_bvh(self, m: int, triangles: int):
        """(node slots, Triangle records in leaf order, height) of mesh m as resident."""
        from .scene import TRIANGLE_DTYPE
        cap = max(2 * triangles, 2)
        nodes = np.zeros(cap, dtype=NODE_DTYPE)
        tris = np.zeros(triangles, dtype=TRIANGLE_DTYPE)
        cnt, h = C.c_uint32(), C.c_uint32()
        _check(lib().st_scene_mesh_bvh(self.handle, C.c_uint32(m), _ptr(nodes), C.c_uint32(cap), _ptr(tris),
                                       C.byref(cnt), C.byref(h)))
        return nodes[:cnt.value], tris, h.value

    def set_prims(self, prims: np.ndarray, first: int = 0):
        """Replace primitive records [first, first+len(prims)) (st_scene_set_prims)."""
        prims = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
        _check(lib().st_scene_set_prims(self.handle, _ptr(prims), C.c_uint32(first), C.c_uint32(len(prims))))

    def rebuild(self, mode: int = ST_REBUILD_LBVH) -> float:
        """Rebuild the scene BVH (st_scene_rebuild); returns the build time in ms."""
        ms = C.c_float()
        _check(lib().st_scene_rebuild(self.handle, C.c_uint32(mode), C.byref(ms)))
        return ms.value


class _Pinned:
    """Owner of one st_host_alloc block."""

    def __init__(self, nbytes):
        p = C.c_void_p()
        _check(lib().st_host_alloc(C.c_size_t(nbytes), C.byref(p)))
        self.ptr = p.value

    def __del__(self):
        try:
            lib().st_host_free(C.c_void_p(self.ptr))
        except Exception:
            pass


def pinned_empty(n: int, dtype) -> np.ndarray:
    """Uninitialised numpy array in page-locked host memory (st_host_alloc)."""
    dtype = np.dtype(dtype)
    owner = _Pinned(max(1, n * dtype.itemsize))
    buf = (C.c_char * max(1, n * dtype.itemsize)).from_address(owner.ptr)
    buf._owner = owner
    return np.frombuffer(buf, dtype=dtype, count=n)


def pinned_split(sizes, dtype) -> list:
    """Arrays of the given lengths carved from one page-locked block."""
    dtype = np.dtype(dtype)
    offs, total = [], 0
    for n in sizes:
        offs.append(total)
        total += (n * dtype.itemsize + 255) // 256 * 256
    block = pinned_empty(max(total, 1), np.uint8)
    return [block[o:o + n * dtype.itemsize].view(dtype) for o, n in zip(offs, sizes)]


def render(scene: DeviceScene, view: np.ndarray, hits: bool = True, out: np.ndarray | None = None):
    """Plain per-request launch with blocking download (renderSceneOnGPU);
    out: a preallocated FloatPixel array to download into."""
    n = grid_size(view)
    pixels = out if out is not None else np.zeros(n, dtype=PIXEL_DTYPE)
    assert len(pixels) >= n and pixels.dtype == PIXEL_DTYPE
    hrec = np.zeros(n, dtype=HIT_DTYPE) if hits else None
    counts = RayCounts()
    _check(lib().st_render(scene.handle, _ptr(view), _ptr(pixels), _ptr(hrec) if hits else None, C.byref(counts)))
    return pixels, hrec, counts


def render_noalloc(scene: DeviceScene, view: np.ndarray, hits: bool = True, out: np.ndarray | None = None):
    """st_render_noalloc: the plain launch with every device buffer taken from
    the library's fixed scratch array (no allocation in the call)."""
    n = grid_size(view)
    pixels = out if out is not None else np.zeros(n, dtype=PIXEL_DTYPE)
    assert len(pixels) >= n and pixels.dtype == PIXEL_DTYPE
    hrec = np.zeros(n, dtype=HIT_DTYPE) if hits else None
    counts = RayCounts()
    _check(lib().st_render_noalloc(scene.handle, _ptr(view), _ptr(pixels), _ptr(hrec) if hits else None,
                                   C.byref(counts)))
    return pixels, hrec, counts


def render_batch(scenes, views: np.ndarray, hits: bool = False, persistent: bool = True, outs=None,
                 graph: bool = False):
    """Fuse len(views) requests into one launch; returns per-request outputs
    (outs: preallocated FloatPixel arrays, e.g. from pinned_split). graph:
    replay a captured HIP graph of the launch while its shape repeats."""
    n = len(views)
    handles = (C.c_void_p * n)(*[s.handle for s in scenes])
    pix = outs if outs is not None else [np.zeros(grid_size(views[i:i + 1]), dtype=PIXEL_DTYPE) for i in range(n)]
    hr = [np.zeros(grid_size(views[i:i + 1]), dtype=HIT_DTYPE) for i in range(n)] if hits else None
    pp = (C.c_void_p * n)(*[p.ctypes.data for p in pix])
    hp = (C.c_void_p * n)(*[h.ctypes.data for h in hr]) if hits else None
    counts = (RayCounts * n)()
    views = np.ascontiguousarray(views, dtype=VIEW_DTYPE)
    flags = (ST_BATCH_PERSISTENT if persistent else 0) | (ST_BATCH_GRAPH if graph else 0)
    _check(lib().st_render_batch(handles, _ptr(views), C.c_uint32(n), pp, hp, counts, C.c_uint32(flags)))
    return pix, hr, list(counts)


def device_ssb(scene, view: np.ndarray):
    """The SSB pass on the device (st_ssb): (rects (P, 4) int32 as computeSSB
    leaves them, flags (P,) uint8)."""
    P = scene.info().prim_count
    rects = np.zeros((P, 4), dtype=np.int32)
    flags = np.zeros(P, dtype=np.uint8)
    view = np.ascontiguousarray(view, dtype=VIEW_DTYPE)
    _check(lib().st_ssb(scene.handle, _ptr(view), _ptr(rects), _ptr(flags)))
    return rects, flags


def pick(scene, view: np.ndarray, xy) -> np.ndarray:
    """Pick at window points (an (n, 2) array of x, y): per point the closest
    primitive or light glow sphere, its world hit point and view-space
    distance (st_pick; scene.PICK_DTYPE records, object_type 0 = nothing)."""
    from .scene import PICK_DTYPE
    pts = np.ascontiguousarray(xy, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros(len(pts), dtype=PICK_DTYPE)
    view = np.ascontiguousarray(view, dtype=VIEW_DTYPE)
    _check(lib().st_pick(scene.handle, _ptr(view), C.c_uint32(len(pts)), _ptr(pts), _ptr(out)))
    return out


def debug_redo(abandon_steps: int = 0, coop_min_tris: int = 0):
    """st_debug_redo: the main stages' give-up step count and the redo stage's
    cooperative-traversal mesh size (0: the library defaults)."""
    _check(lib().st_debug_redo(C.c_uint32(abandon_steps), C.c_uint32(coop_min_tris)))


def last_redo_count() -> int:
    """Pixels the redo stage traced again in the last launch (main context)."""
    c = C.c_uint32()
    _check(lib().st_last_redo_count(C.byref(c)))
    return c.value


def graph_stats():
    """(graphs captured, graph launches) so far in this process."""
    c, l = C.c_uint64(), C.c_uint64()
    _check(lib().st_graph_stats(C.byref(c), C.byref(l)))
    return c.value, l.value


def render_timed(scenes, views: np.ndarray, persistent: bool = True, stages: bool = False, graph: bool = False):
    """One launch without downloads; returns (render ms from HIP events, counts)
    or, with stages=True, (ms, counts, per-stage ms in launch order: primary,
    then shadow/shade/bounce per depth-loop level)."""
    n = len(views)
    handles = (C.c_void_p * n)(*[s.handle for s in scenes])
    counts = (RayCounts * n)()
    ms = C.c_float()
    views = np.ascontiguousarray(views, dtype=VIEW_DTYPE)
    flags = ((ST_BATCH_PERSISTENT if persistent else 0) | (ST_BATCH_STAGE_TIMES if stages else 0)
             | (ST_BATCH_GRAPH if graph else 0))
    _check(lib().st_render_timed(handles, _ptr(views), C.c_uint32(n), C.c_uint32(flags), C.byref(ms), counts))
    if not stages:
        return ms.value, list(counts)
    buf = (C.c_float * 64)()
    cnt = C.c_uint32()
    _check(lib().st_last_stage_ms(buf, C.c_uint32(64), C.byref(cnt)))
    return ms.value, list(counts), [float(buf[i]) for i in range(min(cnt.value, 64))]


def stage_names(levels: int):
    """Names of the stage launches of a frame whose Beauty depth loop has
    `levels` levels (the order render_timed(stages=True) reports). "primary"
    includes the first level's shading entry, shadow rays and shading,
    "bounce" every bounce ray and every further level, and "redo" the pixels
    whose traversal gave up in those two, traced again by whole waves."""
    return ["primary"] + (["bounce"] if levels > 1 else []) + ["redo"]


def render_batch_bgra8(scenes, views: np.ndarray, persistent: bool = True, outs=None):
    """Fuse the requests into one launch and resolve each frame on the device
    to its width x height BGRA8 window image (uint32 per pixel); outs:
    preallocated arrays (e.g. from pinned_split)."""
    n = len(views)
    handles = (C.c_void_p * n)(*[s.handle for s in scenes])
    if outs is None:
        outs = [np.zeros(int(views["width"][i]) * int(views["height"][i]), dtype=np.uint32) for i in range(n)]
    op = (C.c_void_p * n)(*[o.ctypes.data for o in outs])
    counts = (RayCounts * n)()
    views = np.ascontiguousarray(views, dtype=VIEW_DTYPE)
    _check(lib().st_render_batch_bgra8(handles, _ptr(views), C.c_uint32(n), op, counts,
                                       C.c_uint32(ST_BATCH_PERSISTENT if persistent else 0)))
    return outs, list(counts)


def resolve_bgra8(grid: np.ndarray, grid_w: int, grid_h: int) -> np.ndarray:
    """Device display resolve of a traced FloatPixel grid (st_resolve_bgra8)."""
    g = np.ascontiguousarray(grid, dtype=PIXEL_DTYPE)
    out = np.zeros((grid_w - 1) * (grid_h - 1), dtype=np.uint32)
    _check(lib().st_resolve_bgra8(_ptr(g), C.c_uint32(grid_w), C.c_uint32(grid_h), _ptr(out)))
    return out


class Tuning(C.Structure):
    _fields_ = [("tile_shift", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("persistent", C.c_uint32),
                ("cost_order", C.c_uint32), ("ray_sort", C.c_uint32), ("stack_entries", C.c_uint32),
                ("primary_waves", C.c_uint32), ("bounce_order", C.c_uint32)]

    def as_dict(self):
        return {"tile": f"{1 << self.tile_shift}x{64 >> self.tile_shift}", "blocks_per_cu": self.blocks_per_cu,
                "persistent": bool(self.persistent), "xcd_claim": self.persistent == 2,
                "slowest_first": self.cost_order == 1, "super_tiles": self.cost_order == 2,
                "ray_sort": bool(self.ray_sort), "stack_entries": self.stack_entries or "whole",
                "primary_waves_per_simd": self.primary_waves or 3,
                "bounce_first": "others" if self.bounce_order else "refracting"}


def get_tuning() -> Tuning:
    t = Tuning()
    _check(lib().st_get_tuning(C.byref(t)))
    return t


def set_tuning(tile_shift: int = 3, blocks_per_cu: int = 4, persistent: bool = True, slowest_first: bool = False,
               ray_sort: bool = False, stack_entries: int = 16, primary_waves: int = 4, xcd_claim: bool = True,
               super_tiles: bool = False, refracting_first: bool = False):
    """Launch shape (StTuning); the defaults are the library's. xcd_claim:
    a persistent launch's tiles claimed by each XCD from its own eighth of
    the claim order first; super_tiles (when not slowest_first): tiles
    claimed request by request in blocks of 8 x 8 tiles; refracting_first:
    the bounce stage starts with the refracting pixels' depth loops rather
    than the others'."""
    mode = (2 if xcd_claim else 1) if persistent else 0
    order = 1 if slowest_first else 2 if super_tiles else 0
    t = Tuning(tile_shift, blocks_per_cu, mode, order, 1 if ray_sort else 0, stack_entries, primary_waves,
               0 if refracting_first else 1)
    _check(lib().st_set_tuning(C.byref(t)))


def restore_tuning(t: Tuning):
    """Set a Tuning read earlier with get_tuning (every field)."""
    _check(lib().st_set_tuning(C.byref(t)))


def autotune(scenes, views: np.ndarray, reps: int = 3):
    """Pick the fastest launch shape for these requests (st_autotune);
    returns (Tuning, render-kernel ms)."""
    n = len(views)
    handles = (C.c_void_p * n)(*[s.handle for s in scenes])
    views = np.ascontiguousarray(views, dtype=VIEW_DTYPE)
    best, ms = Tuning(), C.c_float()
    _check(lib().st_autotune(handles, _ptr(views), C.c_uint32(n), C.c_uint32(reps), C.byref(best), C.byref(ms)))
    return best, ms.value
