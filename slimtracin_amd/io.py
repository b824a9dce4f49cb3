"""The reference's asset formats through the library's C readers/writers
(include/slimtracin_io.h): .mesh files (mesh, prebuilt BVH, triangle
records; scene/io.h:92-175), .texture files (texel-quad mip chains;
scene/io.h:34-57), and the obj2mesh / bmp2texture conversions
(src/obj2mesh.c, src/bmp2texture.c)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import SlimTracinError, _check, lib
from .scene import NODE_DTYPE, TRIANGLE_DTYPE, Mesh, Texture


class _MeshFile(C.Structure):
    _fields_ = [("aabb", C.c_float * 6)] + [(n, C.c_uint32) for n in (
        "vertex_count", "triangle_count", "edge_count", "uvs_count", "normals_count", "node_count", "height")] + [
        (n, C.c_void_p) for n in ("positions", "position_indices", "edges", "uvs", "uv_indices", "normals",
                                   "normal_indices", "triangles", "nodes", "leaf_ids")]


class _MipView(C.Structure):
    _fields_ = [("width", C.c_uint16), ("height", C.c_uint16), ("quads", C.c_void_p)]


class _TextureFile(C.Structure):
    _fields_ = [("width", C.c_uint16), ("height", C.c_uint16), ("mipmap", C.c_uint8), ("wrap", C.c_uint8),
                ("mip_count", C.c_uint8), ("mips", _MipView * 16)]


def _bytes(src) -> bytes:
    if isinstance(src, (bytes, bytearray, memoryview)):
        return bytes(src)
    with open(os.fspath(src), "rb") as f:
        return f.read()


def _take(buf, base, ptr, dtype, count, shape=None):
    if not ptr or count == 0:
        return None
    a = np.frombuffer(buf, dtype=dtype, count=count, offset=ptr - base).copy()
    return a.reshape(shape) if shape else a


def _owned(ptr, size) -> bytes:
    try:
        return C.string_at(ptr, size)
    finally:
        lib().st_free(C.c_void_p(ptr))


def parse_mesh(src) -> dict:
    """All arrays of a .mesh file (path or bytes), validated by the library."""
    data = _bytes(src)
    buf = C.create_string_buffer(data, len(data))
    base = C.addressof(buf)
    m = _MeshFile()
    _check(lib().st_mesh_file_parse(buf, C.c_size_t(len(data)), C.byref(m)))
    T, V = m.triangle_count, m.vertex_count
    return {
        "aabb": np.array(m.aabb[:], dtype=np.float32), "height": m.height,
        "positions": _take(buf, base, m.positions, "<f4", 3 * V, (V, 3)),
        "position_indices": _take(buf, base, m.position_indices, "<u4", 3 * T, (T, 3)),
        "edges": _take(buf, base, m.edges, "<u4", 2 * m.edge_count, (m.edge_count, 2)),
        "uvs": _take(buf, base, m.uvs, "<f4", 2 * m.uvs_count, (m.uvs_count, 2)),
        "uv_indices": _take(buf, base, m.uv_indices, "<u4", 3 * T, (T, 3)),
        "normals": _take(buf, base, m.normals, "<f4", 3 * m.normals_count, (m.normals_count, 3)),
        "normal_indices": _take(buf, base, m.normal_indices, "<u4", 3 * T, (T, 3)),
        "triangles": _take(buf, base, m.triangles, TRIANGLE_DTYPE, T),
        "nodes": _take(buf, base, m.nodes, NODE_DTYPE, m.node_count),
        "leaf_ids": _take(buf, base, m.leaf_ids, "<u4", T),
    }


def read_mesh(src) -> Mesh:
    """A .mesh file as a scene Mesh carrying its prebuilt BVH and records."""
    d = parse_mesh(src)
    return Mesh(d["positions"], d["position_indices"], d["normals"], d["normal_indices"], d["uvs"], d["uv_indices"],
                aabb=d["aabb"], prebuilt=(d["nodes"], d["triangles"], d["leaf_ids"], d["height"]))


def write_mesh(d: dict) -> bytes:
    """Serialise the arrays of parse_mesh's form to a .mesh image."""
    keep = []

    def ptr(a, dtype):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return a.ctypes.data

    T = len(d["position_indices"])
    m = _MeshFile()
    m.aabb[:] = [float(x) for x in d["aabb"]]
    m.vertex_count = len(d["positions"])
    m.triangle_count = T
    m.edge_count = 0 if d["edges"] is None else len(d["edges"])
    m.uvs_count = 0 if d["uvs"] is None else len(d["uvs"])
    m.normals_count = 0 if d["normals"] is None else len(d["normals"])
    m.node_count = len(d["nodes"])
    m.height = int(d["height"])
    m.positions = ptr(d["positions"], "<f4")
    m.position_indices = ptr(d["position_indices"], "<u4")
    m.edges = ptr(d["edges"], "<u4")
    m.uvs = ptr(d["uvs"], "<f4")
    m.uv_indices = ptr(d["uv_indices"], "<u4")
    m.normals = ptr(d["normals"], "<f4")
    m.normal_indices = ptr(d["normal_indices"], "<u4")
    m.triangles = ptr(d["triangles"], TRIANGLE_DTYPE)
    m.nodes = ptr(d["nodes"], NODE_DTYPE)
    m.leaf_ids = ptr(d["leaf_ids"], "<u4")
    out, size = C.c_void_p(), C.c_size_t()
    _check(lib().st_mesh_file_write(C.byref(m), C.byref(out), C.byref(size)))
    return _owned(out.value, size.value)


def obj_to_mesh(obj_text, invert_winding: bool = False) -> bytes:
    """obj2mesh: Wavefront OBJ text (str or bytes) to a .mesh image."""
    raw = obj_text.encode() if isinstance(obj_text, str) else bytes(obj_text)
    out, size = C.c_void_p(), C.c_size_t()
    _check(lib().st_obj_to_mesh(C.c_char_p(raw), C.c_size_t(len(raw)), C.c_int(int(invert_winding)),
                                C.byref(out), C.byref(size)))
    return _owned(out.value, size.value)


def read_texture(src) -> Texture:
    """A .texture file as a scene Texture (mips of texel quads)."""
    data = _bytes(src)
    buf = C.create_string_buffer(data, len(data))
    base = C.addressof(buf)
    t = _TextureFile()
    _check(lib().st_texture_file_parse(buf, C.c_size_t(len(data)), C.byref(t)))
    mips = []
    for i in range(t.mip_count):
        mv = t.mips[i]
        n = (mv.width + 1) * (mv.height + 1) * 12
        mips.append((mv.width, mv.height, _take(buf, base, mv.quads, np.uint8, n)))
    return Texture(t.width, t.height, mips, wrap=int(t.wrap), mipmap=int(t.mipmap))


def texture_from_texels(texels: np.ndarray, wrap: bool = True, mipmap: bool = True) -> bytes:
    """bmp2texture's conversion of a (height, width, 3 or 4) uint8 texel
    array (first three channels r, g, b) to a .texture image."""
    a = np.ascontiguousarray(texels, dtype=np.uint8)
    if a.ndim != 3 or a.shape[2] not in (3, 4):
        raise SlimTracinError("texels must be (height, width, 3 or 4) uint8")
    h, w, c = a.shape
    out, size = C.c_void_p(), C.c_size_t()
    _check(lib().st_texture_from_texels(a.ctypes.data_as(C.c_void_p), C.c_uint16(w), C.c_uint16(h), C.c_uint32(c),
                                        C.c_int(int(wrap)), C.c_int(int(mipmap)), C.byref(out), C.byref(size)))
    return _owned(out.value, size.value)


def bmp_to_texture(bmp, wrap: bool = True, mipmap: bool = True) -> bytes:
    """bmp2texture: an uncompressed 24/32-bit BMP (path or bytes) to a .texture image."""
    data = _bytes(bmp)
    out, size = C.c_void_p(), C.c_size_t()
    _check(lib().st_bmp_to_texture(C.c_char_p(data), C.c_size_t(len(data)), C.c_int(int(wrap)), C.c_int(int(mipmap)),
                                   C.byref(out), C.byref(size)))
    return _owned(out.value, size.value)
