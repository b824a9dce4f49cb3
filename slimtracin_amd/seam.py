"""ctypes binding of the reference seam (include/slimtracin_seam.h).

One method per function of the reference's device mirror (SlimTracin
core/init.h:36-126) and renderSceneOnGPU (render/raytracer.h:258): a host
that owns its scene uploads arrays when they change and asks for frames with
the viewport's projection plane and camera.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import HIT_DTYPE, PIXEL_DTYPE, RayCounts, _check, _ptr, lib
from .scene import NODE_DTYPE, PRIM_DTYPE, TRIANGLE_DTYPE


class SeamSizes(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("primitives", "materials", "lights", "meshes", "textures")]


class SeamMesh(C.Structure):
    _fields_ = [("aabb", C.c_float * 6), ("triangle_count", C.c_uint32), ("normals_count", C.c_uint32),
                ("uvs_count", C.c_uint32), ("node_count", C.c_uint32), ("height", C.c_uint32),
                ("nodes", C.c_void_p), ("triangles", C.c_void_p), ("leaf_ids", C.c_void_p)]


class SeamMip(C.Structure):
    _fields_ = [("width", C.c_uint16), ("height", C.c_uint16), ("texel_quads", C.c_void_p)]


class SeamTexture(C.Structure):
    _fields_ = [("width", C.c_uint16), ("height", C.c_uint16), ("mip_count", C.c_uint8), ("wrap", C.c_uint8),
                ("mipmap", C.c_uint8), ("pad", C.c_uint8), ("mips", C.POINTER(SeamMip))]


class SeamFrame(C.Structure):
    _fields_ = [("start", C.c_float * 3), ("right", C.c_float * 3), ("down", C.c_float * 3),
                ("cone_angle", C.c_float), ("camera_position", C.c_float * 3),
                ("camera_rotation_inverted", C.c_float * 4), ("ambient", C.c_float * 3),
                ("width", C.c_uint16), ("height", C.c_uint16), ("render_mode", C.c_uint32), ("depth", C.c_uint32)]


def _arr(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class SeamScene:
    """A scene fed through the reference's upload functions."""

    def __init__(self, primitives, materials, lights, meshes=0, textures=0):
        self.sizes = SeamSizes(primitives, materials, lights, meshes, textures)
        h = C.c_void_p()
        _check(lib().st_seam_allocate(C.byref(self.sizes), C.byref(h)))
        self.handle = h
        self._keep = []

    def close(self):
        if self.handle:
            lib().st_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- the seven upload functions -------------------------------------

    def upload_primitives(self, prims, screen_bounds):
        p = _arr(prims, PRIM_DTYPE)
        b = _arr(screen_bounds, np.int32)
        _check(lib().st_seam_upload_primitives(self.handle, _ptr(p), _ptr(b)))

    def upload_materials(self, material_records: bytes):
        m = np.frombuffer(material_records, dtype=np.uint8)
        _check(lib().st_seam_upload_materials(self.handle, _ptr(m)))

    def upload_lights(self, light_records: bytes):
        lr = np.frombuffer(light_records, dtype=np.uint8) if light_records else np.zeros(1, np.uint8)
        _check(lib().st_seam_upload_lights(self.handle, _ptr(lr)))

    def upload_scene(self, light_records: bytes, prims, screen_bounds, material_records: bytes, textures=None):
        """textures: list of (width, height, wrap, mipmap, [(w, h, quads uint8), ...])."""
        lr = np.frombuffer(light_records, dtype=np.uint8) if light_records else np.zeros(1, np.uint8)
        p = _arr(prims, PRIM_DTYPE)
        b = _arr(screen_bounds, np.int32)
        m = np.frombuffer(material_records, dtype=np.uint8)
        tex_arr = None
        if textures:
            tex_arr = (SeamTexture * len(textures))()
            for i, (w, h, wrap, mipmap, mips) in enumerate(textures):
                marr = (SeamMip * len(mips))()
                for k, (mw, mh, quads) in enumerate(mips):
                    q = _arr(quads, np.uint8).reshape(-1)
                    self._keep.append(q)
                    marr[k] = SeamMip(mw, mh, q.ctypes.data)
                self._keep.append(marr)
                tex_arr[i] = SeamTexture(w, h, len(mips), wrap, mipmap, 0, C.cast(marr, C.POINTER(SeamMip)))
        _check(lib().st_seam_upload_scene(self.handle, _ptr(lr), _ptr(p), _ptr(b), _ptr(m),
                                          C.cast(tex_arr, C.c_void_p) if tex_arr is not None else None))
        self._keep = []

    def upload_scene_bvh(self, nodes, leaf_ids, height):
        n = _arr(nodes, NODE_DTYPE)
        leaf = _arr(leaf_ids, np.uint32)
        _check(lib().st_seam_upload_scene_bvh(self.handle, _ptr(n), C.c_uint32(len(n)), C.c_uint32(height),
                                              _ptr(leaf)))

    def upload_mesh_bvhs(self, meshes):
        """meshes: list of dicts aabb, nodes, triangles, height, normals_count, uvs_count."""
        arr = (SeamMesh * max(1, len(meshes)))()
        keep = []
        for i, m in enumerate(meshes):
            nodes = _arr(m["nodes"], NODE_DTYPE)
            tris = _arr(m["triangles"], TRIANGLE_DTYPE)
            keep += [nodes, tris]
            arr[i] = SeamMesh((C.c_float * 6)(*[float(v) for v in m["aabb"]]), len(tris), m["normals_count"],
                              m["uvs_count"], len(nodes), int(m["height"]), nodes.ctypes.data, tris.ctypes.data,
                              None)
        _check(lib().st_seam_upload_mesh_bvhs(self.handle, C.cast(arr, C.c_void_p)))

    # ---- renderSceneOnGPU ------------------------------------------------

    def render(self, frame: SeamFrame, hits=True):
        n = int(frame.width) * int(frame.height)
        pixels = np.zeros(n, dtype=PIXEL_DTYPE)
        hrec = np.zeros(n, dtype=HIT_DTYPE) if hits else None
        counts = RayCounts()
        _check(lib().st_seam_render(self.handle, C.byref(frame), _ptr(pixels), _ptr(hrec) if hits else None,
                                    C.byref(counts)))
        return pixels, hrec, counts


def frame_from_camera(camera17, position, cone_angle, ambient, grid_w, grid_h, mode, depth) -> SeamFrame:
    """SeamFrame from a projection plane + rotation in the 17-float layout
    (start, right, down, rotation_inverted, rotation) and the traced grid."""
    c = np.asarray(camera17, dtype=np.float32)
    f = SeamFrame()
    f.start[:] = [float(v) for v in c[0:3]]
    f.right[:] = [float(v) for v in c[3:6]]
    f.down[:] = [float(v) for v in c[6:9]]
    f.camera_rotation_inverted[:] = [float(v) for v in c[9:13]]
    f.camera_position[:] = [float(v) for v in np.asarray(position, dtype=np.float32)]
    f.cone_angle = float(np.float32(cone_angle))
    f.ambient[:] = [float(v) for v in np.asarray(ambient, dtype=np.float32)]
    f.width, f.height = grid_w, grid_h
    f.render_mode, f.depth = mode, depth
    return f
