"""Scene description for the tracing path: builds the flat scene blob of
include/st_blob.h and the per-request StView.

The builder mirrors the reference's scene set-up API: defaults come from
initPrimitive / initMaterial (SlimTracin core/init.h:461-493) and the light
defaults of initScene (app.h:164-173); the quaternion helpers restate
getRotationAroundAxis / mulQuat / normQuat (math/quat.h:15-42,132-139) in
float32 so example scenes get bit-identical rotations.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

f32 = np.float32

PRIM_NONE, PRIM_MESH, PRIM_BOX, PRIM_TETRAHEDRON, PRIM_QUAD, PRIM_SPHERE = range(6)
MODE_NORMALS, MODE_BEAUTY, MODE_DEPTH, MODE_UVS = range(4)
BRDF_LAMBERT, BRDF_PHONG, BRDF_BLINN, BRDF_COOK_TORRANCE = range(4)

IS_VISIBLE, IS_TRANSLATED, IS_ROTATED, IS_SCALED = 1, 2, 4, 8
IS_SCALED_NON_UNIFORMLY, IS_SHADOWING, IS_TRANSPARENT = 16, 32, 64
ALL_FLAGS = 63
REFLECTIVE, REFRACTIVE, EMISSIVE = 1, 2, 4
ALBEDO_MAP, NORMAL_MAP = 1, 2
IOR_AIR, IOR_GLASS = f32(1.0003), f32(1.52)

BLOB_MAGIC = 0x43535453
BLOB_VERSION = 1
BLOB_HAS_SCENE_BVH = 1

VIEW_DTYPE = np.dtype([("position", "<f4", 3), ("yaw", "<f4"), ("pitch", "<f4"), ("roll", "<f4"),
                       ("focal_length", "<f4"), ("cone_angle", "<f4"), ("width", "<u2"), ("height", "<u2"),
                       ("render_mode", "<u4"), ("depth", "<u4"), ("reserved", "<u4")])
PIXEL_DTYPE = np.dtype([("r", "<f4"), ("g", "<f4"), ("b", "<f4"), ("opacity", "<f4"), ("depth", "<f8")])
HIT_DTYPE = np.dtype([("prim_id", "<i4"), ("object_type", "<i4"), ("triangle_id", "<i4"), ("distance", "<f4")])
PICK_DTYPE = np.dtype([("object_type", "<i4"), ("object_id", "<u4"), ("position", "<f4", 3),
                       ("object_distance", "<f4")])
PRIM_LIGHT = 6    # pick result: a light's glow sphere
NODE_DTYPE = np.dtype([("min", "<f4", 3), ("max", "<f4", 3), ("first_child_id", "<u4"),
                       ("child_count", "<u2"), ("depth", "<u2")])
TRIANGLE_DTYPE = np.dtype([("world_to_tangent", "<f4", 9), ("position", "<f4", 3), ("normal", "<f4", 3),
                           ("vertex_normals", "<f4", 9), ("uvs", "<f4", 6),
                           ("area_of_parallelogram", "<f4"), ("area_of_uv", "<f4")])
PRIM_DTYPE = np.dtype([("position", "<f4", 3), ("scale", "<f4", 3), ("rotation", "<f4", 4), ("type", "<u4"),
                       ("mesh_id", "<u4"), ("material_id", "<u4"), ("flags", "<u4")])
BLOB_HEADER_BYTES = 56
assert VIEW_DTYPE.itemsize == 48 and PIXEL_DTYPE.itemsize == 24 and HIT_DTYPE.itemsize == 16
assert PRIM_DTYPE.itemsize == 56
assert NODE_DTYPE.itemsize == 32 and TRIANGLE_DTYPE.itemsize == 128


# ---- float32 rotation helpers (core/base.h:404-411, math/quat.h) ----------

def unit_circle(t):
    t = f32(t)
    t2 = t * t
    fac = f32(1) / (f32(1) + t2)
    return fac - fac * t2, fac * f32(2) * t


def quat_norm(q):
    x, y, z, w = (f32(c) for c in q)
    fac = f32(1) / np.sqrt(x * x + y * y + z * z + w * w)
    return (x * fac, y * fac, z * fac, w * fac)


def quat_axis(axis, amount):
    """getRotationAroundAxis: axis scaled by the sine term, amount = cosine term."""
    c, s = unit_circle(amount)
    ax = [f32(a) for a in axis]
    return quat_norm((ax[0] * s, ax[1] * s, ax[2] * s, c))


def quat_mul(a, b):
    ax, ay, az, aw = (f32(c) for c in a)
    bx, by, bz, bw = (f32(c) for c in b)
    w = aw * bw - ax * bx - ay * by - az * bz
    x = aw * bx + ax * bw + ay * bz - az * by
    y = aw * by - ax * bz + ay * bw + az * bx
    z = aw * bz + ax * by - ay * bx + az * bw
    return (x, y, z, w)


def quat_mul_norm(a, b):
    return quat_norm(quat_mul(a, b))


IDENTITY_QUAT = (f32(0), f32(0), f32(0), f32(1))


@dataclass
class Primitive:
    type: int = PRIM_NONE
    position: tuple = (0.0, 0.0, 0.0)
    scale: tuple = (1.0, 1.0, 1.0)
    rotation: tuple = IDENTITY_QUAT
    mesh_id: int = 0
    material_id: int = 0
    flags: int = ALL_FLAGS


@dataclass
class Material:
    albedo: tuple = (1.0, 1.0, 1.0)
    reflectivity: tuple = (1.0, 1.0, 1.0)
    emission: tuple = (0.0, 0.0, 0.0)
    uv_repeat: tuple = (1.0, 1.0)
    metallic: float = 0.0
    roughness: float = 1.0
    n1_over_n2: float = 1.0
    n2_over_n1: float = 1.0
    brdf: int = BRDF_LAMBERT
    is_: int = 0
    use: int = NORMAL_MAP | ALBEDO_MAP
    texture_count: int = 0
    texture_ids: tuple = ()


@dataclass
class Light:
    position: tuple = (0.0, 0.0, 0.0)
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0


@dataclass
class Mesh:
    positions: np.ndarray                    # (V,3) float32
    indices: np.ndarray                      # (T,3) uint32
    normals: np.ndarray | None = None        # (N,3)
    normal_indices: np.ndarray | None = None  # (T,3)
    uvs: np.ndarray | None = None            # (U,2)
    uv_indices: np.ndarray | None = None     # (T,3)
    aabb: np.ndarray | None = None           # (6,) min xyz, max xyz
    prebuilt: tuple | None = None            # (nodes NODE_DTYPE, triangles TRIANGLE_DTYPE, leaf_ids u32, height)


@dataclass
class Texture:
    width: int
    height: int
    mips: list                               # list of (w, h, quads uint8 array of (h+1)*(w+1)*12)
    wrap: int = 1
    mipmap: int = 1


def mesh_aabb(positions: np.ndarray) -> np.ndarray:
    """Min/max over the vertices as obj2mesh does (src/obj2mesh.c:193-201)."""
    p = np.asarray(positions, dtype=f32)
    return np.concatenate([p.min(axis=0), p.max(axis=0)]).astype(f32)


@dataclass
class Scene:
    primitives: list = field(default_factory=list)
    materials: list = field(default_factory=list)
    lights: list = field(default_factory=list)
    meshes: list = field(default_factory=list)
    textures: list = field(default_factory=list)
    ambient: tuple = (0.0, 0.0, 0.0)
    scene_bvh: tuple | None = None           # (nodes NODE_DTYPE, leaf_ids u32, height)

    def to_blob(self) -> bytes:
        out = bytearray()
        flags = BLOB_HAS_SCENE_BVH if self.scene_bvh is not None else 0
        out += struct.pack("<7I3fI3I", BLOB_MAGIC, BLOB_VERSION, len(self.primitives), len(self.materials),
                           len(self.lights), len(self.meshes), len(self.textures), *map(float, self.ambient),
                           flags, 0, 0, 0)
        for p in self.primitives:
            out += np.asarray(list(p.position) + list(p.scale) + list(p.rotation), dtype=f32).tobytes()
            out += struct.pack("<4I", p.type, p.mesh_id, p.material_id, p.flags)
        for m in self.materials:
            vals = list(m.albedo) + list(m.reflectivity) + list(m.emission) + list(m.uv_repeat) + \
                [m.metallic, m.roughness, m.n1_over_n2, m.n2_over_n1]
            out += np.asarray(vals, dtype=f32).tobytes()
            tids = list(m.texture_ids) + [0] * (16 - len(m.texture_ids))
            out += struct.pack("<I4B16B", m.brdf, m.is_, m.use, m.texture_count, 0, *tids)
        for light in self.lights:
            out += np.asarray(list(light.position) + list(light.color) + [light.intensity], dtype=f32).tobytes()
        for mesh in self.meshes:
            pos = np.ascontiguousarray(mesh.positions, dtype=f32)
            idx = np.ascontiguousarray(mesh.indices, dtype=np.uint32)
            nrm = None if mesh.normals is None else np.ascontiguousarray(mesh.normals, dtype=f32)
            uvs = None if mesh.uvs is None else np.ascontiguousarray(mesh.uvs, dtype=f32)
            aabb = mesh.aabb if mesh.aabb is not None else mesh_aabb(pos)
            pre = mesh.prebuilt
            node_count = 0 if pre is None else len(pre[0])
            height = 0 if pre is None else int(pre[3])
            out += struct.pack("<8I", len(pos), len(idx), 0 if nrm is None else len(nrm),
                               0 if uvs is None else len(uvs), 0 if pre is None else 1, node_count, height, 0)
            out += np.asarray(aabb, dtype=f32).tobytes()
            out += pos.tobytes() + idx.tobytes()
            if nrm is not None:
                out += nrm.tobytes() + np.ascontiguousarray(mesh.normal_indices, dtype=np.uint32).tobytes()
            if uvs is not None:
                out += uvs.tobytes() + np.ascontiguousarray(mesh.uv_indices, dtype=np.uint32).tobytes()
            if pre is not None:
                out += np.ascontiguousarray(pre[0], dtype=NODE_DTYPE).tobytes()
                out += np.ascontiguousarray(pre[1], dtype=TRIANGLE_DTYPE).tobytes()
                out += np.ascontiguousarray(pre[2], dtype=np.uint32).tobytes()
        for tex in self.textures:
            out += struct.pack("<2H4B", tex.width, tex.height, len(tex.mips), tex.wrap, tex.mipmap, 0)
            for (w, h, quads) in tex.mips:
                q = np.ascontiguousarray(quads, dtype=np.uint8).reshape(-1)
                assert q.size == (w + 1) * (h + 1) * 12
                out += struct.pack("<2H", w, h) + q.tobytes()
        if self.scene_bvh is not None:
            nodes, leaf_ids, height = self.scene_bvh
            out += struct.pack("<2I", len(nodes), int(height))
            out += np.ascontiguousarray(nodes, dtype=NODE_DTYPE).tobytes()
            out += np.ascontiguousarray(leaf_ids, dtype=np.uint32).tobytes()
        return bytes(out)


def make_view(position=(0.0, 0.0, 0.0), yaw=0.0, pitch=0.0, roll=0.0, width=64, height=48,
              mode=MODE_BEAUTY, depth=2, focal_length=2.0, cone_angle=0.0) -> np.ndarray:
    """One StView record (camera defaults of initCamera, core/init.h:337-345)."""
    v = np.zeros(1, dtype=VIEW_DTYPE)
    v["position"] = position
    v["yaw"], v["pitch"], v["roll"] = yaw, pitch, roll
    v["focal_length"] = focal_length
    v["cone_angle"] = cone_angle
    v["width"], v["height"] = width, height
    v["render_mode"] = mode
    v["depth"] = depth
    return v


def grid_size(view) -> int:
    """Traced pixel count: (width+1)*(height+1) (QCAA, core/base.h:344-358)."""
    return (int(view["width"][0]) + 1) * (int(view["height"][0]) + 1)


# ---- blob edits (the primitive array follows the 56-byte header) ----------

def blob_prims(blob: bytes) -> np.ndarray:
    """Copy of the blob's StPrimRec array."""
    n = struct.unpack_from("<I", blob, 8)[0]
    return np.frombuffer(blob, dtype=PRIM_DTYPE, count=n, offset=BLOB_HEADER_BYTES).copy()


def blob_with_prims(blob: bytes, prims: np.ndarray) -> bytes:
    """The same scene with its primitive records replaced (same count)."""
    n = struct.unpack_from("<I", blob, 8)[0]
    prims = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    assert len(prims) == n
    end = BLOB_HEADER_BYTES + n * PRIM_DTYPE.itemsize
    return blob[:BLOB_HEADER_BYTES] + prims.tobytes() + blob[end:]


def blob_with_scene_bvh(blob: bytes, nodes: np.ndarray, leaf_ids: np.ndarray, height: int) -> bytes:
    """The same scene carrying a prebuilt scene BVH (e.g. one built by the
    device LBVH), used instead of the SAH build when the blob is loaded."""
    flags = struct.unpack_from("<I", blob, 40)[0]
    assert not flags & BLOB_HAS_SCENE_BVH, "blob already carries a scene BVH"
    head = bytearray(blob[:BLOB_HEADER_BYTES])
    struct.pack_into("<I", head, 40, flags | BLOB_HAS_SCENE_BVH)
    return (bytes(head) + blob[BLOB_HEADER_BYTES:] + struct.pack("<2I", len(nodes), int(height))
            + np.ascontiguousarray(nodes, dtype=NODE_DTYPE).tobytes()
            + np.ascontiguousarray(leaf_ids, dtype=np.uint32).tobytes())


def blob_records(blob: bytes) -> dict:
    """The blob's header counts, ambient light and the raw primitive,
    material (84 B) and light (28 B) record arrays."""
    magic, version, P, M, L, NM, NT = struct.unpack_from("<7I", blob, 0)
    ambient = struct.unpack_from("<3f", blob, 28)
    o = BLOB_HEADER_BYTES
    prims = np.frombuffer(blob, dtype=PRIM_DTYPE, count=P, offset=o).copy()
    o += P * PRIM_DTYPE.itemsize
    mats = bytes(blob[o:o + 84 * M])
    o += 84 * M
    lights = bytes(blob[o:o + 28 * L])
    return dict(prim_count=P, material_count=M, light_count=L, mesh_count=NM, texture_count=NT, ambient=ambient,
                prims=prims, materials=mats, lights=lights)


def blob_mesh_info(blob: bytes) -> list:
    """Per mesh: dict of the StMeshRec counts and the byte offset of its record."""
    magic, version, P, M, L, NM, NT = struct.unpack_from("<7I", blob, 0)
    o = BLOB_HEADER_BYTES + P * PRIM_DTYPE.itemsize + 84 * M + 28 * L
    out = []
    for _ in range(NM):
        V, T, N, U, pre, nc, h, _r = struct.unpack_from("<8I", blob, o)
        out.append(dict(offset=o, vertex_count=V, triangle_count=T, normals_count=N, uvs_count=U, prebuilt=pre,
                        node_count=nc, height=h))
        o += 56 + 12 * V + 12 * T
        if N:
            o += 12 * N + 12 * T
        if U:
            o += 8 * U + 12 * T
        if pre:
            o += 32 * nc + 128 * T + 4 * T
    return out


def blob_with_prebuilt_meshes(blob: bytes, builds: list) -> bytes:
    """The same scene with each mesh carrying a prebuilt BVH: builds[m] =
    (nodes, triangles, leaf_ids, height) or None to keep mesh m as it is."""
    info = blob_mesh_info(blob)
    out = bytearray()
    pos = 0
    for m, mi in enumerate(info):
        o = mi["offset"]
        out += blob[pos:o]
        T, V, N, U = mi["triangle_count"], mi["vertex_count"], mi["normals_count"], mi["uvs_count"]
        end = o + 56 + 12 * V + 12 * T + (12 * N + 12 * T if N else 0) + (8 * U + 12 * T if U else 0)
        body_end = end + (32 * mi["node_count"] + 132 * T if mi["prebuilt"] else 0)
        if builds[m] is None:
            out += blob[o:body_end]
        else:
            nodes, tris, leaf, h = builds[m]
            rec = bytearray(blob[o:o + 56])
            struct.pack_into("<3I", rec, 16, 1, len(nodes), int(h))
            out += rec + blob[o + 56:end]
            out += np.ascontiguousarray(nodes, dtype=NODE_DTYPE).tobytes()
            out += np.ascontiguousarray(tris, dtype=TRIANGLE_DTYPE).tobytes()
            out += np.ascontiguousarray(leaf, dtype=np.uint32).tobytes()
        pos = body_end
    out += blob[pos:]
    return bytes(out)
