"""Ports of the reference examples' setupScene bodies as test scenes.

Each function returns (Scene, view kwargs). Values follow
src/examples/0N_*.c of the reference; float constants are computed with
float32 arithmetic where the reference computes them in C (e.g. 1.1f * 40.0f).
"""
from __future__ import annotations

import os

import numpy as np

from slimtracin_amd.scene import (
    ALBEDO_MAP, BRDF_BLINN, BRDF_COOK_TORRANCE, BRDF_PHONG, EMISSIVE, IOR_AIR, IOR_GLASS, IS_TRANSPARENT,
    NORMAL_MAP, PRIM_BOX, PRIM_MESH, PRIM_QUAD, PRIM_SPHERE, PRIM_TETRAHEDRON, REFLECTIVE, REFRACTIVE,
    Light, Material, Mesh, Primitive, Scene, Texture, f32, quat_axis, quat_mul_norm,
)

HALF_SQRT2 = f32(0.70710678118)
CAMERA = dict(position=(0.0, 7.0, -11.0), pitch=-0.2)


def _three_point_lights(scale=(1.1, 1.2, 0.9)):
    # 02_Geometry.c:137-151 (key/rim/fill)
    k, r, f = (f32(s) * f32(40.0) for s in scale)
    return [Light((10, 10, -5), (1.0, 1.0, 0.65), k),
            Light((2, 5, 12), (1.0, 0.25, 0.25), r),
            Light((-10, 10, -5), (0.65, 0.65, 1.0), f)]


def _box_tet_rotations():
    # 02_Geometry.c:170-180
    rx = quat_axis((1, 0, 0), 0.02)
    ry = quat_axis((0, 1, 0), 0.04)
    rz = quat_axis((0, 0, 1), 0.06)
    box = quat_mul_norm(rx, ry)
    tet = quat_mul_norm(box, rz)
    return box, tet


def ex01_lights():
    s = Scene(ambient=(0.004, 0.004, 0.007))
    s.lights = _three_point_lights((3.3, 3.5, 3.1))
    s.materials = [Material(brdf=BRDF_COOK_TORRANCE, roughness=0.5)]
    s.primitives = [Primitive(PRIM_QUAD, scale=(40, 1, 40))]
    return s, dict(CAMERA)


def ex02_geometry(transparent=False):
    s = Scene(ambient=(0.004, 0.004, 0.007))
    s.lights = _three_point_lights()
    s.materials = [Material(brdf=BRDF_COOK_TORRANCE, roughness=0.5)]
    box_q, tet_q = _box_tet_rotations()
    flags = 63 | (IS_TRANSPARENT if transparent else 0)
    s.primitives = [
        Primitive(PRIM_QUAD, scale=(40, 1, 40)),
        Primitive(PRIM_BOX, position=(-9, 4, 3), scale=(2.5, 2.5, 2.5), rotation=box_q, mesh_id=1, flags=flags),
        Primitive(PRIM_TETRAHEDRON, position=(-3, 4, 12), scale=(2.5, 2.5, 2.5), rotation=tet_q, mesh_id=2, flags=flags),
        Primitive(PRIM_SPHERE, position=(3, 4, 0), scale=(2.5, 2.5, 2.5), mesh_id=3, flags=flags),
    ]
    return s, dict(CAMERA)


def ex03_blinn_phong():
    s = Scene(ambient=(0.004, 0.004, 0.007))
    s.lights = _three_point_lights()
    lambert = Material(albedo=(0.8, 1.0, 0.8))
    phong = Material(brdf=BRDF_PHONG, roughness=0.5, albedo=(0.2, 0.2, 0.3), reflectivity=(1.0, 1.0, 0.4))
    blinn = Material(brdf=BRDF_BLINN, roughness=0.5, albedo=(1.0, 0.3, 1.0), reflectivity=(1.0, 0.4, 1.0))
    s.materials = [lambert, phong, blinn]
    box_q, tet_q = _box_tet_rotations()
    s.primitives = [
        Primitive(PRIM_QUAD, scale=(40, 1, 40)),
        Primitive(PRIM_BOX, position=(-9, 5, 3), scale=(2.5, 2.5, 2.5), rotation=box_q, mesh_id=1, material_id=1),
        Primitive(PRIM_TETRAHEDRON, position=(-3, 4, 12), scale=(2.5, 2.5, 2.5), rotation=tet_q, mesh_id=2, material_id=1),
        Primitive(PRIM_SPHERE, position=(3, 4, 0), scale=(2.5, 2.5, 2.5), mesh_id=3, material_id=2),
    ]
    return s, dict(CAMERA)


def ex04_glass_mirror():
    s = Scene(ambient=(0.004, 0.004, 0.007))
    s.lights = _three_point_lights()
    rough = Material(albedo=(0.8, 1.0, 0.8))
    mirror = Material(is_=REFLECTIVE, reflectivity=(0.9, 0.9, 0.9), roughness=0.0, brdf=BRDF_BLINN)
    glass = Material(is_=REFRACTIVE, n1_over_n2=IOR_AIR / IOR_GLASS, n2_over_n1=IOR_GLASS / IOR_AIR,
                     reflectivity=(0.7, 0.9, 0.7), roughness=0.0, brdf=BRDF_BLINN)
    s.materials = [rough, mirror, glass]
    box_q, tet_q = _box_tet_rotations()
    s.primitives = [
        Primitive(PRIM_QUAD, scale=(40, 1, 40)),
        Primitive(PRIM_BOX, position=(-9, 3, 3), scale=(2.5, 2.5, 2.5), rotation=box_q, mesh_id=1, material_id=2),
        Primitive(PRIM_TETRAHEDRON, position=(-3, 4, 12), scale=(2.5, 2.5, 2.5), rotation=tet_q, mesh_id=2, material_id=2),
        Primitive(PRIM_SPHERE, position=(3, 3, 0), scale=(2.5, 2.5, 2.5), mesh_id=3, material_id=2),
        Primitive(PRIM_QUAD, position=(-15, 5, 5), scale=(4, 1, 8), rotation=(0, 0, -HALF_SQRT2, HALF_SQRT2),
                  mesh_id=4, material_id=1),
    ]
    return s, dict(CAMERA, depth=4)


def ex05_pbr():
    grid, spacing = 7, f32(2.5)
    s = Scene()
    s.lights = [Light((-10.0 if i % 2 else 10.0, 10.0 if i < 2 else -20.0, -10.0), (1, 1, 1), 300.0)
                for i in range(4)]
    plastic = f32(0.04)
    for row in range(grid):
        for col in range(grid):
            metallic = f32(row) / f32(grid)
            rough = f32(col) / f32(grid)
            albedo = (f32(0.5), f32(0), f32(0))
            # lerpVec3(plastic, albedo, metallic) = (albedo - plastic) * metallic + plastic
            refl = tuple((a - plastic) * metallic + plastic for a in albedo)
            mn = rough if rough < f32(1) else f32(1)
            rough = mn if mn > f32(0.05) else f32(0.05)
            s.materials.append(Material(brdf=BRDF_COOK_TORRANCE, metallic=metallic, roughness=rough,
                                        albedo=albedo, reflectivity=refl))
            i = row * grid + col
            x = (f32(col) - f32(grid) / f32(2)) * spacing
            y = (f32(row) - f32(grid) / f32(2)) * spacing
            s.primitives.append(Primitive(PRIM_SPHERE, position=(x, y, 0.0), mesh_id=i, material_id=i))
    return s, dict(position=(0.0, 0.0, -11.0))


def ex06_area_lights(transparent=False):
    s = Scene(ambient=(0.004, 0.004, 0.007))
    s.lights = _three_point_lights()
    s.materials = [Material(albedo=(0.8, 1.0, 0.8)),
                   Material(is_=EMISSIVE, emission=(0.9, 0.7, 0.7)),
                   Material(is_=EMISSIVE, emission=(0.7, 0.7, 0.9))]
    flags = 63 | (IS_TRANSPARENT if transparent else 0)
    s.primitives = [
        Primitive(PRIM_QUAD, scale=(40, 1, 40)),
        Primitive(PRIM_QUAD, position=(-15, 5, 5), scale=(4, 1, 8), rotation=(0, 0, -HALF_SQRT2, HALF_SQRT2),
                  mesh_id=1, material_id=1),
        Primitive(PRIM_QUAD, position=(15, 5, 5), scale=(4, 1, 8), rotation=(0, 0, -HALF_SQRT2, -HALF_SQRT2),
                  mesh_id=2, material_id=2),
        Primitive(PRIM_SPHERE, position=(3, 3, 0), mesh_id=3, flags=flags),
        Primitive(PRIM_SPHERE, position=(-3, 3, 0), mesh_id=4, flags=flags),
    ]
    return s, dict(CAMERA)


# ---- reference .mesh reader (Windows layout, scene/io.h:95-144) ----------

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def read_reference_mesh(path):
    """Parse a .mesh file written by the reference's obj2mesh with 4-byte u32."""
    from slimtracin_amd.scene import NODE_DTYPE, TRIANGLE_DTYPE
    data = open(path, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        arr = np.frombuffer(data, dtype=dtype, count=count, offset=off)
        off += arr.nbytes
        return arr

    aabb = take("<f4", 6)
    vc, tc, ec, uc, nc, node_count, height = (int(x) for x in take("<u4", 7))
    pos = take("<f4", 3 * vc).reshape(vc, 3)
    idx = take("<u4", 3 * tc).reshape(tc, 3)
    take("<u4", 2 * ec)
    uvs = uv_idx = nrm = n_idx = None
    if uc:
        uvs = take("<f4", 2 * uc).reshape(uc, 2)
        uv_idx = take("<u4", 3 * tc).reshape(tc, 3)
    if nc:
        nrm = take("<f4", 3 * nc).reshape(nc, 3)
        n_idx = take("<u4", 3 * tc).reshape(tc, 3)
    tris = take(TRIANGLE_DTYPE, tc)
    nodes = take(NODE_DTYPE, node_count)
    leaf = take("<u4", tc)
    assert off == len(data), (off, len(data))
    return Mesh(pos, idx, nrm, n_idx, uvs, uv_idx, aabb=aabb.copy(), prebuilt=(nodes, tris, leaf, height))


def ex07_monkey(prebuilt=True):
    """07_Meshes with the only shipped mesh (monkey.mesh); dog/dragon are missing blobs."""
    mesh = read_reference_mesh(os.path.join(GOLDEN, "monkey.mesh"))
    if not prebuilt:
        mesh.prebuilt = None
        mesh.aabb = None
    s = Scene(ambient=(0.004, 0.004, 0.007))
    s.lights = _three_point_lights()
    s.materials = [Material(brdf=BRDF_COOK_TORRANCE, roughness=0.5)]
    box_q, tet_q = _box_tet_rotations()
    s.meshes = [mesh]
    s.primitives = [
        Primitive(PRIM_QUAD, scale=(40, 1, 40)),
        Primitive(PRIM_MESH, position=(-4, 1.2, 8), scale=(0.3, 0.3, 0.3), mesh_id=0),
        Primitive(PRIM_MESH, position=(2, 3, 2), scale=(1.5, 1.5, 1.5), rotation=box_q, mesh_id=0),
        Primitive(PRIM_BOX, position=(-9, 3, 3), rotation=box_q, mesh_id=3),
        Primitive(PRIM_TETRAHEDRON, position=(-3, 4, 12), rotation=tet_q, mesh_id=4),
        Primitive(PRIM_SPHERE, position=(3, 3, 0), mesh_id=5),
    ]
    return s, dict(CAMERA)


def checker_texture(size=64, seed=0, mips=True):
    """Procedural RGB texture as a TexelQuad mip chain (core/types.h:56-74).

    Each quad (x,y) holds the 4 texels around grid corner (x,y) with clamped
    neighbours; mips average 2x2 parents, mirroring src/bmp2texture.c.
    """
    rng = np.random.default_rng(seed)
    base = (rng.random((size, size, 3)) * 255).astype(np.uint8)
    out = []
    img = base.astype(np.float64)
    w = size
    while True:
        tex = np.clip(np.rint(img), 0, 255).astype(np.uint8)
        out.append((w, w, texel_quads(tex)))
        if not mips or w <= 4:
            break
        img = 0.25 * (img[0::2, 0::2] + img[1::2, 0::2] + img[0::2, 1::2] + img[1::2, 1::2])
        w //= 2
    return Texture(size, size, out, wrap=1, mipmap=1 if mips else 0)


def texel_quads(img: np.ndarray) -> np.ndarray:
    """(h,w,3) uint8 -> (h+1)*(w+1) quads of 12 bytes: R{TL,TR,BL,BR},G{..},B{..}."""
    h, w, _ = img.shape
    pad = np.pad(img, ((1, 1), (1, 1), (0, 0)), mode="edge")
    tl = pad[0:h + 1, 0:w + 1]
    tr = pad[0:h + 1, 1:w + 2]
    bl = pad[1:h + 2, 0:w + 1]
    br = pad[1:h + 2, 1:w + 2]
    q = np.stack([tl, tr, bl, br], axis=-1)          # (h+1, w+1, 3, 4)
    return np.ascontiguousarray(q.reshape(h + 1, w + 1, 12))


def ex09_textures(mips=True):
    """09_Textures-like scene with procedural albedo/normal maps (the .texture assets are missing)."""
    s, cam = ex02_geometry()
    albedo = checker_texture(64, 1, mips)
    normal = checker_texture(64, 2, mips)
    s.textures = [albedo, normal]
    s.materials = [
        Material(brdf=BRDF_COOK_TORRANCE, roughness=0.5, albedo=(0.7, 0.7, 0.7)),
        Material(brdf=BRDF_COOK_TORRANCE, texture_count=2, texture_ids=(0, 1), uv_repeat=(4, 4)),
        Material(is_=REFLECTIVE, brdf=BRDF_BLINN, reflectivity=(0.3, 0.3, 0.3), metallic=1.0, roughness=0.0,
                 use=NORMAL_MAP, texture_count=2, texture_ids=(0, 1)),
    ]
    s.primitives[0].material_id = 1
    s.primitives[3].material_id = 2
    s.primitives[1].material_id = 0
    return s, dict(cam, depth=3)


ALL_EXAMPLES = {
    "ex01": ex01_lights, "ex02": ex02_geometry, "ex03": ex03_blinn_phong, "ex04": ex04_glass_mirror,
    "ex05": ex05_pbr, "ex06": ex06_area_lights,
}


def crowd(seed=0, n=300, meshes=True):
    """A field of n randomly placed primitives of every type (boxes, spheres,
    tetrahedra, quads, monkey instances) over a floor quad, with lambert,
    Blinn/Phong, Cook-Torrance, glass and mirror materials: the scene-BVH
    rebuild workload (primitives that move every frame)."""
    rng = np.random.default_rng(seed)
    s = Scene(ambient=(0.004, 0.004, 0.007))
    s.lights = _three_point_lights()
    s.materials = [
        Material(albedo=(0.8, 1.0, 0.8)),
        Material(brdf=BRDF_PHONG, roughness=0.5, albedo=(0.2, 0.2, 0.3), reflectivity=(1.0, 1.0, 0.4)),
        Material(brdf=BRDF_BLINN, roughness=0.5, albedo=(1.0, 0.3, 1.0), reflectivity=(1.0, 0.4, 1.0)),
        Material(brdf=BRDF_COOK_TORRANCE, roughness=0.4, metallic=0.3, albedo=(0.6, 0.5, 0.3)),
        Material(is_=REFRACTIVE, n1_over_n2=IOR_AIR / IOR_GLASS, n2_over_n1=IOR_GLASS / IOR_AIR,
                 reflectivity=(0.7, 0.9, 0.7), roughness=0.0, brdf=BRDF_BLINN),
        Material(is_=REFLECTIVE, reflectivity=(0.9, 0.9, 0.9), roughness=0.0, brdf=BRDF_BLINN),
    ]
    types = [PRIM_BOX, PRIM_SPHERE, PRIM_TETRAHEDRON, PRIM_QUAD]
    if meshes:
        s.meshes = [read_reference_mesh(os.path.join(GOLDEN, "monkey.mesh"))]
        types.append(PRIM_MESH)
    s.primitives = [Primitive(PRIM_QUAD, scale=(60, 1, 60))]
    for i in range(n - 1):
        t = types[int(rng.integers(len(types)))]
        s.primitives.append(Primitive(t, position=random_position(rng), scale=tuple(rng.uniform(0.4, 1.6, 3)),
                                      rotation=random_rotation(rng), material_id=int(rng.integers(len(s.materials))),
                                      flags=63))
    return s, dict(position=(0.0, 14.0, -24.0), pitch=-0.35, depth=3)


def random_position(rng):
    return (float(rng.uniform(-25, 25)), float(rng.uniform(0.5, 9)), float(rng.uniform(-8, 40)))


def random_rotation(rng):
    """A float32 unit rotation as the reference composes them (axis-angle
    quaternions multiplied and normalised)."""
    q = quat_axis((1, 0, 0), float(rng.uniform(-3, 3)))
    q = quat_mul_norm(q, quat_axis((0, 1, 0), float(rng.uniform(-3, 3))))
    return tuple(quat_mul_norm(q, quat_axis((0, 0, 1), float(rng.uniform(-3, 3)))))


def move_prims(prims, rng, first=1):
    """Random rigid motion of every primitive from `first` on: new position and
    rotation, as an animation step would set them."""
    out = prims.copy()
    for i in range(first, len(out)):
        out["position"][i] = random_position(rng)
        out["rotation"][i] = random_rotation(rng)
    return out
