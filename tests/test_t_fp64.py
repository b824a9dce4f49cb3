"""Hit distance t against an FP64 oracle.

For every primary hit of the CPU oracle (whose hit records the GPU matches
bit for bit, test_gpu_parity.py), the ray is rebuilt exactly as the
reference builds it (sequential fp32 sums of the projection plane, fp32
normalisation, raytracer.h:156-170) and intersected in float64 with the hit
primitive's exact surface: unit quad plane, unit box slabs, unit sphere,
the tetrahedron's four face planes (tetrahedra.h:6-7,16-24), or the hit
triangle of a mesh, all through the primitive's transform (scale, then the
quaternion rotation, then translation; scene/primitive.h:8-51). The fp32 t
is compared with the float64 root nearest to it; the worst relative error
is printed and bounded.
"""
from __future__ import annotations

import numpy as np
import pytest

import checkers as ck
import scenes
from slimtracin_amd.scene import MODE_DEPTH, blob_mesh_info, blob_records, make_view

pytestmark = pytest.mark.skipif(not ck.oracle_available(), reason="oracle not built")
f32, f64 = np.float32, np.float64
C577 = f64(f32(0.577350259))
# worst relative error of the fp32 hit distance over all the cases below
REL_TOL = 2e-5


def primary_directions(view):
    """fp32 primary directions in the reference's order of operations."""
    cam = ck.oracle_camera(view)
    start, right, down = cam[0:3].astype(f32), cam[3:6].astype(f32), cam[6:9].astype(f32)
    w1, h1 = int(view["width"][0]) + 1, int(view["height"][0]) + 1
    dirs = np.zeros((h1, w1, 3), dtype=f32)
    s = start.copy()
    for y in range(h1):
        cur = s.copy()
        row = np.zeros((w1, 3), dtype=f32)
        for x in range(w1):
            row[x] = cur
            cur = (cur + right).astype(f32)
        dirs[y] = row
        s = (s + down).astype(f32)
    # normVec3: v * (1 / sqrt(x*x + y*y + z*z)), left to right, fp32
    d = dirs.reshape(-1, 3)
    l2 = ((d[:, 0] * d[:, 0]).astype(f32) + (d[:, 1] * d[:, 1]).astype(f32)).astype(f32)
    l2 = (l2 + (d[:, 2] * d[:, 2]).astype(f32)).astype(f32)
    inv = (f32(1) / np.sqrt(l2)).astype(f32)
    return (d * inv[:, None]).astype(f32)


def qrot(v, q):
    """mulVec3Quat in float64: v + 2*(w*c + a x c), c = a x v."""
    a, w = q[:3], q[3]
    c = np.cross(a, v)
    return v + 2.0 * (w * c + np.cross(a, c))


def candidates(kind, ro, rd, tri=None):
    out = []
    if kind == "quad":
        if rd[1] != 0:
            out.append(-ro[1] / rd[1])
    elif kind == "box":
        for ax in range(3):
            if rd[ax] != 0:
                out += [(-1 - ro[ax]) / rd[ax], (1 - ro[ax]) / rd[ax]]
    elif kind == "sphere":
        a, b, c = rd @ rd, 2 * (ro @ rd), ro @ ro - 1
        disc = b * b - 4 * a * c
        if disc >= 0:
            out += [(-b - np.sqrt(disc)) / (2 * a), (-b + np.sqrt(disc)) / (2 * a)]
    elif kind == "tet":
        planes = [((-C577, -C577, -C577), (-C577, C577, -C577)), ((-C577, -C577, -C577), (C577, -C577, -C577)),
                  ((-C577, -C577, -C577), (-C577, -C577, C577)), ((C577, -C577, C577), (C577, C577, C577))]
        for p, n in planes:
            p, n = np.array(p), np.array(n)
            if rd @ n != 0:
                out.append(((p - ro) @ n) / (rd @ n))
    elif kind == "tri":
        v1, v2, v3 = tri
        n = np.cross(v3 - v1, v2 - v1)
        if rd @ n != 0:
            out.append(((v1 - ro) @ n) / (rd @ n))
    return out


KIND = {1: "mesh", 2: "box", 3: "tet", 4: "quad", 5: "sphere"}


def mesh_triangles(blob, m):
    info = blob_mesh_info(blob)[m]
    o = info["offset"] + 56
    V, T = info["vertex_count"], info["triangle_count"]
    pos = np.frombuffer(blob, dtype=f32, count=3 * V, offset=o).reshape(V, 3).astype(f64)
    idx = np.frombuffer(blob, dtype=np.uint32, count=3 * T, offset=o + 12 * V).reshape(T, 3)
    return pos, idx


def worst_rel_error(blob, view, hits=None):
    if hits is None:
        hits = ck.oracle_render(blob, view)[1]
    rec = blob_records(blob)
    prims = rec["prims"]
    dirs = primary_directions(view)
    origin = view["position"][0].astype(f64)
    meshes = {}
    worst, n = 0.0, 0
    for i in np.nonzero(hits["prim_id"] >= 0)[0]:
        h = hits[i]
        p = prims[h["prim_id"]]
        kind = KIND[int(p["type"])]
        d = dirs[i].astype(f64)
        q = p["rotation"].astype(f64)
        qc = np.array([-q[0], -q[1], -q[2], q[3]])
        pos, scale = p["position"].astype(f64), p["scale"].astype(f64)
        ro = qrot(origin - pos, qc) / scale
        rd = qrot(d, qc) / scale
        tri = None
        if kind == "mesh":
            m = int(p["mesh_id"])
            if m not in meshes:
                T = blob_mesh_info(blob)[m]["triangle_count"]
                leaf = ck.oracle_mesh_bvh(blob, m, T)[2]
                meshes[m] = (mesh_triangles(blob, m), np.argsort(leaf))
            (vp, vi), to_tri = meshes[m]
            t_id = to_tri[h["triangle_id"]]
            tri = [vp[vi[t_id, 0]], vp[vi[t_id, 1]], vp[vi[t_id, 2]]]
            kind = "tri"
        ts = candidates(kind, ro, rd, tri)
        t32 = float(h["distance"])
        # world distance of the object-space hit point (the affine t is the world t)
        dists = [np.linalg.norm(qrot((ro + t * rd) * scale, q) + pos - origin) for t in ts if t > 0]
        t64 = min(dists, key=lambda v: abs(v - t32))
        worst = max(worst, abs(t32 - t64) / t64)
        n += 1
    return worst, n


CASES = {
    "ex02": lambda: scenes.ex02_geometry(),
    "ex04": lambda: scenes.ex04_glass_mirror(),
    "ex07": lambda: scenes.ex07_monkey(prebuilt=False),
    "crowd": lambda: scenes.crowd(6, 60),
}


@pytest.mark.parametrize("name", list(CASES))
def test_hit_distance_against_fp64(name):
    s, cam = CASES[name]()
    view = make_view(width=64, height=48, mode=MODE_DEPTH, **{k: v for k, v in cam.items() if k != "depth"})
    worst, n = worst_rel_error(s.to_blob(), view)
    print(f"{name}: {n} primary hits, worst relative error of t vs float64: {worst:.3e}", flush=True)
    assert n > 100
    assert worst < REL_TOL, worst


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_hit_distance_against_fp64(name):
    import slimtracin_amd as st
    s, cam = CASES[name]()
    blob = s.to_blob()
    view = make_view(width=64, height=48, mode=MODE_DEPTH, **{k: v for k, v in cam.items() if k != "depth"})
    ds = st.DeviceScene(blob)
    try:
        _, hits, _ = st.render(ds, view)
    finally:
        ds.close()
    worst, n = worst_rel_error(blob, view, hits)
    print(f"{name} (GPU): {n} primary hits, worst relative error of t vs float64: {worst:.3e}", flush=True)
    assert n > 100 and worst < REL_TOL, worst
