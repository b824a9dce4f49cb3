"""The binned SAH builder (host, for large meshes; no counterpart in the
reference): the library's build equals the oracle's restatement field for
field; the trees are valid reference-layout BVHs over every box once; the
reference renderer finds the same hits over them as over its own SAH trees;
and their SAH cost is that of the reference's exact SAH builder within 2%
at the same leaf cap, below the Morton LBVH's."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import checkers as ck
import slimtracin_amd as st
from slimtracin_amd.scene import (MODE_BEAUTY, MODE_NORMALS, NODE_DTYPE, blob_mesh_info, blob_with_prebuilt_meshes,
                                  make_view)


def lib_binned_build(boxes, max_leaf):
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = len(boxes)
    nodes = np.zeros(2 * n + 2, dtype=NODE_DTYPE)
    leaf = np.zeros(n + 1, dtype=np.uint32)
    h = C.c_uint32()
    fn = st.lib().st_binned_build_boxes
    fn.restype = C.c_uint32
    cnt = fn(boxes.ctypes.data_as(C.c_void_p), C.c_uint32(n), C.c_uint32(max_leaf),
             nodes.ctypes.data_as(C.c_void_p), leaf.ctypes.data_as(C.c_void_p), C.byref(h))
    return nodes[:cnt], leaf[:n], h.value


def _boxes(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        lo = rng.uniform(-5, 5, size=(n, 3))
    elif kind == "clusters":                 # a few tight clusters and many equal centroids
        lo = rng.normal(size=(4, 3))[rng.integers(0, 4, n)] + rng.normal(scale=1e-3, size=(n, 3)).round(2)
    elif kind == "equal":                    # every centroid the same: halving in range order
        lo = np.zeros((n, 3))
    else:                                    # a flat sheet: zero centroid extent on one axis
        lo = rng.uniform(-1, 1, size=(n, 3))
        lo[:, 1] = 0.5
    ext = rng.exponential(0.05, size=(n, 3)) if kind != "equal" else np.full((n, 3), 0.1)
    return np.concatenate([lo, lo + ext], 1).astype(np.float32)


@pytest.mark.parametrize("kind", ["uniform", "clusters", "equal", "sheet"])
@pytest.mark.parametrize("n", [1, 4, 5, 17, 1000, 40000])
def test_library_build_equals_oracle(kind, n):
    b = _boxes(kind, n, n)
    for max_leaf in (1, 4):
        got, want = lib_binned_build(b, max_leaf), ck.oracle_binned_build(b, max_leaf)
        assert len(got[0]) == len(want[0]) and got[0].tobytes() == want[0].tobytes(), (kind, n, max_leaf)
        assert (got[1] == want[1]).all() and got[2] == want[2]


def _lopsided_boxes(n):
    """Centroids spaced exponentially along x: every split peels a few boxes
    off the top bin, a long chain of uneven splits."""
    x = np.power(np.float64(1.00025), np.arange(n))
    lo = np.stack([x, np.zeros(n), np.zeros(n)], 1)
    return np.concatenate([lo, lo + 0.5], 1).astype(np.float32)


@pytest.mark.parametrize("kind", ["uniform", "clusters", "sphere", "lopsided"])
def test_large_build_equals_oracle(kind):
    """300,000 boxes: the library builds the top of the tree, then subtrees
    on worker threads, and places them where a single-threaded build puts
    them (the lopsided input's top is a long chain of uneven splits); the
    tree equals the oracle's slot for slot."""
    n = 300_000
    b = {"uniform": lambda: _boxes("uniform", n, 11), "clusters": lambda: _boxes("clusters", n, 12),
         "sphere": lambda: _sphere_boxes(n, 13), "lopsided": lambda: _lopsided_boxes(n)}[kind]()
    for max_leaf in ((1, 4) if kind == "uniform" else (1,)):
        got, want = lib_binned_build(b, max_leaf), ck.oracle_binned_build(b, max_leaf)
        assert len(got[0]) == len(want[0]) and got[0].tobytes() == want[0].tobytes(), (kind, max_leaf)
        assert (got[1] == want[1]).all() and got[2] == want[2]


@pytest.mark.parametrize("kind", ["uniform", "clusters", "equal", "sheet"])
def test_tree_invariants(kind):
    n = 5000
    b = _boxes(kind, n, 7)
    nodes, leaf, height = ck.oracle_binned_build(b, 4)
    assert sorted(leaf.tolist()) == list(range(n))                 # a permutation
    assert len(nodes) <= 2 * n - 1
    reach = ck.reachable_nodes(nodes)
    assert len(reach) == len(nodes)                                # compact: every slot in the tree
    covered = np.zeros(n, dtype=np.int64)
    for o in reach:
        nd = nodes[o]
        if nd["child_count"]:
            f, c = int(nd["first_child_id"]), int(nd["child_count"])
            assert 1 <= c <= 4
            covered[f:f + c] += 1
            mine = b[leaf[f:f + c]]
            assert (nd["min"] == mine[:, :3].min(0)).all() and (nd["max"] == mine[:, 3:].max(0)).all()
        else:
            l, r = nodes[nd["first_child_id"]], nodes[nd["first_child_id"] + 1]
            # the parent's box is exactly the union of its children's
            assert (nd["min"] == np.minimum(l["min"], r["min"])).all()
            assert (nd["max"] == np.maximum(l["max"], r["max"])).all()
            assert l["depth"] == r["depth"] == nd["depth"] + 1
    assert (covered == 1).all()
    assert height == int(nodes["depth"].max()) + 1


def test_sah_split_beats_halving():
    """Two far-apart groups of unequal size: the root split separates the
    groups (a count median would cut the larger group)."""
    rng = np.random.default_rng(3)
    a = rng.uniform(0, 1, size=(30, 3))
    c = rng.uniform(100, 101, size=(10, 3))
    lo = np.concatenate([a, c])
    b = np.concatenate([lo, lo + 0.01], 1).astype(np.float32)
    nodes, leaf, _ = ck.oracle_binned_build(b, 4)
    l, r = nodes[nodes[0]["first_child_id"]], nodes[nodes[0]["first_child_id"] + 1]
    assert l["max"][0] < 2 and r["min"][0] > 99


def _sphere_boxes(n, seed):
    """Boxes of small triangles over a unit sphere (a mesh-like distribution)."""
    v = np.random.default_rng(seed).normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return np.concatenate([v - 0.004, v + 0.004], 1).astype(np.float32)


@pytest.mark.parametrize("kind", ["uniform", "sphere"])
def test_sah_cost_against_reference_and_lbvh(kind):
    """At the same leaf cap the binned tree costs what the reference's exact
    SAH tree costs within 2%, and with one box per leaf clearly less than the
    Morton LBVH's."""
    b = _boxes("uniform", 20000, 1) if kind == "uniform" else _sphere_boxes(20000, 5)
    for max_leaf in (1, 4):
        cost = ck.sah_cost(ck.oracle_binned_build(b, max_leaf)[0])
        ref = ck.sah_cost(ck.oracle_sah_build(b, max_leaf)[0])
        assert cost <= 1.02 * ref, (kind, max_leaf, cost, ref)
    lbvh = ck.sah_cost(ck.oracle_lbvh_build(b)[0])
    assert ck.sah_cost(ck.oracle_binned_build(b, 1)[0]) < 0.95 * lbvh


@pytest.mark.parametrize("tris", [1, 5, 2000, 30000])
def test_mesh_tree_and_records(tris):
    """Mesh trees hold one triangle per leaf; each triangle record equals the
    reference SAH build's record of the same triangle."""
    blob = st.gen_mesh_scene(21, tris)
    T = blob_mesh_info(blob)[0]["triangle_count"]
    nodes, trirecs, leaf, h = ck.oracle_mesh_binned(blob, 0, T)
    snodes, stris, sleaf, _ = ck.oracle_mesh_bvh(blob, 0, T)        # reference SAH build (pinned elsewhere)
    assert len(nodes) == 2 * T - 1
    leaves = nodes[nodes["child_count"] > 0]
    assert len(leaves) == T and (leaves["child_count"] == 1).all()
    assert sorted(leaves["first_child_id"].tolist()) == list(range(T))
    # leaf order permutes the triangles
    tid = np.argsort(leaf)
    assert sorted(tid.tolist()) == list(range(T))
    assert trirecs.tobytes() == stris[sleaf[tid]].tobytes()
    assert h == int(nodes["depth"].max()) + 1


def _mesh_arrays(blob, m=0):
    """Mesh m's geometry arrays as the blob stores them after its record."""
    mi = blob_mesh_info(blob)[m]
    V, T, N, U = mi["vertex_count"], mi["triangle_count"], mi["normals_count"], mi["uvs_count"]
    o = mi["offset"] + 56
    def take(dtype, count):
        nonlocal o
        a = np.frombuffer(blob, dtype=dtype, count=count, offset=o).copy()
        o += a.nbytes
        return a
    pos, pidx = take(np.float32, 3 * V), take(np.uint32, 3 * T)
    nrm = nidx = uvs = uidx = None
    if N:
        nrm, nidx = take(np.float32, 3 * N), take(np.uint32, 3 * T)
    if U:
        uvs, uidx = take(np.float32, 2 * U), take(np.uint32, 3 * T)
    return T, N, U, pos, pidx, nrm, nidx, uvs, uidx


@pytest.mark.parametrize("tris", [5, 30000, 200000])
def test_library_mesh_build_equals_oracle(tris):
    """st_build_mesh_binned (tree, triangle records in leaf order, leaf ids)
    equals the oracle's binned mesh build; 200k triangles take the threaded
    path."""
    from slimtracin_amd.scene import TRIANGLE_DTYPE
    blob = st.gen_mesh_scene(23, tris)
    T, N, U, pos, pidx, nrm, nidx, uvs, uidx = _mesh_arrays(blob)
    nodes = np.zeros(2 * T + 2, dtype=NODE_DTYPE)
    trirecs = np.zeros(T, dtype=TRIANGLE_DTYPE)
    leaf = np.zeros(T, dtype=np.uint32)
    cnt, h = C.c_uint32(), C.c_uint32()
    P = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
    rc = st.lib().st_build_mesh_binned(C.c_uint32(T), P(pos), P(pidx), P(nrm), P(nidx), C.c_uint32(N), P(uvs),
                                       P(uidx), C.c_uint32(U), P(nodes), C.byref(cnt), C.byref(h), P(trirecs),
                                       P(leaf))
    assert rc == 0
    onodes, otris, oleaf, oh = ck.oracle_mesh_binned(blob, 0, T)
    assert cnt.value == len(onodes) and h.value == oh
    assert nodes[:cnt.value].tobytes() == onodes.tobytes()
    assert trirecs.tobytes() == otris.tobytes() and (leaf == oleaf).all()


@pytest.mark.skipif(not ck.ref_available(), reason="reference build missing")
def test_reference_traversal_over_mesh_binned():
    """The reference renderer over binned mesh BVHs finds the same primary
    hits as over its own SAH trees, and the oracle renders the binned-tree
    scene as the reference does."""
    blob = st.gen_mesh_scene(8, 20000)
    T = blob_mesh_info(blob)[0]["triangle_count"]
    built = ck.oracle_mesh_binned(blob, 0, T)
    bb = blob_with_prebuilt_meshes(blob, [built])
    view = make_view(position=(0.0, 1.5, -4.0), pitch=-0.2, width=96, height=72, mode=MODE_NORMALS)
    a, b = ck.ref_primary_hits(blob, view), ck.ref_primary_hits(bb, view)
    assert (a["prim_id"] == b["prim_id"]).all() and a["distance"].tobytes() == b["distance"].tobytes()
    sah_leaf = ck.oracle_mesh_bvh(blob, 0, T)[2]
    to_tri_sah, to_tri_binned = np.argsort(sah_leaf), np.argsort(built[2])
    hit = a["triangle_id"] >= 0
    assert (hit == (b["triangle_id"] >= 0)).all() and hit.sum() > 1000
    assert (to_tri_sah[a["triangle_id"][hit]] == to_tri_binned[b["triangle_id"][hit]]).all()
    view["render_mode"] = MODE_BEAUTY
    view["depth"] = 3
    assert ck.pixels_equal(ck.ref_render(bb, view), ck.oracle_render(bb, view)[0])
    # the oracle's own binned build inside a scene equals the prebuilt one
    osc = ck.OracleScene(blob, mesh_binned=True)
    try:
        got_px, got_hits = osc.render(view, threads=4)[:2]
    finally:
        osc.close()
    want_px, want_hits, _ = ck.oracle_render(bb, view)
    assert ck.pixels_equal(got_px, want_px) and got_hits.tobytes() == want_hits.tobytes()
