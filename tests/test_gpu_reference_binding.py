"""GPU: the reference's own renderer through the reference-side binding.

oracle/_ref/libslimref_seam.so is the reference (its headers, compiled in
place) patched as integration/init_amd.h describes: the device-mirror
functions of core/init.h and renderSceneOnGPU come from the binding, which
calls libslimtracin.so, and renderScene picks the GPU when
viewport->settings.use_GPU is set (render/raytracer.h:302-309). The harness
replays _initApp's order (allocateDeviceScene, uploadScene, updateSceneBVH ->
uploadSceneBVH, uploadMeshBVHs, updateSceneSSB -> uploadPrimitives) and
renders each frame with renderScene twice, use_GPU set and clear: the two
frames must agree (Depth/Normal/UV bit for bit, Beauty within the radiance
tolerance), also after frames that turn primitives and only call
uploadPrimitives (examples/02_Geometry.c:70-75: stale scene BVH and SSB
rects, as the reference keeps them).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

import scenes
from parity import compare_pixels
from slimtracin_amd.scene import MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS, PIXEL_DTYPE, grid_size, make_view

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SEAM_SO = os.path.join(ROOT, "oracle", "_ref", "libslimref_seam.so")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(SEAM_SO), reason="seam test build not made")]


def _seam(blob, view, frames=1):
    lib = C.CDLL(SEAM_SO)
    lib.ref_seam_error.restype = C.c_char_p
    n = grid_size(view)
    gpu = np.zeros(n, dtype=PIXEL_DTYPE)
    cpu = np.zeros(n, dtype=PIXEL_DTYPE)
    buf = np.frombuffer(blob, dtype=np.uint8)
    rc = lib.ref_seam_render(buf.ctypes.data_as(C.c_void_p), C.c_size_t(len(blob)), view.ctypes.data_as(C.c_void_p),
                             C.c_int(frames), gpu.ctypes.data_as(C.c_void_p), cpu.ctypes.data_as(C.c_void_p))
    assert rc == 0, (rc, lib.ref_seam_error())
    return gpu, cpu


# ex09 without mip chains: with them, a bounce ray after a tetrahedron was
# tested reads the cone angle hitTetrahedron leaves uninitialised
# (intersection/tetrahedra.h:11), so the reference's own frame depends on
# stack contents (the oracle and the library read zero there)
SCENES = {"ex02": scenes.ex02_geometry, "ex04": scenes.ex04_glass_mirror, "ex06": scenes.ex06_area_lights,
          "ex07": scenes.ex07_monkey, "ex09": lambda: scenes.ex09_textures(mips=False)}


@pytest.mark.parametrize("name", sorted(SCENES))
@pytest.mark.parametrize("mode", [MODE_BEAUTY, MODE_DEPTH, MODE_NORMALS, MODE_UVS])
def test_render_scene_gpu_equals_cpu(name, mode):
    s, cam = SCENES[name]()
    view = make_view(width=160, height=120, mode=mode, **cam)
    gpu, cpu = _seam(s.to_blob(), view)
    ps = compare_pixels(gpu, cpu, mode)
    print(name, mode, ps, flush=True)
    assert ps["ok"], ps
    assert (cpu["depth"] != np.inf).any()


@pytest.mark.parametrize("name", ["ex02", "ex07"])
def test_upload_primitives_without_bvh_rebuild(name):
    s, cam = SCENES[name]()
    view = make_view(width=128, height=96, mode=MODE_BEAUTY, **cam)
    gpu, cpu = _seam(s.to_blob(), view, frames=4)
    ps = compare_pixels(gpu, cpu, MODE_BEAUTY)
    assert ps["ok"], ps
