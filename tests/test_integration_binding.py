"""The reference-side binding integration/init_amd.h: the reference's device
mirror (core/init.h:36-126) and renderSceneOnGPU (render/raytracer.h:258-299)
over the library's C ABI (include/slimtracin_seam.h).

CPU: the header compiles against the reference's own core/types.h with
warnings on, and the test build of the patched reference
(oracle/_ref/libslimref_seam.so, made by oracle/Makefile's `ref` target from
the reference's headers in place) exports every function of the seam from
the binding, plus renderScene and renderSceneOnCPU from the reference.
The frames themselves are checked on the GPU (test_gpu_reference_binding.py).
"""
from __future__ import annotations

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src/SlimTracin"
SEAM_SO = os.path.join(ROOT, "oracle", "_ref", "libslimref_seam.so")
SEAM_FUNCTIONS = ["allocateDeviceScene", "uploadScene", "uploadPrimitives", "uploadMaterials", "uploadLights",
                  "uploadSceneBVH", "uploadMeshBVHs", "renderSceneOnGPU", "renderScene", "renderSceneOnCPU"]


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout not present")
def test_binding_compiles_against_reference_types(tmp_path):
    tu = tmp_path / "tu.c"
    tu.write_text('#include "core/types.h"\n#include "init_amd.h"\n')
    r = subprocess.run(["gcc", "-std=gnu11", "-fsyntax-only", "-Wall", "-Wextra", "-Werror=implicit-function-declaration",
                        "-Werror=incompatible-pointer-types", "-Wno-unused-parameter", "-Wno-unused-function",
                        "-Wno-missing-braces", f"-I{REF}", f"-I{ROOT}/include", f"-I{ROOT}/integration", str(tu)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # no warning from the binding itself (the reference's own headers may warn)
    assert "init_amd.h" not in r.stderr, r.stderr


@pytest.mark.skipif(not os.path.exists(SEAM_SO), reason="seam test build not made (needs the reference checkout)")
def test_patched_reference_exports_the_seam():
    out = subprocess.run(["nm", "-D", "--defined-only", SEAM_SO], capture_output=True, text=True, check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [f for f in SEAM_FUNCTIONS + ["ref_seam_render"] if f not in names]
    assert not missing, missing
