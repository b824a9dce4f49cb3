"""Frame and hit-record comparison used by the parity tests.

Bar (stated here so every test applies the same one):
- hit records: primitive id, object type and triangle id identical; the hit
  distance t identical bit for bit (it only involves + - * / and sqrt);
- pixels: opacity and depth identical bit for bit; colour identical bit for
  bit in the Depth / Normal / UV modes (no transcendental functions there);
  in Beauty mode the colour may differ where powf / acosf (device libm vs
  glibc) feed radiance, so it is compared as the displayed value
  sqrt(pixel)/255 (the tone-mapped colour in [0,1], viewport.h:42-45) with
  max-abs tolerance 1e-4.
"""
from __future__ import annotations

import numpy as np

from slimtracin_amd.scene import MODE_BEAUTY

BEAUTY_TOL = 1e-4


def compare_hits(got, want):
    bad = ((got["prim_id"] != want["prim_id"]) | (got["object_type"] != want["object_type"])
           | (got["triangle_id"] != want["triangle_id"]))
    dist_bits = got["distance"].view(np.uint32) != want["distance"].view(np.uint32)
    return {"id_mismatch": int(bad.sum()), "t_mismatch": int(dist_bits.sum()), "n": len(got)}


def compare_pixels(got, want, mode):
    out = {"n": len(got)}
    out["opacity_mismatch"] = int((got["opacity"].view(np.uint32) != want["opacity"].view(np.uint32)).sum())
    out["depth_mismatch"] = int((got["depth"].view(np.uint64) != want["depth"].view(np.uint64)).sum())
    rgb_bits = np.zeros(len(got), dtype=bool)
    for c in "rgb":
        rgb_bits |= got[c].view(np.uint32) != want[c].view(np.uint32)
    out["rgb_bit_mismatch"] = int(rgb_bits.sum())
    if mode == MODE_BEAUTY:
        err = 0.0
        for c in "rgb":
            a = np.sqrt(np.maximum(got[c].astype(np.float64), 0)) / 255.0
            b = np.sqrt(np.maximum(want[c].astype(np.float64), 0)) / 255.0
            both_nan = np.isnan(a) & np.isnan(b)
            d = np.where(both_nan, 0.0, np.abs(a - b))
            d = np.where(np.isnan(d), np.inf, d)
            err = max(err, float(d.max()) if len(d) else 0.0)
        out["max_abs"] = err
        out["ok"] = out["opacity_mismatch"] == 0 and out["depth_mismatch"] == 0 and err <= BEAUTY_TOL
    else:
        out["max_abs"] = 0.0 if out["rgb_bit_mismatch"] == 0 else float("inf")
        out["ok"] = out["opacity_mismatch"] == 0 and out["depth_mismatch"] == 0 and out["rgb_bit_mismatch"] == 0
    return out
