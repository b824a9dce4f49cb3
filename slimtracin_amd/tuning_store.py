"""Launch shapes chosen by the autotuner, kept per GPU and workload in a
small JSON file, so that a workload tuned once starts at its shape in every
later process (st_autotune times dozens of candidate launches).

File layout (version 1):
    {"version": 1,
     "gpus": {"<gpu key>": {"<workload key>": {"tuning": {StTuning fields},
                                                "ms": render ms when tuned,
                                                "workload": readable summary}}}}
The GPU key is the architecture, compute-unit count and device name; the
workload key a hash of the batch's shape: per request its scene's size
(primitives, meshes, triangles, mesh BVH height), window size, render mode,
depth and whether it samples with ray cones, in batch order. A file that
cannot be read or does not have this layout is treated as empty and
rewritten on the next store."""
from __future__ import annotations

import hashlib
import json
import os
import tempfile

import numpy as np

from . import Tuning, autotune, device_info, get_tuning, restore_tuning

VERSION = 1
FIELDS = [f for f, _ in Tuning._fields_]


def default_path() -> str:
    """$SLIMTRACIN_TUNING, else ~/.cache/slimtracin_amd/tuning.json."""
    p = os.environ.get("SLIMTRACIN_TUNING")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "slimtracin_amd", "tuning.json")


def gpu_key(info: dict | None = None) -> str:
    info = info if info is not None else device_info()
    return f"{info['arch'].split(':')[0]}/{info['compute_units']}cu/{info['name']}"


def workload_summary(scenes, views) -> list:
    """Per request: (primitives, meshes, triangles, max mesh BVH height,
    width, height, render mode, depth, ray cones)."""
    out = []
    stats = {}
    for s, v in zip(scenes, np.atleast_1d(views)):
        if id(s) not in stats:
            i = s.info()
            stats[id(s)] = (int(i.prim_count), int(i.mesh_count), int(i.triangle_count), int(i.max_mesh_height))
        out.append(list(stats[id(s)]) + [int(v["width"]), int(v["height"]), int(v["render_mode"]), int(v["depth"]),
                                         int(float(v["cone_angle"]) > 0)])
    return out


def workload_key(summary: list) -> str:
    return hashlib.sha1(json.dumps(summary, separators=(",", ":")).encode()).hexdigest()[:20]


def tuning_to_dict(t: Tuning) -> dict:
    return {f: int(getattr(t, f)) for f in FIELDS}


def tuning_from_dict(d: dict) -> Tuning:
    if set(d) != set(FIELDS) or not all(isinstance(d[f], int) and d[f] >= 0 for f in FIELDS):
        raise ValueError(f"not a launch shape: {d!r}")
    return Tuning(*[d[f] for f in FIELDS])


def load(path: str | None = None) -> dict:
    """The store's contents ({"version": 1, "gpus": {}} when the file is
    missing, unreadable or of another layout)."""
    path = path or default_path()
    try:
        with open(path) as f:
            data = json.load(f)
        if isinstance(data, dict) and data.get("version") == VERSION and isinstance(data.get("gpus"), dict):
            return data
    except (OSError, ValueError):
        pass
    return {"version": VERSION, "gpus": {}}


def save(data: dict, path: str | None = None) -> None:
    """Write the store atomically (a temporary file renamed over it)."""
    path = path or default_path()
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=os.path.dirname(os.path.abspath(path)), prefix=".tuning.", suffix=".json")
    try:
        with os.fdopen(fd, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except BaseException:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise


def lookup(data: dict, gkey: str, wkey: str) -> Tuning | None:
    e = data.get("gpus", {}).get(gkey, {}).get(wkey)
    if not isinstance(e, dict):
        return None
    try:
        return tuning_from_dict(e.get("tuning", {}))
    except ValueError:
        return None


def store(data: dict, gkey: str, wkey: str, t: Tuning, ms: float, summary: list) -> dict:
    data.setdefault("gpus", {}).setdefault(gkey, {})[wkey] = {"tuning": tuning_to_dict(t), "ms": round(float(ms), 4),
                                                              "workload": summary}
    return data


def autotune_cached(scenes, views, path: str | None = None, reps: int = 3):
    """The stored shape for this GPU and workload, set as the library's
    tuning; without one, st_autotune's choice, stored. Returns (Tuning,
    render ms when tuned, True if it came from the store)."""
    views = np.atleast_1d(views)
    summary = workload_summary(scenes, views)
    gkey, wkey = gpu_key(), workload_key(summary)
    data = load(path)
    t = lookup(data, gkey, wkey)
    if t is not None:
        saved = get_tuning()
        try:
            restore_tuning(t)
        except Exception:
            restore_tuning(saved)
            t = None
        else:
            return t, data["gpus"][gkey][wkey].get("ms"), True
    best, ms = autotune(scenes, views, reps=reps)
    save(store(data, gkey, wkey, best, ms, summary), path)
    return best, ms, False
