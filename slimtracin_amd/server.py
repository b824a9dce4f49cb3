"""ctypes binding of the serving loop (include/slimtracin_server.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import PIXEL_DTYPE, VIEW_DTYPE, RayCounts, SlimTracinError, _check, _ptr, lib, scene_flags

OUTPUT_BGRA8 = 0
OUTPUT_FLOAT_PIXELS = 1
ST_ERR_BUSY = -5


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_batch_requests", C.c_uint32), ("output", C.c_uint32),
                ("max_batch_pixels", C.c_uint64), ("max_in_flight", C.c_uint32), ("reserved", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("launches", C.c_uint64),
                ("rays", C.c_uint64), ("resident_bytes", C.c_uint64), ("budget_bytes", C.c_uint64),
                ("resident_scenes", C.c_uint32), ("queued", C.c_uint32), ("latency_p50_ms", C.c_double),
                ("latency_p99_ms", C.c_double), ("latency_max_ms", C.c_double), ("mean_batch", C.c_double),
                ("prep_ms", C.c_double), ("launch_ms", C.c_double), ("kernel_ms", C.c_double),
                ("alloc_ms", C.c_double), ("overlapped_launches", C.c_uint64), ("max_in_flight", C.c_uint32),
                ("reserved", C.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Server:
    """Continuous-batching server (st_server_*): max_in_flight launches run at
    once, each on its own stream; 0 = the library default (3)."""

    def __init__(self, hbm_budget=0, max_batch_requests=0, output=OUTPUT_BGRA8, max_batch_pixels=0,
                 max_in_flight=0):
        self.output = output
        cfg = ServerConfig(hbm_budget, max_batch_requests, output, max_batch_pixels, max_in_flight, 0)
        h = C.c_void_p()
        _check(lib().st_server_create(C.byref(cfg), C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().st_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, blob: bytes, mesh_lbvh: bool = False, mesh_binned: bool = False) -> int:
        b = np.frombuffer(blob, dtype=np.uint8)
        sid = C.c_uint32()
        _check(lib().st_server_add_scene_ex(self.handle, _ptr(b), C.c_size_t(len(blob)),
                                            C.c_uint32(scene_flags(mesh_lbvh, mesh_binned)), C.byref(sid)))
        return sid.value

    def try_add_scene(self, blob: bytes, mesh_lbvh: bool = False, mesh_binned: bool = False):
        """Scene id, or None when the scene does not fit the HBM budget."""
        b = np.frombuffer(blob, dtype=np.uint8)
        sid = C.c_uint32()
        rc = lib().st_server_add_scene_ex(self.handle, _ptr(b), C.c_size_t(len(blob)),
                                          C.c_uint32(scene_flags(mesh_lbvh, mesh_binned)), C.byref(sid))
        if rc == ST_ERR_BUSY:
            return None
        _check(rc)
        return sid.value

    def remove_scene(self, scene_id: int):
        _check(lib().st_server_remove_scene(self.handle, C.c_uint32(scene_id)))

    def submit(self, scene_id: int, view: np.ndarray) -> int:
        v = np.ascontiguousarray(view, dtype=VIEW_DTYPE)
        t = C.c_uint64()
        _check(lib().st_server_submit(self.handle, C.c_uint32(scene_id), _ptr(v), C.byref(t)))
        return t.value

    def poll(self, ticket: int) -> bool:
        r = C.c_int()
        _check(lib().st_server_poll(self.handle, C.c_uint64(ticket), C.byref(r)))
        return bool(r.value)

    def wait(self, ticket: int, timeout_ms: int = -1, copy: bool = True):
        """(image, counts, latency_ms); image is a copy (uint32 BGRA8 or FloatPixels)."""
        img, nbytes, counts, lat = C.c_void_p(), C.c_size_t(), RayCounts(), C.c_double()
        _check(lib().st_server_wait(self.handle, C.c_uint64(ticket), C.c_int(timeout_ms), C.byref(img),
                                    C.byref(nbytes), C.byref(counts), C.byref(lat)))
        out = None
        if copy:
            raw = C.string_at(img.value, nbytes.value)
            out = np.frombuffer(raw, dtype=np.uint32 if self.output == OUTPUT_BGRA8 else PIXEL_DTYPE).copy()
        return out, counts, lat.value

    def release(self, ticket: int):
        _check(lib().st_server_release(self.handle, C.c_uint64(ticket)))

    def set_limits(self, max_batch_requests: int = 0, max_in_flight: int = 0):
        """Requests per launch and launches in flight (0: the configured value)."""
        _check(lib().st_server_set_limits(self.handle, C.c_uint32(max_batch_requests), C.c_uint32(max_in_flight)))

    def plain_rate(self, views: np.ndarray, scene_ids) -> dict:
        """The requests served as plain launches - one request per launch, one
        launch at a time - over the same resident scenes; requests/s and
        Mrays/s (the limits are restored afterwards)."""
        import time
        self.set_limits(1, 1)
        try:
            t0 = time.perf_counter()
            ts = [self.submit(scene_ids[k], views[k:k + 1]) for k in range(len(views))]
            rays = 0
            lat = []
            for t in ts:
                _, c, l_ms = self.wait(t, copy=False)
                rays += c.primary + c.shadow + c.bounce
                lat.append(l_ms)
                self.release(t)
            dt = time.perf_counter() - t0
        finally:
            self.set_limits(0, 0)
        return {"requests": len(views), "req_per_s": round(len(views) / dt, 1), "mrays": round(rays / dt / 1e6, 2),
                "wall_s": round(dt, 3)}

    def stats(self) -> dict:
        s = ServerStats()
        _check(lib().st_server_stats(self.handle, C.byref(s)))
        return s.as_dict()


__all__ = ["Server", "ServerStats", "OUTPUT_BGRA8", "OUTPUT_FLOAT_PIXELS", "SlimTracinError"]
