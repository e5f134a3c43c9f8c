"""Server: the continuous-batching request loop of the sph_server_* C-ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .abi import SphParams, check, load_library


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_scenes", C.c_uint32),
                ("max_particles", C.c_uint32), ("max_table_cells", C.c_uint32),
                ("max_bricks", C.c_uint32), ("substeps", C.c_uint32), ("list_entries", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("frames", C.c_uint64), ("resident_scenes", C.c_uint32), ("queued", C.c_uint32),
                ("resident_particles", C.c_uint64), ("resident_bytes", C.c_uint64),
                ("step_p50_ms", C.c_double), ("step_p99_ms", C.c_double),
                ("scene_p50_ms", C.c_double), ("scene_p99_ms", C.c_double),
                ("wait_p50_ms", C.c_double), ("wait_p99_ms", C.c_double),
                ("particle_steps_per_s", C.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class Server:
    def __init__(self, *, hbm_budget_bytes: int, max_scenes: int, max_particles: int,
                 max_table_cells: int, substeps: int, max_bricks: int = 0, list_entries: int = 0):
        self._lib = load_library()
        self._h = C.c_void_p()
        cfg = ServerConfig(hbm_budget_bytes, max_scenes, max_particles, max_table_cells,
                           max_bricks, substeps, list_entries)
        check(self._lib.sph_server_create(C.byref(cfg), C.byref(self._h)))
        self._n = {}

    def close(self) -> None:
        if self._h:
            self._lib.sph_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, params: SphParams, pos: np.ndarray, vel: np.ndarray, steps: int) -> int:
        n = pos.shape[0]
        pos = np.ascontiguousarray(pos, np.float32).reshape(n, 4)
        vel = np.ascontiguousarray(vel, np.float32).reshape(n, 4)
        rid = C.c_uint64()
        check(self._lib.sph_server_submit(self._h, C.byref(params), n, pos.ctypes.data,
                                          vel.ctypes.data, steps, C.byref(rid)))
        self._n[rid.value] = n
        return rid.value

    def poll(self, rid: int) -> str:
        r = self._lib.sph_server_poll(self._h, rid)
        check(min(r, 0))
        return STATES[r]

    def wait(self, rid: int, timeout_s: float = -1.0) -> None:
        check(self._lib.sph_server_wait(self._h, rid, timeout_s))

    def result(self, rid: int):
        n = self._n.pop(rid)
        pos = np.empty((n, 4), np.float32)
        vel = np.empty((n, 4), np.float32)
        check(self._lib.sph_server_result(self._h, rid, pos.ctypes.data, vel.ctypes.data))
        return pos, vel

    def stats(self) -> dict:
        st = ServerStats()
        check(self._lib.sph_server_get_stats(self._h, C.byref(st)))
        return st.as_dict()
