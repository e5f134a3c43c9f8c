"""Batch: many independent scenes stepped together through the sph_batch_* C-ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .abi import DUMP_STAGES, KERNELS, PassTimes, SphParams, check, density_map_dims, load_library


class Batch:
    def __init__(self, params: list, counts: list):
        self._lib = load_library()
        self._h = C.c_void_p()
        arr = (SphParams * len(params))(*params)
        cnt = (C.c_uint32 * len(counts))(*counts)
        check(self._lib.sph_batch_create(len(params), arr, cnt, C.byref(self._h)))
        self.params = list(params)
        self.counts = list(counts)
        self.tables = [p.table_cells or n for p, n in zip(params, counts)]

    def close(self) -> None:
        if self._h:
            self._lib.sph_batch_destroy(self._h)
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

    @property
    def particles(self) -> int:
        return int(self._lib.sph_batch_particles(self._h))

    def set_kernel(self, variant: str) -> None:
        check(self._lib.sph_batch_set_kernel(self._h, KERNELS[variant]))

    def set_tile_config(self, config: int) -> None:
        """TILE at one fixed exact tile configuration (0 or 1)."""
        check(self._lib.sph_batch_set_tile_config(self._h, config))

    @property
    def kernel(self) -> str:
        """The kernels the next step uses once tuned ("auto" before the measurement)."""
        r = self._lib.sph_batch_kernel(self._h)
        check(min(r, 0))
        return {0: "auto", 1: "direct", 2: "brick", 3: "tile"}[r]

    def set_list_capacity(self, entries: int) -> None:
        """Fixes the force-list entries per particle (1..254); 0: adaptive sizing."""
        check(self._lib.sph_batch_set_list_capacity(self._h, entries))

    @property
    def list_capacity(self) -> int:
        rc = self._lib.sph_batch_list_capacity(self._h)
        if rc < 0:
            check(rc)
        return rc

    def upload(self, scene: int, pos: np.ndarray, vel: np.ndarray) -> None:
        n = self.counts[scene]
        pos = np.ascontiguousarray(pos, np.float32).reshape(n, 4)
        vel = np.ascontiguousarray(vel, np.float32).reshape(n, 4)
        check(self._lib.sph_batch_upload_scene(self._h, scene, pos.ctypes.data, vel.ctypes.data))

    def uses_bricks(self, scene: int) -> bool:
        r = self._lib.sph_batch_scene_uses_bricks(self._h, scene)
        check(min(r, 0))
        return bool(r)

    def step(self, frames: int = 1) -> None:
        check(self._lib.sph_batch_step(self._h, frames))

    def sync(self) -> None:
        check(self._lib.sph_batch_sync(self._h))

    def download(self, scene: int):
        n = self.counts[scene]
        pos = np.empty((n, 4), np.float32)
        vel = np.empty((n, 4), np.float32)
        check(self._lib.sph_batch_download_scene(self._h, scene, pos.ctypes.data, vel.ctypes.data, None))
        return pos, vel

    def set_debug(self, on: bool) -> None:
        check(self._lib.sph_batch_set_debug(self._h, 1 if on else 0))

    def dump(self, scene: int, stage: str) -> np.ndarray:
        n, T = self.counts[scene], self.tables[scene]
        if stage in ("cell_keys", "sorted_keys", "sorted_index"):
            out = np.empty(n, np.uint32)
        elif stage in ("start", "end"):
            out = np.empty(T, np.int32)
        elif stage in ("density", "near_density"):
            out = np.empty(n, np.float32)
        else:
            out = np.empty((n, 4), np.float32)
        check(self._lib.sph_batch_debug_dump(self._h, scene, DUMP_STAGES[stage], out.ctypes.data, out.nbytes))
        return out

    def density_map(self, scene: int, voxel_size: float) -> np.ndarray:
        """updateDensityMap of one scene after the last substep: shape (dz, dy, dx)."""
        dx, dy, dz = density_map_dims(self.params[scene], voxel_size)
        out = np.empty((dz, dy, dx), np.float32)
        check(self._lib.sph_batch_density_map(self._h, scene, voxel_size, out.ctypes.data, out.size))
        return out

    def pass_times(self, substeps: int) -> dict:
        t = PassTimes()
        check(self._lib.sph_batch_pass_times(self._h, substeps, C.byref(t)))
        return {name: getattr(t, name) for name, _ in PassTimes._fields_}
