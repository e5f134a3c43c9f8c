"""World: one SPH scene on one GPU, through the sph_world_* C-ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .abi import DUMP_STAGES, KERNELS, PassTimes, SphParams, check, load_library


class World:
    """Mirror of the reference's per-frame flow (main.c:142-199): set params
    (updateUniforms), step (recordCommands' substep loop), read back."""

    def __init__(self, params: SphParams, n: int):
        self._lib = load_library()
        self._h = C.c_void_p()
        check(self._lib.sph_world_create(C.byref(params), n, C.byref(self._h)))
        self.n = n
        self.params = params
        self.table_cells = params.table_cells or n

    def close(self) -> None:
        if self._h:
            self._lib.sph_world_destroy(self._h)
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

    def set_params(self, params: SphParams) -> None:
        check(self._lib.sph_world_set_params(self._h, C.byref(params)))
        self.params = params

    def upload(self, pos: np.ndarray, vel: np.ndarray) -> None:
        pos = np.ascontiguousarray(pos, np.float32).reshape(self.n, 4)
        vel = np.ascontiguousarray(vel, np.float32).reshape(self.n, 4)
        check(self._lib.sph_world_upload_state(self._h, pos.ctypes.data, vel.ctypes.data))

    def step(self, frames: int = 1) -> None:
        check(self._lib.sph_world_step(self._h, frames))

    def sync(self) -> None:
        check(self._lib.sph_world_sync(self._h))

    def density_map_dims(self, voxel_size: float) -> tuple:
        dims = (C.c_uint32 * 3)()
        check(self._lib.sph_world_density_map_dims(self._h, voxel_size, C.byref(dims)))
        return tuple(int(d) for d in dims)

    def density_map(self, voxel_size: float) -> np.ndarray:
        """updateDensityMap after the last substep: float32 array of shape (dz, dy, dx)."""
        dx, dy, dz = self.density_map_dims(voxel_size)
        out = np.empty((dz, dy, dx), np.float32)
        check(self._lib.sph_world_density_map(self._h, voxel_size, out.ctypes.data, out.size))
        return out

    def density_map_device(self, voxel_size: float, dev_ptr: int, count: int) -> None:
        """updateDensityMap into device memory at dev_ptr (count floats, x fastest), enqueued
        on the world's stream: complete after sync()."""
        check(self._lib.sph_world_density_map_device(self._h, voxel_size, C.c_void_p(dev_ptr), count))

    def download(self, color: bool = False):
        pos = np.empty((self.n, 4), np.float32)
        vel = np.empty((self.n, 4), np.float32)
        col = np.empty((self.n, 4), np.float32) if color else None
        check(self._lib.sph_world_download_state(
            self._h, pos.ctypes.data, vel.ctypes.data, col.ctypes.data if color else None))
        return (pos, vel, col) if color else (pos, vel)

    def snapshot_async(self, pos: np.ndarray, vel: np.ndarray) -> None:
        """Starts copying the current state into pos / vel (n x 4 float32, C order) and
        returns; the arrays are held until snapshot_wait()."""
        for a in (pos, vel):
            if a.dtype != np.float32 or a.shape != (self.n, 4) or not a.flags.c_contiguous:
                raise ValueError("snapshot arrays must be C-ordered float32 of shape (n, 4)")
        self._snap = (pos, vel)
        check(self._lib.sph_world_snapshot_async(self._h, pos.ctypes.data, vel.ctypes.data))

    def bind_state(self, pos_dev: int | None, vel_dev: int | None) -> None:
        """Steps the caller's device buffers (n vec4 each) in place; None, None unbinds."""
        check(self._lib.sph_world_bind_state(self._h, C.c_void_p(pos_dev) if pos_dev else None,
                                             C.c_void_p(vel_dev) if vel_dev else None))

    def device_state(self) -> tuple:
        """Device addresses (positions, velocities) of the buffers the world steps."""
        p, v = C.c_void_p(), C.c_void_p()
        check(self._lib.sph_world_device_state(self._h, C.byref(p), C.byref(v)))
        return p.value, v.value

    def snapshot_wait(self) -> None:
        check(self._lib.sph_world_snapshot_wait(self._h))
        self._snap = None

    def set_kernel(self, variant: str) -> None:
        """Neighbour-pass kernels: "auto", "direct" (one lane per particle), "brick" or
        "tile" (cell tiles staged in LDS)."""
        check(self._lib.sph_world_set_kernel(self._h, KERNELS[variant]))

    def autotune(self, substeps: int = 3):
        """Times every kernel variant on this world's state (restored afterwards) and keeps
        the fastest; returns (variant name, ms per substep)."""
        v = C.c_int()
        ms = C.c_double()
        check(self._lib.sph_world_autotune(self._h, substeps, C.byref(v), C.byref(ms)))
        names = {b: a for a, b in KERNELS.items()}
        return names[v.value], ms.value

    @property
    def kernel(self) -> str:
        r = self._lib.sph_world_kernel(self._h)
        check(min(r, 0))
        return {1: "direct", 2: "brick", 3: "tile"}[r]

    @property
    def kernel_config(self) -> str:
        buf = C.create_string_buffer(256)
        check(self._lib.sph_world_kernel_config(self._h, buf, 256))
        return buf.value.decode()

    def set_lane_config(self, config: int) -> None:
        """DIRECT with one fixed per-lane launch configuration (see lane_configs())."""
        check(self._lib.sph_world_set_lane_config(self._h, config))

    @staticmethod
    def lane_configs() -> int:
        return int(load_library().sph_lane_configs())

    def set_tile_config(self, config: int) -> None:
        """TILE with one fixed tile configuration (see tile_configs())."""
        check(self._lib.sph_world_set_tile_config(self._h, config))

    @staticmethod
    def tile_configs() -> int:
        return int(load_library().sph_tile_configs())

    def tile_stats(self) -> dict:
        """What the last substep's tiled passes did: tiles, strays, tiles streamed through
        LDS in more than one group of rows, and tiles handed over to the strays (all 0 after
        other kernels)."""
        out = (C.c_uint32 * 4)()
        check(self._lib.sph_world_tile_stats(self._h, out))
        return dict(zip(("tiles", "strays", "streamed", "handed_over"), (int(v) for v in out)))

    def set_precision(self, mode: str) -> None:
        """"exact" (bit-identical to the reference's float order, the default) or
        "tolerance" (approximate square root / reciprocal allowed, 1e-5 relative)."""
        check(self._lib.sph_world_set_precision(self._h, {"exact": 0, "tolerance": 1}[mode]))

    def set_radix_bits(self, bits: int) -> None:
        """Pins the radix digit width (8, 11 or 12 bits)."""
        check(self._lib.sph_world_set_radix_bits(self._h, bits))

    def set_list_capacity(self, entries: int) -> None:
        """Fixes the force-list entries per particle (1..254); 0: adaptive sizing."""
        check(self._lib.sph_world_set_list_capacity(self._h, entries))

    @property
    def list_capacity(self) -> int:
        rc = self._lib.sph_world_list_capacity(self._h)
        if rc < 0:
            check(rc)
        return rc

    @property
    def device_bytes(self) -> int:
        return int(self._lib.sph_world_device_bytes(self._h))

    def set_debug(self, on: bool) -> None:
        check(self._lib.sph_world_set_debug(self._h, 1 if on else 0))

    def dump(self, stage: str) -> np.ndarray:
        sid = DUMP_STAGES[stage]
        if stage in ("cell_keys", "sorted_keys", "sorted_index"):
            out = np.empty(self.n, np.uint32)
        elif stage in ("start", "end"):
            out = np.empty(self.table_cells, np.int32)
        elif stage in ("density", "near_density"):
            out = np.empty(self.n, np.float32)
        else:
            out = np.empty((self.n, 4), np.float32)
        check(self._lib.sph_world_debug_dump(self._h, sid, out.ctypes.data, out.nbytes))
        return out

    def pass_times(self, substeps: int) -> dict:
        t = PassTimes()
        check(self._lib.sph_world_pass_times(self._h, substeps, C.byref(t)))
        return {name: getattr(t, name) for name, _ in PassTimes._fields_}
