"""The allocation-free step (sph_step_noalloc) against the existing world path
(sph_world_create + upload + sph_world_step + download) on the same state: bit-exact on two
calls in a row, the first at the entry point's maximum size, the second smaller (the static
scratch array keeps the first call's contents, so the second checks that what must start
cleared is cleared)."""
import ctypes as C

import numpy as np
import pytest

from test_gpu_device_state import Hip
from test_gpu_parity import _assert_bitexact

pytestmark = pytest.mark.gpu


def _world_result(gpu, params, pos, vel, frames):
    with gpu.World(params, pos.shape[0]) as w:
        w.upload(pos, vel)
        w.step(frames)
        return w.download()


def _noalloc_result(gpu, hip, params, pos, vel, frames):
    lib = gpu.load_library()
    dpos, dvel = hip.malloc(pos.nbytes), hip.malloc(vel.nbytes)
    try:
        hip.h2d(dpos, pos)
        hip.h2d(dvel, vel)
        rc = lib.sph_step_noalloc(C.byref(params), pos.shape[0], C.c_void_p(dpos), C.c_void_p(dvel),
                                  frames, None)
        gpu.abi.check(rc)
        hip.sync()
        gp = np.empty_like(pos)
        gv = np.empty_like(vel)
        hip.d2h(gp, dpos)
        hip.d2h(gv, dvel)
        return gp, gv
    finally:
        hip.sync()
        hip.free(dpos)
        hip.free(dvel)


def test_noalloc_matches_world_two_calls(gpu):
    hip = Hip()
    n_max = 262144
    # call 1: the maximum particle count and table size, a uniform fill at density 55
    a = float(np.float32((n_max / 110.0) ** (1.0 / 3.0)))
    p1 = gpu.params_for(particle_radius=0.3, smoothing_radius=0.35, g=9.8, box_min=(-1.5, 0.25, -3.0),
                        box_max=(-1.5 + 2 * a, 0.25 + a, -3.0 + a), frame_dt=1 / 120, substeps=3,
                        pressure_multiplier=500.0, target_density=55.0, table_cells=1 << 19)
    pos1, vel1 = gpu.fill_block(n_max, 21, tuple(p1.min_xyz), tuple(p1.max_xyz), 0.1)
    print(f"noalloc call 1: n={n_max} T={p1.table_cells}", flush=True)
    want1 = _world_result(gpu, p1, pos1, vel1, 2)
    got1 = _noalloc_result(gpu, hip, p1, pos1, vel1, 2)
    _assert_bitexact("call 1 pos", got1[0], want1[0])
    _assert_bitexact("call 1 vel", got1[1], want1[1])
    # call 2: smaller, another box and table (T = N)
    from sph_hip.workloads import DAM_BREAK_16K as W
    p2 = W.params()
    pos2, vel2 = W.state(12000)
    print("noalloc call 2: n=12000", flush=True)
    want2 = _world_result(gpu, p2, pos2, vel2, 3)
    got2 = _noalloc_result(gpu, hip, p2, pos2, vel2, 3)
    _assert_bitexact("call 2 pos", got2[0], want2[0])
    _assert_bitexact("call 2 vel", got2[1], want2[1])


def test_noalloc_rejects_beyond_maximum(gpu):
    lib = gpu.load_library()
    from sph_hip.workloads import DAM_BREAK_16K as W
    p = W.params()
    hip = Hip()
    d = hip.malloc(16 * 16)
    try:
        assert lib.sph_step_noalloc(C.byref(p), 262145, C.c_void_p(d), C.c_void_p(d + 128), 1, None) == -5
        p.table_cells = 1 << 20
        assert lib.sph_step_noalloc(C.byref(p), 8, C.c_void_p(d), C.c_void_p(d + 128), 1, None) == -5
        assert b"allocation-free maximum" in lib.sph_last_error()
    finally:
        hip.free(d)
