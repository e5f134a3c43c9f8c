"""C-ABI checks that need no GPU: the library loads, exports every entry point that
include/*.h declares, derives parameters like the reference host, validates arguments
before touching a device, and the seeded scene initialiser is deterministic."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(sph_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import sph_hip
    lib = sph_hip.load_library()
    names = _declared_functions()
    assert "sph_world_step" in names and "sph_world_create" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_params_derive_matches_reference_host():
    """main.c:57 dt = DELTA_TIME / (float)NUM_SUBSTEPS in double, stored as float;
    main.c:349 volume = 2 pi R^5 / 15 in double from the double radius literal."""
    import sph_hip
    p = sph_hip.params_for(particle_radius=0.3, smoothing_radius=0.35, g=9.8, box_min=(0, 0, 0),
                           box_max=(10, 6, 6), frame_dt=1 / 120, substeps=3,
                           pressure_multiplier=500, target_density=55)
    assert p.dt == np.float32((1 / 120) / 3.0)
    assert p.volume == np.float32(2.0 * math.pi * 0.35 ** 5 / 15.0)
    assert p.substeps == 3
    p2 = sph_hip.params_for(particle_radius=0.3, smoothing_radius=2.5, g=100, box_min=(0, 0, 0),
                            box_max=(1, 1, 1), frame_dt=0.01, substeps=1,
                            pressure_multiplier=15000, target_density=1.0)
    assert p2.dt == np.float32(0.01)
    assert p2.volume == np.float32(2.0 * math.pi * 2.5 ** 5 / 15.0)


def test_create_rejects_bad_arguments_without_a_device():
    import sph_hip
    lib = sph_hip.load_library()
    h = C.c_void_p()
    assert lib.sph_world_create(None, 16, C.byref(h)) == -1
    p = sph_hip.params_for(particle_radius=0.3, smoothing_radius=0.35, g=9.8, box_min=(0, 0, 0),
                           box_max=(10, 6, 6), frame_dt=1 / 120, substeps=3,
                           pressure_multiplier=500, target_density=55)
    assert lib.sph_world_create(C.byref(p), 0, C.byref(h)) == -1
    assert b"particle count" in lib.sph_last_error()
    bad = sph_hip.SphParams.from_buffer_copy(p)
    bad.smoothing_radius = float("nan")
    assert lib.sph_world_create(C.byref(bad), 16, C.byref(h)) == -1
    bad = sph_hip.SphParams.from_buffer_copy(p)
    bad.max_xyz[1] = -1.0
    assert lib.sph_world_create(C.byref(bad), 16, C.byref(h)) == -1
    assert lib.sph_world_step(None, 1) == -1


def test_batch_rejects_bad_arguments_without_a_device():
    import sph_hip
    lib = sph_hip.load_library()
    h = C.c_void_p()
    p = sph_hip.params_for(particle_radius=0.3, smoothing_radius=0.35, g=9.8, box_min=(0, 0, 0),
                           box_max=(10, 6, 6), frame_dt=1 / 120, substeps=3,
                           pressure_multiplier=500, target_density=55)
    params = (sph_hip.SphParams * 2)(p, p)
    counts = (C.c_uint32 * 2)(100, 0)
    create = lib.sph_batch_create
    assert create(C.c_uint32(2), params, counts, C.byref(h)) == -1
    assert b"no particles" in lib.sph_last_error()
    assert create(C.c_uint32(0), params, counts, C.byref(h)) == -1
    assert create(C.c_uint32(1), None, counts, C.byref(h)) == -1
    q = sph_hip.SphParams.from_buffer_copy(p)
    q.substeps = 2
    counts = (C.c_uint32 * 2)(100, 100)
    assert create(C.c_uint32(2), (sph_hip.SphParams * 2)(p, q), counts, C.byref(h)) == -1
    assert b"same substep count" in lib.sph_last_error()
    assert lib.sph_batch_step(None, 1) == -1


def test_fill_block_is_seeded_and_in_range():
    import sph_hip
    a = sph_hip.fill_block(5000, 3, (1, 2, 3), (4, 5, 6), 0.1)
    b = sph_hip.fill_block(5000, 3, (1, 2, 3), (4, 5, 6), 0.1)
    c = sph_hip.fill_block(5000, 4, (1, 2, 3), (4, 5, 6), 0.1)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])
    pos, vel = a
    assert np.all(pos[:, :3] >= (1, 2, 3)) and np.all(pos[:, :3] <= (4, 5, 6))
    assert np.all(pos[:, 3] == 0) and np.all(vel[:, 3] == 0)
    assert abs(vel[:, :3].std() - 0.1) < 0.005
    assert abs(pos[:, 0].mean() - 2.5) < 0.05


def test_footprint_is_monotone():
    import sph_hip
    lib = sph_hip.load_library()
    f1 = lib.sph_world_footprint(1 << 20, 1 << 20)
    f2 = lib.sph_world_footprint(1 << 20, 1 << 21)
    assert f2 - f1 == 24 * (1 << 20)
    assert lib.sph_world_footprint(1 << 20, 0) == f1


def test_server_rejects_bad_config_without_a_device():
    import sph_hip
    from sph_hip.server import ServerConfig
    lib = sph_hip.load_library()
    h = C.c_void_p()
    cfg = ServerConfig(1 << 30, 0, 1024, 1024, 0, 3)
    assert lib.sph_server_create(C.byref(cfg), C.byref(h)) == -1
    cfg = ServerConfig(1 << 30, 4, 0, 1024, 0, 3)
    assert lib.sph_server_create(C.byref(cfg), C.byref(h)) == -1
    cfg = ServerConfig(1 << 30, 4, 1024, 1024, 0, 0)
    assert lib.sph_server_create(C.byref(cfg), C.byref(h)) == -1


def test_lane_configs_without_a_device():
    import sph_hip
    lib = sph_hip.load_library()
    assert lib.sph_lane_configs() == 10
    assert lib.sph_world_set_lane_config(None, 0) == -1
