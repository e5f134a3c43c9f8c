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
    assert lib.sph_lane_configs() == 12
    assert lib.sph_world_set_lane_config(None, 0) == -1
    assert lib.sph_tile_configs() == 5
    assert lib.sph_world_set_tile_config(None, 0) == -1
    assert lib.sph_world_tile_stats(None, None) == -1


def _reference_block(params, draws, n):
    """renderer.c:573-582 from a list of rand() outputs (4 per particle), in the
    reference's rounding: float arithmetic up to `+ min`, then a double sum (0.5 is a double
    literal) rounded to float on assignment."""
    f32 = np.float32
    scale = (f32(0.25), f32(0.45), f32(0.45))
    r = f32(params.particle_radius)
    out = np.zeros((n, 4), np.float32)
    for i in range(n):
        for a in range(3):
            rf = f32(f32(draws[4 * i + a]) / f32(2147483648.0))
            ext = f32(f32(params.max_xyz[a]) - f32(params.min_xyz[a]))
            head = f32(f32(f32(rf * f32(f32(scale[a] * ext) - f32(f32(2.0) * r))) + r) + f32(params.min_xyz[a]))
            out[i, a] = f32(float(head) + float(f32(f32(1.0) - scale[a])) * 0.5 * float(ext))
    return out


def test_reference_fill_matches_libc_rand():
    """sph_scene_fill_reference with the glibc generator draws exactly what this system's
    libc rand() draws after srand(seed), four numbers per particle, and a reset continues the
    same stream; the velocities are (0, 0, -100)."""
    import sph_hip
    libc = C.CDLL("libc.so.6")
    p = sph_hip.reference_default_params()
    assert tuple(p.min_xyz) == (-140.0, -40.0, -40.0) and tuple(p.max_xyz) == (140.0, 40.0, 40.0)
    assert (p.smoothing_radius, p.g, p.pressure_multiplier, p.target_density) == (2.5, 100.0, 15000.0, 1.0)
    assert p.dt == np.float32(0.01) and p.substeps == 1
    n = 500
    for seed in (0, 7, 3000000000):
        libc.srand(C.c_uint(seed))
        draws = [libc.rand() for _ in range(8 * n)]
        pos, vel = sph_hip.fill_reference(p, n, seed=seed)
        assert np.array_equal(pos.view(np.uint32), _reference_block(p, draws, n).view(np.uint32)), seed
        assert (vel[:, :3] == (0.0, 0.0, -100.0)).all() and (vel[:, 3] == 0).all() and (pos[:, 3] == 0).all()
        pos1, _ = sph_hip.fill_reference(p, n, seed=seed, resets=1)
        assert np.array_equal(pos1.view(np.uint32), _reference_block(p, draws[4 * n:], n).view(np.uint32))
    # the block: 25 / 45 / 45 % of the box, centred, inset by the particle radius
    pos, _ = sph_hip.fill_reference(p, 300000)
    lo = np.array([-35.0 + 0.3, -18.0 + 0.3, -18.0 + 0.3])
    hi = np.array([35.0 - 0.3, 18.0 - 0.3, 18.0 - 0.3])
    assert (pos[:, :3] >= lo - 1e-4).all() and (pos[:, :3] <= hi + 1e-4).all()
    assert (pos[:, :3].min(0) < lo + 0.1).all() and (pos[:, :3].max(0) > hi - 0.1).all()


def test_reference_fill_msvcrt_generator():
    """The MSVCRT generator (the reference's Windows build): its well-known first outputs
    after srand(1), 41 18467 6334 26500 19169, give the first particle's coordinates."""
    import sph_hip
    p = sph_hip.reference_default_params()
    pos, _ = sph_hip.fill_reference(p, 2, seed=1, rng="msvcrt")
    f32 = np.float32
    draws = [41, 18467, 6334, 26500, 19169]
    scale = (f32(0.25), f32(0.45), f32(0.45))
    for a in range(3):
        rf = f32(f32(draws[a]) / f32(32768.0))
        ext = f32(f32(p.max_xyz[a]) - f32(p.min_xyz[a]))
        head = f32(f32(f32(rf * f32(f32(scale[a] * ext) - f32(0.6))) + f32(0.3)) + f32(p.min_xyz[a]))
        assert pos[0, a] == f32(float(head) + float(f32(1.0) - scale[a]) * 0.5 * float(ext))
    # the second particle starts at the fifth draw (the fourth is the unused shift draw)
    rf = f32(f32(draws[4]) / f32(32768.0))
    assert pos[1, 0] == f32(float(f32(f32(f32(rf * f32(f32(f32(0.25) * f32(280.0)) - f32(0.6))) + f32(0.3)) +
                                  f32(-140.0))) + float(f32(0.75)) * 0.5 * 280.0)
