"""The tiled neighbour passes (SPH_KERNEL_TILE: cell tiles staged in LDS, sph_tiles.hip)
against the FP32 oracle, bit for bit: every exact tile configuration on the dam break (one
substep, every stage, and ten frames), the 1M scene over three frames (its first clamp
piles particles onto the margin planes, so some tiles' regions exceed the LDS and stream
through it in groups of rows), lists pinned far below the neighbour counts (the force pass
walks the cells instead), scenes whose local density is far above the mean the tiles were
sized for (row groups; a single row too dense for LDS hands the tile's particles to the
stray kernels), the near-pressure term, and the tolerance configuration within 1e-5 of
the FP64 oracle.  Configurations 0 and 1 pair
the tiled density pass with the per-lane force pass over slot lists, 2 and 3 with the tiled
force pass over staged-index lists; the cases that stress the force pass run both."""
import numpy as np
import pytest

from test_gpu_parity import (REL_TOL, STAGES_F32, STAGES_INT, _assert_bitexact, _fp64_errors,
                             _run_one_substep)

pytestmark = pytest.mark.gpu


def _tile_world(gpu, params, n, config):
    w = gpu.World(params, n)
    w.set_tile_config(config)
    assert w.kernel == "tile" and "tiles of" in w.kernel_config
    return w


@pytest.mark.parametrize("config", [0, 1, 2, 3])
def test_tiles_dam_break_every_stage_and_ten_frames(gpu, orc, config):
    from sph_hip.workloads import DAM_BREAK_16K as W
    params = W.params()
    pos, vel = W.state()
    p1 = gpu.SphParams.from_buffer_copy(params)
    p1.substeps = 1
    print(f"case dam break, tile config {config}", flush=True)
    with _tile_world(gpu, p1, W.n, config) as w:
        w.set_debug(True)
        w.upload(pos, vel)
        w.step(1)
        got = {s: w.dump(s) for s in STAGES_INT + STAGES_F32}
        got["pos"], got["vel"] = w.download()
        stats = w.tile_stats()
    rpos, rvel, ref = orc.substep(p1, pos, vel)
    for s in STAGES_INT + STAGES_F32:
        _assert_bitexact(f"tile{config}:{s}", got[s], ref[s])
    _assert_bitexact("pos", got["pos"], rpos)
    _assert_bitexact("vel", got["vel"], rvel)
    assert stats["tiles"] > 0, stats
    with _tile_world(gpu, params, W.n, config) as w:
        w.upload(pos, vel)
        w.step(10)
        gp, gv = w.download()
    rp, rv = orc.step(params, pos, vel, 10)
    _assert_bitexact("pos", gp, rp)
    _assert_bitexact("vel", gv, rv)


@pytest.mark.parametrize("config", [0, 2])
def test_tiles_uniform_1m_three_frames(gpu, orc, config):
    """Three frames of the 1M scene (T = 2^21) on tiles: the margin piles the first clamp
    forms overflow their tiles' LDS; with slot lists such a region streams through LDS in
    groups of rows, with the tiled force pass the tile is handed over; bit-exact
    throughout."""
    from sph_hip.workloads import UNIFORM_1M as W
    params = W.params(table_cells=1 << 21)
    pos, vel = W.state()
    print(f"case n={W.n} frames=3 tile config {config}", flush=True)
    with _tile_world(gpu, params, W.n, config) as w:
        w.upload(pos, vel)
        w.step(3)
        gp, gv = w.download()
        stats = w.tile_stats()
    print("tile stats after 3 frames:", stats, flush=True)
    assert stats["streamed" if config == 0 else "handed_over"] > 0, stats
    rp, rv = orc.step(params, pos, vel, 3)
    _assert_bitexact("pos", gp, rp)
    _assert_bitexact("vel", gv, rv)


@pytest.mark.parametrize("config", [0, 2])
@pytest.mark.parametrize("entries", [1, 16])
def test_tiles_short_lists(gpu, orc, entries, config):
    """Lists pinned far below the dam break's neighbour counts: the force pass walks the
    27 cells instead (through global memory, or of the staged region); ten frames
    bit-exact."""
    from sph_hip.workloads import DAM_BREAK_16K as W
    params = W.params()
    pos, vel = W.state()
    print(f"case short lists {entries}, tile config {config}", flush=True)
    with _tile_world(gpu, params, W.n, config) as w:
        w.set_list_capacity(entries)
        w.upload(pos, vel)
        w.step(10)
        gp, gv = w.download()
    rp, rv = orc.step(params, pos, vel, 10)
    _assert_bitexact("pos", gp, rp)
    _assert_bitexact("vel", gv, rv)


def _background_and_block(gpu, n_bg, n_block, lo, hi):
    import sph_hip
    params = sph_hip.params_for(particle_radius=0.3, smoothing_radius=0.35, g=9.8,
                                box_min=(0, 0, 0), box_max=(10, 6, 6), frame_dt=1 / 120,
                                substeps=1, pressure_multiplier=500, target_density=55)
    p0, v0 = gpu.fill_block(n_bg, 7, (0, 0, 0), (10, 6, 6), 0.0)
    p1, v1 = gpu.fill_block(n_block, 8, lo, hi, 0.0)
    return params, np.concatenate([p0, p1]), np.concatenate([v0, v1])


@pytest.mark.parametrize("config", [0, 2])
@pytest.mark.parametrize("case", ["chunks", "hand_over"])
def test_tiles_dense_blocks(gpu, orc, case, config):
    """A uniform background plus a block far denser than the mean the tiles are sized for:
    ~6 particles per cell (the regions of tiles crossing it stream through LDS in groups of
    rows, or with the tiled force pass are handed over) or ~60 (a single row of the region
    exceeds the LDS: the tile's particles are handed over to the stray kernels).  Two
    substeps bit-exact, and the statistics show the path taken."""
    if case == "chunks":
        params, pos, vel = _background_and_block(gpu, 8000, 4000, (4, 2, 1), (7, 5, 4))
    else:
        params, pos, vel = _background_and_block(gpu, 8000, 12000, (4, 2, 1), (5.5, 3.5, 2.5))
    n = pos.shape[0]
    print(f"case dense block {case} n={n}, tile config {config}", flush=True)
    with _tile_world(gpu, params, n, config) as w:
        w.upload(pos, vel)
        w.step(1)
        stats = w.tile_stats()
        w.step(1)
        gp, gv = w.download()
    print("tile stats after one substep:", stats, flush=True)
    assert stats["streamed" if config == 0 else "handed_over"] > 0, stats
    if case == "hand_over":
        assert stats["handed_over"] > 0 and stats["strays"] > 0, stats
    rp, rv = orc.step(params, pos, vel, 2)
    _assert_bitexact("pos", gp, rp)
    _assert_bitexact("vel", gv, rv)


def test_tiles_near_pressure(gpu, orc):
    """k_near = 50 on tiles: the near-pressure term through the LDS lists, one substep every
    stage and ten frames."""
    from sph_hip.workloads import DAM_BREAK_16K as W
    params = W.params()
    params.near_pressure_multiplier = 50.0
    pos, vel = W.state()
    p1 = gpu.SphParams.from_buffer_copy(params)
    p1.substeps = 1
    got = _run_one_substep(gpu, p1, pos, vel, "tile")
    rpos, rvel, ref = orc.substep(p1, pos, vel)
    for s in STAGES_INT + STAGES_F32:
        _assert_bitexact(f"tile:{s}", got[s], ref[s])
    with _tile_world(gpu, params, W.n, 0) as w:
        w.upload(pos, vel)
        w.step(10)
        gp, gv = w.download()
    rp, rv = orc.step(params, pos, vel, 10)
    _assert_bitexact("pos", gp, rp)
    _assert_bitexact("vel", gv, rv)


def test_tiles_tolerance(gpu, orc):
    """Tile configuration 4 (approximate square root and reciprocal, opt-in): integer stages
    bit-exact, one substep within 1e-5 of the FP64 oracle."""
    from sph_hip.workloads import DAM_BREAK_16K as W
    params = W.params()
    pos, vel = W.state()
    p1 = gpu.SphParams.from_buffer_copy(params)
    p1.substeps = 1
    with gpu.World(p1, W.n) as w:
        with pytest.raises(gpu.SphError):
            w.set_tile_config(4)  # needs the tolerance precision
        w.set_precision("tolerance")
        w.set_tile_config(4)
        assert "approximate" in w.kernel_config
        w.set_debug(True)
        w.upload(pos, vel)
        w.step(1)
        got = {s: w.dump(s) for s in STAGES_INT + STAGES_F32}
        got["pos"], got["vel"] = w.download()
    rpos, rvel, ref = orc.substep(p1, pos, vel)
    for s in STAGES_INT + ("pred_pos", "pred_vel"):
        _assert_bitexact(f"tile tolerance:{s}", got[s], ref[s])
    err = _fp64_errors(gpu, orc, params, pos, vel, got, ref)
    for name, e in err.items():
        assert e.max() <= REL_TOL, (name, e.max())
