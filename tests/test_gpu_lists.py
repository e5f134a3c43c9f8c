"""Force-list capacity on the GPU (sph_world_set_list_capacity, the adaptive sizing, the
batch's and the server's capacities).  A particle whose neighbours do not fit its list
walks its cells in the force pass, so every capacity gives the same bits as the oracle;
uploaded states size the lists from their density, lists that overflowed grow (between
frames of a step call, at a synchronisation, during online tuning) and lists that no
particle comes near shrink."""
import numpy as np
import pytest

from test_gpu_parity import _assert_bitexact

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("entries", [1, 16, 40])
def test_short_lists_stay_bitexact(gpu, orc, entries):
    """The dam break (29 neighbours per particle at rest, up to ~100 while it breaks) on a
    list configuration with lists pinned far below that: 10 frames bit-exact."""
    from sph_hip.workloads import DAM_BREAK_16K as W
    params = W.params()
    pos, vel = W.state()
    print(f"case dam break, lists pinned at {entries}", flush=True)
    with gpu.World(params, W.n) as w:
        w.set_lane_config(3)
        w.set_list_capacity(entries)
        w.upload(pos, vel)
        assert w.list_capacity == entries  # pinned: the upload does not resize
        w.step(10)
        gp, gv = w.download()
        assert w.list_capacity == entries
    rp, rv = orc.step(params, pos, vel, 10)
    _assert_bitexact("pos", gp, rp)
    _assert_bitexact("vel", gv, rv)


def test_upload_sizes_the_lists(gpu):
    """64 entries at creation; an upload raises them to five times the neighbours of a
    uniform fill of the state's bounding box, in steps of 32: 64 for the 1M fill (~12
    neighbours), 160 for the dam break's resting block (~31), 254 for the reference's own
    scene (~216)."""
    from sph_hip.workloads import DAM_BREAK_16K, UNIFORM_1M
    for W, p, want in ((UNIFORM_1M, UNIFORM_1M.params(table_cells=1 << 21), 64),
                       (DAM_BREAK_16K, DAM_BREAK_16K.params(), 160)):
        with gpu.World(p, W.n) as w:
            assert w.list_capacity == 64
            w.upload(*W.state())
            assert w.list_capacity == want, W.name
            if W is UNIFORM_1M:
                # 190 entries per particle fewer than the largest capacity
                saved = gpu.abi.load_library().sph_world_footprint(W.n, 1 << 21) - w.device_bytes
                assert saved >= W.n * 190 * 4, saved
    p = gpu.reference_default_params()
    n = 300000
    with gpu.World(p, n) as w:
        w.upload(*gpu.fill_reference(p, n))
        assert w.list_capacity == 254


def test_lists_grow_at_sync_and_stay_exact(gpu, orc):
    """The dam break from 32 entries on the adaptive path: lists overflow from the start
    (up to ~54 neighbours at rest, ~100 while breaking); the step call itself grows them
    between frames once it sees a list that did not fit, a sync does the same, and the
    12-frame trajectory stays bit-exact across the changes."""
    from sph_hip.workloads import DAM_BREAK_16K as W
    params = W.params()
    pos, vel = W.state()
    with gpu.World(params, W.n) as w:
        w.set_lane_config(3)
        w.upload(pos, vel)
        w.set_list_capacity(32)
        w.set_list_capacity(0)  # adaptive again, from 32 entries
        assert w.list_capacity == 32
        w.step(1)
        w.sync()
        first = w.list_capacity
        w.step(11)
        gp, gv = w.download()
        last = w.list_capacity
    print(f"lists 32 -> {first} -> {last}", flush=True)
    assert first > 32 and last > 32
    rp, rv = orc.step(params, pos, vel, 12)
    _assert_bitexact("pos", gp, rp)
    _assert_bitexact("vel", gv, rv)


def test_lists_shrink_when_no_list_comes_near(gpu, orc):
    """A fill at the 1M scene's number density (~12 neighbours) that starts one unit inside
    every wall, so that no particle is clamped onto a wall (the integrate margin piles
    particles that start nearer onto planes and edges, where lists run long), from 254
    entries on the adaptive path, synchronising after every frame: once eight records in a row
    show no list above half the capacity it halves, down to 64, and the 30-frame trajectory
    stays bit-exact across the reallocations."""
    import dataclasses
    from sph_hip.workloads import UNIFORM_1M
    W = dataclasses.replace(UNIFORM_1M, name="inner_block", n=12480, box_max=(10.0, 5.0, 10.0),
                            block_lo=(1.0, 1.0, 1.0), block_hi=(9.0, 4.0, 9.0))
    p = W.params(table_cells=1 << 14)
    pos, vel = W.state()
    seen = []
    with gpu.World(p, pos.shape[0]) as w:
        w.set_lane_config(3)
        w.upload(pos, vel)
        w.set_list_capacity(254)
        w.set_list_capacity(0)  # adaptive again, from 254 entries
        for f in range(30):  # an interactive host: one frame, then a synchronisation
            w.step(1)
            w.sync()
            if f % 5 == 4:
                seen.append(w.list_capacity)
        gp, gv = w.download()
    print(f"list capacity every 5 frames: {seen}", flush=True)
    assert seen[-1] == 64 and seen[0] > seen[-1]
    rp, rv = orc.step(p, pos, vel, 30)
    _assert_bitexact("pos", gp, rp)
    _assert_bitexact("vel", gv, rv)


def test_lists_grow_during_online_tuning(gpu, orc):
    """An AUTO world whose lists overflow while the online tuner times the candidates: the
    lists grow when the kernel phase ends, the phase runs again, and the trajectory (every
    tuning substep is a real one) is bit-exact."""
    from sph_hip.workloads import DAM_BREAK_16K as W
    params = W.params()
    pos, vel = W.state()
    with gpu.World(params, W.n) as w:
        w.upload(pos, vel)
        w.set_list_capacity(16)
        w.set_list_capacity(0)
        w.step(20)
        gp, gv = w.download()
        print(f"online tuning with growing lists: {w.kernel_config}, {w.list_capacity} entries", flush=True)
        assert w.list_capacity > 16
    rp, rv = orc.step(params, pos, vel, 20)
    _assert_bitexact("pos", gp, rp)
    _assert_bitexact("vel", gv, rv)


def _batch_specs(gpu):
    """A dense resting block beside two sparse fills, own boxes and tables."""
    from sph_hip.workloads import DAM_BREAK_16K as W
    out = []
    for i, (n, box, hi) in enumerate(((6000, (6.0, 4.0, 4.0), (2.0, 4.0, 2.0)),
                                      (3000, (8.0, 5.0, 5.0), (8.0, 5.0, 5.0)),
                                      (2000, (5.0, 5.0, 3.0), (5.0, 5.0, 3.0)))):
        p = gpu.params_for(particle_radius=W.particle_radius, smoothing_radius=W.smoothing_radius,
                           g=W.g, box_min=(0, 0, 0), box_max=box, frame_dt=W.frame_dt,
                           substeps=W.substeps, pressure_multiplier=W.pressure_multiplier,
                           target_density=W.target_density,
                           table_cells=1 << max(0, (n - 1).bit_length()))
        out.append((p, gpu.fill_block(n, 70 + i, (0, 0, 0), hi, 0.0)))
    return out


@pytest.mark.parametrize("entries", [0, 8])
def test_batch_list_capacity(gpu, orc, entries):
    """A batch's one capacity: sized by its densest scene (adaptive), or pinned short; every
    scene bit-exact against the oracle run alone."""
    specs = _batch_specs(gpu)
    with gpu.Batch([p for p, _ in specs], [s[0].shape[0] for _, s in specs]) as b:
        b.set_kernel("direct")
        b.set_list_capacity(entries)
        for i, (_, (pp, vv)) in enumerate(specs):
            b.upload(i, pp, vv)
        if entries:
            assert b.list_capacity == entries
        else:
            assert b.list_capacity >= 64
        b.step(4)
        got = [b.download(i) for i in range(len(specs))]
    for i, (p, (pp, vv)) in enumerate(specs):
        rp, rv = orc.step(p, pp, vv, 4)
        _assert_bitexact(f"scene {i} pos", got[i][0], rp)
        _assert_bitexact(f"scene {i} vel", got[i][1], rv)


def test_server_short_lists(gpu, orc):
    """A server whose lists hold 8 entries: the served scenes walk most of their force cells
    and match the oracle run alone; admission charges 136 + 4 x 8 B per particle."""
    specs = _batch_specs(gpu)
    with gpu.Server(hbm_budget_bytes=8 << 30, max_scenes=8, max_particles=1 << 15,
                    max_table_cells=1 << 15, substeps=3, list_entries=8) as srv:
        ids = [srv.submit(p, pp, vv, 3) for p, (pp, vv) in specs]
        for rid in ids:
            srv.wait(rid, 60.0)
        got = [srv.result(rid) for rid in ids]
    for i, (p, (pp, vv)) in enumerate(specs):
        rp, rv = orc.step(p, pp, vv, 3)
        _assert_bitexact(f"served {i} pos", got[i][0], rp)
        _assert_bitexact(f"served {i} vel", got[i][1], rv)
