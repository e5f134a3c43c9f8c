"""Batched step on the GPU: every scene of a mixed batch matches the FP32 oracle run on that
scene alone, bit for bit (same arithmetic, own table segment, stable segmented sort), for
the per-lane, brick and automatic kernel choices."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bitexact(name, got, ref):
    bad = np.flatnonzero(got.view(np.uint32).ravel() != ref.view(np.uint32).ravel())
    assert bad.size == 0, f"{name}: {bad.size} mismatches, first {bad[:5]}"


def _specs():
    from sph_hip.workloads import mixed_scenes
    # small counts keep the oracle fast; boxes span dense (bricks) and sparse (per lane)
    return mixed_scenes(count=10, seed=100, n_min=300, n_max=6000, extent=(2.0, 12.0))


@pytest.mark.parametrize("kernel", ["auto", "direct", "brick", "tile"])
def test_batch_matches_oracle_per_scene(gpu, orc, kernel):
    specs = _specs()
    params = [s.params() for s in specs]
    states = [s.state() for s in specs]
    print(f"batch kernel={kernel} counts={[s.n for s in specs]}", flush=True)
    with gpu.Batch(params, [s.n for s in specs]) as b:
        b.set_kernel(kernel)
        for i, (p, v) in enumerate(states):
            b.upload(i, p, v)
        if kernel == "auto":
            choice = [b.uses_bricks(i) for i in range(len(specs))]
            print("bricks per scene:", choice)
        b.step(2)
        got = [b.download(i) for i in range(len(specs))]
    for i, (spec, (p, v)) in enumerate(zip(specs, states)):
        rp, rv = orc.step(params[i], p, v, 2)
        _bitexact(f"scene {i} pos", got[i][0], rp)
        _bitexact(f"scene {i} vel", got[i][1], rv)


def test_batch_intermediates_per_scene(gpu, orc):
    specs = _specs()[:4]
    params = []
    for s in specs:
        p = s.params()
        p.substeps = 1
        params.append(p)
    states = [s.state() for s in specs]
    with gpu.Batch(params, [s.n for s in specs]) as b:
        b.set_debug(True)
        for i, (p, v) in enumerate(states):
            b.upload(i, p, v)
        b.step(1)
        for i in range(len(specs)):
            _, _, ref = orc.substep(params[i], *states[i])
            for stage in ("cell_keys", "sorted_keys", "sorted_index", "start", "end", "pred_pos",
                          "pred_vel", "density", "near_density", "vel_force"):
                got = b.dump(i, stage)
                if got.dtype.kind == "f":
                    _bitexact(f"scene {i} {stage}", got, ref[stage])
                else:
                    assert np.array_equal(got, ref[stage]), f"scene {i} {stage}"


def test_scene_alone_equals_scene_in_mix(gpu):
    specs = _specs()
    params = [s.params() for s in specs]
    states = [s.state() for s in specs]
    k = 3
    with gpu.World(params[k], specs[k].n) as w:
        w.upload(*states[k])
        w.step(3)
        alone = w.download()
    with gpu.Batch(params, [s.n for s in specs]) as b:
        for i, (p, v) in enumerate(states):
            b.upload(i, p, v)
        b.step(3)
        mixed = b.download(k)
    _bitexact("pos", mixed[0], alone[0])
    _bitexact("vel", mixed[1], alone[1])


@pytest.mark.parametrize("kernel", ["auto", "tile"])
def test_full_mixed_batch_one_frame(gpu, orc, kernel):
    """Full size of the mixed workload: its 256 scenes (5.4M particles, 47 of them packed
    above their target density, whose particles record up to ~200 neighbours) stepped one
    frame together, through the autotuned kernels and through tiles (every scene tiled); every
    scene's state is bit-exact against the oracle run on it alone."""
    from sph_hip.workloads import mixed_scenes
    specs = mixed_scenes()
    params = [s.params() for s in specs]
    states = [s.state() for s in specs]
    print(f"full mixed batch: {len(specs)} scenes, {sum(s.n for s in specs)} particles, kernel={kernel}",
          flush=True)
    with gpu.Batch(params, [s.n for s in specs]) as b:
        if kernel == "tile":
            b.set_tile_config(0)
        for i, (p, v) in enumerate(states):
            b.upload(i, p, v)
        b.step(1)
        print("batch kernels:", b.kernel, flush=True)
        got = [b.download(i) for i in range(len(specs))]
    print("oracle", flush=True)
    for i, (p, v) in enumerate(states):
        rp, rv = orc.step(params[i], p, v, 1)
        _bitexact(f"scene {i} pos", got[i][0], rp)
        _bitexact(f"scene {i} vel", got[i][1], rv)


def shifted(gpu, params, pos, offset):
    """The same scene translated by `offset`: a box whose minimum is non-zero (negative on
    some axes, not a multiple of R), positions moved with it (rounded to float32 and kept
    inside the new box).  Cell coordinates floor((p - min) / R) round differently here than
    at the origin, which exercises the sample-point bounds skip and the bucket of each run."""
    p2 = gpu.SphParams.from_buffer_copy(params)
    for a in range(3):
        p2.min_xyz[a] = float(np.float32(offset[a]))
        p2.max_xyz[a] = float(np.float32(np.float32(params.max_xyz[a]) + np.float32(offset[a])))
    q = pos.copy()
    lo = np.array(p2.min_xyz, np.float32)
    hi = np.array(p2.max_xyz, np.float32)
    q[:, :3] = np.clip((pos[:, :3] + np.asarray(offset, np.float32)).astype(np.float32), lo, hi)
    return p2, q


OFFSETS = [(-7.31, 2.77, -0.123), (13.5, -41.05, 0.6), (-0.35, -0.7, -3.3), (1.0e3, 5.25, -2.0e2)]


@pytest.mark.parametrize("kernel", ["direct", "brick", "tile"])
def test_batch_scenes_with_offset_boxes(gpu, orc, kernel):
    """Scenes whose boxes start away from the origin (each its own asymmetric offset), two
    frames in one batch: every scene bit-exact against the oracle run on it alone."""
    specs = _specs()[:len(OFFSETS)]
    params, states = [], []
    for sp, off in zip(specs, OFFSETS):
        pos, vel = sp.state()
        p2, q = shifted(gpu, sp.params(), pos, off)
        params.append(p2)
        states.append((q, vel))
    with gpu.Batch(params, [s.n for s in specs]) as b:
        b.set_kernel(kernel)
        for i, (p, v) in enumerate(states):
            b.upload(i, p, v)
        b.step(2)
        got = [b.download(i) for i in range(len(specs))]
    for i, (p, v) in enumerate(states):
        rp, rv = orc.step(params[i], p, v, 2)
        _bitexact(f"scene {i} pos", got[i][0], rp)
        _bitexact(f"scene {i} vel", got[i][1], rv)
