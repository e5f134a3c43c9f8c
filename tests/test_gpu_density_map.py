"""Density map (updateDensityMap.comp) on the GPU: bit for bit the FP32 oracle's map of the
positions and cell table the last substep left, for a single world on either kernel
family, for the 1M-particle scene, and for scenes of a batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bitexact(name, got, ref):
    assert got.shape == ref.shape, (got.shape, ref.shape)
    bad = np.flatnonzero(got.view(np.uint32).ravel() != ref.view(np.uint32).ravel())
    assert bad.size == 0, f"{name}: {bad.size} mismatches, first {bad[:5]}"


def _oracle_map(orc, params, pos, vel, frames, voxel):
    """frames * substeps oracle substeps; the map of the last one's positions and table."""
    total = frames * params.substeps
    if frames > 1:
        pos, vel = orc.step(params, pos, vel, frames - 1)
    for _ in range(params.substeps - 1):
        pos, vel, _ = orc.substep(params, pos, vel)
    pos, vel, inter = orc.substep(params, pos, vel)
    assert total >= 1
    return orc.density_map(params, pos, inter["start"], inter["end"], voxel)


@pytest.mark.parametrize("kernel", ["direct", "brick"])
def test_world_density_map_dam_break(gpu, orc, kernel):
    from sph_hip.workloads import DAM_BREAK_16K as W
    p = W.params()
    pos, vel = W.state()
    print(f"density map dam_break_16k kernel={kernel}", flush=True)
    with gpu.World(p, W.n) as w:
        w.set_kernel(kernel)
        w.upload(pos, vel)
        with pytest.raises(gpu.SphError):
            w.density_map(0.1)  # no substep since the upload
        w.step(2)
        got = w.density_map(0.1)
    ref = _oracle_map(orc, p, pos, vel, 2, 0.1)
    assert got.shape == (60, 60, 100)
    assert np.count_nonzero(ref) > 1000
    _bitexact("density map", got, ref)


def test_world_density_map_device_matches_host(gpu):
    """The device-memory entry point writes the same map, and nothing past it, into memory
    the caller owns (allocated here through the HIP runtime the library runs on), and
    refuses a buffer that is too small."""
    import ctypes as C
    from sph_hip.workloads import DAM_BREAK_16K as W
    hip = C.CDLL("libamdhip64.so.7")
    p = W.params()
    pos, vel = W.state()
    print("density map device dam_break_16k", flush=True)
    with gpu.World(p, W.n) as w:
        w.upload(pos, vel)
        w.step(2)
        host = w.density_map(0.1)
        count = host.size + 64
        dev = C.c_void_p()
        assert hip.hipMalloc(C.byref(dev), C.c_size_t(count * 4)) == 0
        try:
            assert hip.hipMemset(dev, 0xFF, C.c_size_t(count * 4)) == 0  # NaN pattern
            assert hip.hipDeviceSynchronize() == 0
            w.density_map_device(0.1, dev.value, count)
            w.sync()
            got = np.empty(count, np.float32)
            assert hip.hipMemcpy(C.c_void_p(got.ctypes.data), dev, C.c_size_t(count * 4), 2) == 0
            with pytest.raises(gpu.SphError):
                w.density_map_device(0.1, dev.value, host.size - 1)
        finally:
            hip.hipFree(dev)
    _bitexact("device map", got[:host.size].reshape(host.shape), host)
    assert np.all(got[host.size:].view(np.uint32) == 0xFFFFFFFF)


def test_world_density_map_uniform_1m(gpu, orc):
    from sph_hip.workloads import UNIFORM_1M as W
    p = W.params(table_cells=1 << 21)
    pos, vel = W.state()
    print("density map uniform_1m T=2^21", flush=True)
    with gpu.World(p, W.n) as w:
        w.upload(pos, vel)
        w.step(1)
        got = w.density_map(0.5)
    ref = _oracle_map(orc, p, pos, vel, 1, 0.5)
    _bitexact("density map", got, ref)


def test_batch_density_map_per_scene(gpu, orc):
    from sph_hip.workloads import DAM_BREAK_16K, mixed_scenes
    specs = mixed_scenes(count=3, seed=200, n_min=500, n_max=4000, extent=(2.0, 6.0))
    params = [DAM_BREAK_16K.params()] + [s.params() for s in specs]
    states = [DAM_BREAK_16K.state()] + [s.state() for s in specs]
    counts = [DAM_BREAK_16K.n] + [s.n for s in specs]
    print(f"density map batch counts={counts}", flush=True)
    with gpu.Batch(params, counts) as b:
        for i, (q, v) in enumerate(states):
            b.upload(i, q, v)
        b.step(1)
        got = [b.density_map(i, 0.2) for i in range(len(counts))]
    for i, (q, v) in enumerate(states):
        _bitexact(f"scene {i}", got[i], _oracle_map(orc, params[i], q, v, 1, 0.2))


@pytest.mark.parametrize("kernel", ["direct", "brick"])
def test_world_density_map_after_box_change(gpu, orc, kernel):
    """set_params may swap two box extents (same cell, brick and bucket totals, different
    grid dimensions); a map taken before the next substep reads the last substep's runs,
    laid out for the old dimensions, under the new box."""
    from sph_hip.workloads import DAM_BREAK_16K as W
    pa = W.params()
    pb = gpu.SphParams.from_buffer_copy(pa)
    pb.max_xyz[0], pb.max_xyz[1] = pa.max_xyz[1], pa.max_xyz[0]  # 10 x 6 x 6 -> 6 x 10 x 6
    pos, vel = W.state()
    print(f"density map after a box change kernel={kernel}", flush=True)
    with gpu.World(pa, W.n) as w:
        w.set_kernel(kernel)
        w.upload(pos, vel)
        w.step(1)
        w.set_params(pb)
        got = w.density_map(0.2)
    for _ in range(pa.substeps - 1):
        pos, vel, _ = orc.substep(pa, pos, vel)
    pos, vel, inter = orc.substep(pa, pos, vel)
    ref = orc.density_map(pb, pos, inter["start"], inter["end"], 0.2)
    assert np.count_nonzero(ref) > 1000
    _bitexact("density map", got, ref)
