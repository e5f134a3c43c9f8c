"""GPU parity of the near-pressure force term (applyPressureAndViscosityForce.comp:253-261,
316-321; its multiplier is NEAR_PRESSURE_MULTIPLIER, main.c:22, 0 in the reference's own
scene).  Every other parity case runs at k_near = 0, where the term vanishes; here it runs
at k_near = 50 on the dam break and the 1M fill, through every stage of one substep on both
kernel families, through every per-lane launch configuration, and over multi-frame
trajectories, bit-exact against the FP32 oracle."""
import dataclasses

import numpy as np
import pytest

from test_gpu_parity import STAGES_F32, STAGES_INT, _assert_bitexact, _check_substep, _fp64_errors

pytestmark = pytest.mark.gpu

K_NEAR = 50.0


def _near(w):
    return dataclasses.replace(w, near_pressure_multiplier=K_NEAR)


def test_near_pressure_changes_the_step(orc):
    """The term is live at k_near = 50: the oracle's force pass differs from k_near = 0."""
    from sph_hip.workloads import DAM_BREAK_16K
    W = _near(DAM_BREAK_16K)
    pos, vel = W.state(4096)
    p = W.params()
    p.substeps = 1
    _, _, on = orc.substep(p, pos, vel)
    p.near_pressure_multiplier = 0.0
    _, _, off = orc.substep(p, pos, vel)
    assert np.abs(on["vel_force"] - off["vel_force"]).max() > 1e-3


def test_near_pressure_dam_break_one_substep(gpu, orc):
    from sph_hip.workloads import DAM_BREAK_16K
    W = _near(DAM_BREAK_16K)
    params = W.params()
    pos, vel = W.state()
    got, ref = _check_substep(gpu, orc, params, pos, vel)
    err = _fp64_errors(gpu, orc, params, pos, vel, got, ref)
    for name, e in err.items():
        assert e.max() <= 1e-5, (name, e.max())


def test_near_pressure_uniform_1m_one_substep(gpu, orc):
    from sph_hip.workloads import UNIFORM_1M
    W = _near(UNIFORM_1M)
    params = W.params(table_cells=1 << 21)
    pos, vel = W.state()
    _check_substep(gpu, orc, params, pos, vel)


@pytest.mark.parametrize("config", range(10))
def test_near_pressure_every_lane_config(gpu, orc, config):
    """Every per-lane launch configuration, 3 frames x 3 substeps of the dam break."""
    from sph_hip.workloads import DAM_BREAK_16K
    W = _near(DAM_BREAK_16K)
    params = W.params()
    pos, vel = W.state()
    with gpu.World(params, W.n) as w:
        w.set_lane_config(config)
        print(f"near pressure, lane config {config}: {w.kernel_config}", flush=True)
        w.upload(pos, vel)
        w.step(3)
        gp, gv = w.download()
    rp, rv = orc.step(params, pos, vel, 3)
    _assert_bitexact("pos", gp, rp)
    _assert_bitexact("vel", gv, rv)


@pytest.mark.parametrize("kernel", ["direct", "brick", "tile"])
def test_near_pressure_dam_break_ten_frames(gpu, orc, kernel):
    from sph_hip.workloads import DAM_BREAK_16K
    W = _near(DAM_BREAK_16K)
    params = W.params()
    pos, vel = W.state()
    with gpu.World(params, W.n) as w:
        w.set_kernel(kernel)
        w.upload(pos, vel)
        w.step(10)
        gp, gv = w.download()
    rp, rv = orc.step(params, pos, vel, 10)
    _assert_bitexact("pos", gp, rp)
    _assert_bitexact("vel", gv, rv)


def test_near_pressure_uniform_1m_three_frames(gpu, orc):
    from sph_hip.workloads import UNIFORM_1M
    W = _near(UNIFORM_1M)
    params = W.params(table_cells=1 << 21)
    pos, vel = W.state()
    print(f"case near pressure n={W.n} frames=3", flush=True)
    with gpu.World(params, W.n) as w:
        w.upload(pos, vel)
        w.step(3)
        gp, gv = w.download()
    rp, rv = orc.step(params, pos, vel, 3)
    _assert_bitexact("pos", gp, rp)
    _assert_bitexact("vel", gv, rv)


def test_near_pressure_batch(gpu, orc):
    """A batch of three scenes at k_near = 50 with different boxes: each scene bit-exact
    against the oracle run alone (the batch reads k_near from each scene's constants)."""
    from sph_hip.workloads import DAM_BREAK_16K
    W = _near(DAM_BREAK_16K)
    specs = [(3000, (10.0, 6.0, 6.0)), (5000, (8.0, 5.0, 7.0)), (2500, (6.0, 6.0, 4.0))]
    params, states = [], []
    for i, (n, box) in enumerate(specs):
        p = gpu.params_for(particle_radius=W.particle_radius, smoothing_radius=W.smoothing_radius,
                           g=W.g, box_min=(0, 0, 0), box_max=box, frame_dt=W.frame_dt,
                           substeps=W.substeps, pressure_multiplier=W.pressure_multiplier,
                           target_density=W.target_density, near_pressure_multiplier=K_NEAR,
                           table_cells=1 << max(0, (n - 1).bit_length()))
        params.append(p)
        states.append(gpu.fill_block(n, 40 + i, (0, 0, 0), (box[0] / 2, box[1], box[2] / 2), 0.0))
    with gpu.Batch(params, [s[0].shape[0] for s in states]) as b:
        for i, (pp, vv) in enumerate(states):
            b.upload(i, pp, vv)
        b.step(2)
        got = [b.download(i) for i in range(len(specs))]
    for i, ((pp, vv), p) in enumerate(zip(states, params)):
        rp, rv = orc.step(p, pp, vv, 2)
        _assert_bitexact(f"scene {i} pos", got[i][0], rp)
        _assert_bitexact(f"scene {i} vel", got[i][1], rv)
