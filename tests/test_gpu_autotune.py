"""Autotuner: the measurement leaves the state untouched and whichever variant it picks
stays bit-exact with the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_keeps_state_and_parity(gpu, orc):
    from sph_hip.workloads import DAM_BREAK_16K as W
    params = W.params()
    pos, vel = W.state()
    with gpu.World(params, W.n) as w:
        w.upload(pos, vel)
        before = w.download()
        chosen, ms = w.autotune(2)
        after = w.download()
        assert chosen in ("direct", "brick", "tile") and ms > 0
        assert w.kernel == chosen
        assert np.array_equal(before[0].view(np.uint32), after[0].view(np.uint32))
        assert np.array_equal(before[1].view(np.uint32), after[1].view(np.uint32))
        w.step(2)
        gp, gv = w.download()
    rp, rv = orc.step(params, pos, vel, 2)
    assert np.array_equal(gp.view(np.uint32), rp.view(np.uint32))
    assert np.array_equal(gv.view(np.uint32), rv.view(np.uint32))


def test_auto_world_steps_like_oracle(gpu, orc):
    from sph_hip.workloads import DAM_BREAK_16K as W
    params = W.params()
    pos, vel = W.state(6000)
    with gpu.World(params, 6000) as w:  # AUTO: tunes on the first step
        w.upload(pos, vel)
        w.step(3)
        gp, gv = w.download()
    rp, rv = orc.step(params, pos, vel, 3)
    assert np.array_equal(gp.view(np.uint32), rp.view(np.uint32))
    assert np.array_equal(gv.view(np.uint32), rv.view(np.uint32))


def test_online_tuning_runs_real_substeps(gpu, orc):
    """An AUTO world of a size bucket no other test uses tunes over its own first substeps
    (each configuration twice, then the radix widths), then steps through the chosen
    configuration's graph: 30 frames (90 substeps) bit-exact against the oracle."""
    from sph_hip.workloads import DAM_BREAK_16K as W
    params = W.params(table_cells=1 << 17)
    pos, vel = W.state(12345)
    with gpu.World(params, 12345) as w:
        w.upload(pos, vel)
        w.step(2)    # 6 substeps: still tuning
        w.step(28)   # the rest of the candidates, then the graph of the fastest
        gp, gv = w.download()
        cfg = w.kernel_config
    print("chosen:", cfg)
    assert "radix digits" in cfg
    rp, rv = orc.step(params, pos, vel, 30)
    assert np.array_equal(gp.view(np.uint32), rp.view(np.uint32))
    assert np.array_equal(gv.view(np.uint32), rv.view(np.uint32))
