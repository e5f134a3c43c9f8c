"""Radix digit widths (8, 11, 12 bits) against the FP32 oracle: the stable sort of
(cell key, input index) pairs and every later stage of one substep are bit-exact for each
width at 16-, 21-, 24- and 25-bit tables (one to four passes), at the full size of the 1M
scene, for a table of 5 cells (one partial pass) and over a multi-frame trajectory.  The
reference sorts with 4 x 8-bit passes (commands.c:106-161); any stable LSD width gives the
same order."""
import numpy as np
import pytest

from test_gpu_parity import STAGES_F32, STAGES_INT, _assert_bitexact

pytestmark = pytest.mark.gpu


def _substep(gpu, orc, params, pos, vel, bits):
    p1 = gpu.SphParams.from_buffer_copy(params)
    p1.substeps = 1
    with gpu.World(p1, pos.shape[0]) as w:
        w.set_lane_config(3)
        w.set_radix_bits(bits)
        assert f"{bits}-bit radix digits" in w.kernel_config, w.kernel_config
        w.set_debug(True)
        w.upload(pos, vel)
        w.step(1)
        got = {s: w.dump(s) for s in STAGES_INT + STAGES_F32}
        got["pos"], got["vel"] = w.download()
    rpos, rvel, ref = orc.substep(p1, pos, vel)
    for s in STAGES_INT + STAGES_F32:
        _assert_bitexact(f"{bits}-bit:{s}", got[s], ref[s])
    _assert_bitexact(f"{bits}-bit:pos", got["pos"], rpos)
    _assert_bitexact(f"{bits}-bit:vel", got["vel"], rvel)


@pytest.mark.parametrize("bits", [8, 11, 12])
@pytest.mark.parametrize("table_bits", [16, 21, 24, 25])
def test_widths_and_table_sizes(gpu, orc, bits, table_bits):
    from sph_hip.workloads import DAM_BREAK_16K as W
    print(f"radix {bits}-bit, T = 2^{table_bits}", flush=True)
    _substep(gpu, orc, W.params(table_cells=1 << table_bits), *W.state(20000), bits)


@pytest.mark.parametrize("bits", [11, 12])
def test_widths_uniform_1m(gpu, orc, bits):
    from sph_hip.workloads import UNIFORM_1M as W
    print(f"radix {bits}-bit, uniform_1m", flush=True)
    _substep(gpu, orc, W.params(table_cells=1 << 21), *W.state(), bits)


def test_width_tiny_table(gpu, orc):
    from sph_hip.workloads import DAM_BREAK_16K as W
    _substep(gpu, orc, W.params(table_cells=5), *W.state(3000), 12)


@pytest.mark.parametrize("bits", [11, 12])
def test_width_trajectory(gpu, orc, bits):
    from sph_hip.workloads import DAM_BREAK_16K as W
    params = W.params(table_cells=1 << 21)
    pos, vel = W.state()
    with gpu.World(params, W.n) as w:
        w.set_radix_bits(bits)
        w.upload(pos, vel)
        w.step(5)
        gp, gv = w.download()
    rp, rv = orc.step(params, pos, vel, 5)
    _assert_bitexact("pos", gp, rp)
    _assert_bitexact("vel", gv, rv)


def test_width_rejected_without_look_back_memory(gpu):
    from sph_hip.workloads import DAM_BREAK_16K as W
    with gpu.World(W.params(), W.n) as w:
        with pytest.raises(gpu.SphError):
            w.set_radix_bits(10)
