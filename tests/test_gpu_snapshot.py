"""Overlapped readback (sph_world_snapshot_async / _wait): a snapshot taken after a frame and
copied to the host while the next frames step holds that frame's state, bit for bit the
oracle's, and back-to-back snapshots wait for each other."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bitexact(name, got, ref):
    bad = np.flatnonzero(got.view(np.uint32).ravel() != ref.view(np.uint32).ravel())
    assert bad.size == 0, f"{name}: {bad.size} mismatches, first {bad[:5]}"


def test_snapshot_overlaps_the_next_steps(gpu, orc):
    from sph_hip.workloads import DAM_BREAK_16K as W
    p = W.params()
    pos, vel = W.state()
    r1 = orc.step(p, pos, vel, 1)
    r3 = orc.step(p, r1[0], r1[1], 2)
    print(f"snapshot n={W.n}", flush=True)
    with gpu.World(p, W.n) as w:
        w.upload(pos, vel)
        w.step(1)
        sp = np.empty((W.n, 4), np.float32)
        sv = np.empty_like(sp)
        w.snapshot_async(sp, sv)
        w.step(2)  # enqueued while the snapshot of frame 1 is being copied
        w.snapshot_wait()
        _bitexact("snapshot pos", sp, r1[0])
        _bitexact("snapshot vel", sv, r1[1])
        a = np.empty_like(sp)
        b = np.empty_like(sp)
        w.snapshot_async(a, b)
        w.snapshot_async(sp, sv)  # waits for the first
        w.snapshot_wait()
        gp, gv = w.download()
    for name, got in (("pos", gp), ("first pos", a), ("second pos", sp)):
        _bitexact(name, got, r3[0])
    for name, got in (("vel", gv), ("first vel", b), ("second vel", sv)):
        _bitexact(name, got, r3[1])
