"""Device-resident state (the reference's renderer draws from the device buffers it steps,
commands.c:521-525): a world steps buffers the caller allocated on the device, filled and
read back with plain hipMemcpy, bit-exact against the FP32 oracle; the world's own buffers
come back after unbinding."""
import ctypes as C

import numpy as np
import pytest

from test_gpu_parity import _assert_bitexact

pytestmark = pytest.mark.gpu


class Hip:
    """The few HIP runtime calls a renderer-side caller makes (the runtime the library runs on)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        return p.value

    def free(self, p):
        assert self.lib.hipFree(C.c_void_p(p)) == 0

    def h2d(self, dst, a):
        assert self.lib.hipMemcpy(C.c_void_p(dst), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def d2h(self, a, src):
        assert self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(src), C.c_size_t(a.nbytes), 2) == 0

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0


def test_world_steps_caller_device_buffers(gpu, orc):
    from sph_hip.workloads import DAM_BREAK_16K as W
    params = W.params()
    pos, vel = W.state()
    hip = Hip()
    dpos, dvel = hip.malloc(pos.nbytes), hip.malloc(vel.nbytes)
    try:
        with gpu.World(params, W.n) as w:
            own = w.device_state()
            assert own[0] and own[1] and own != (dpos, dvel)
            w.bind_state(dpos, dvel)
            assert w.device_state() == (dpos, dvel)
            hip.h2d(dpos, pos)     # the caller fills its buffers itself
            hip.h2d(dvel, vel)
            w.step(3)
            w.sync()
            gp = np.empty_like(pos)
            gv = np.empty_like(vel)
            hip.d2h(gp, dpos)      # and draws / reads them without the library
            hip.d2h(gv, dvel)
            rp, rv = orc.step(params, pos, vel, 3)
            _assert_bitexact("bound pos", gp, rp)
            _assert_bitexact("bound vel", gv, rv)
            # the library's own download reads the bound buffers too
            lp, lv = w.download()
            _assert_bitexact("download pos", lp, rp)
            # unbinding returns to the world's own buffers, which an upload refills
            w.bind_state(None, None)
            assert w.device_state() == own
            w.upload(pos, vel)
            w.step(3)
            op, ov = w.download()
            _assert_bitexact("own pos", op, rp)
            _assert_bitexact("own vel", ov, rv)
            # and the caller's buffers were left as the bound steps wrote them
            hip.d2h(gp, dpos)
            _assert_bitexact("caller buffer untouched", gp, rp)
            with pytest.raises(gpu.SphError):
                w.bind_state(dpos, None)
            with pytest.raises(gpu.SphError):
                w.bind_state(dpos + 4, dvel)   # not 16-byte aligned
            host = np.zeros((W.n, 4), np.float32)
            with pytest.raises(gpu.SphError):
                w.bind_state(host.ctypes.data, dvel)  # host memory
    finally:
        hip.sync()
        hip.free(dpos)
        hip.free(dvel)
