"""Host-side argument checks of the Python wrappers (no GPU needed): a call
whose counts need more bytes than a buffer holds is refused with
ERR_INVALID_INPUT before the C ABI is entered (the ABI would read past the
buffer: a host over-read, or a device memory fault for device buffers)."""
import ctypes

import numpy as np
import pytest

import zkmetal_amd as zk


class _FakeDevice(zk.DeviceBuffer):
    """A DeviceBuffer that only knows its size (nothing is allocated)."""

    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        self.nbytes = nbytes

    def free(self):
        pass


def _engine():
    e = zk.MetalMSM.__new__(zk.MetalMSM)  # no device: only the host checks run
    e.handle = ctypes.c_void_p()
    return e


def _refused(fn):
    with pytest.raises(zk.MSMError) as ei:
        fn()
    assert ei.value.status == zk.ERR_INVALID_INPUT


def test_short_host_buffers_refused():
    e = _engine()
    pts, sc = np.zeros(10 * 64, np.uint8), np.zeros(10 * 32, np.uint8)
    _refused(lambda: e.msm_batch(pts, sc, [4, 7]))          # 11 points, 10 given
    _refused(lambda: e.msm_batch(pts, sc[:-1], [4, 6]))     # one scalar byte short
    _refused(lambda: e.grumpkin_msm_batch(pts, sc, [11]))
    _refused(lambda: e.g2_msm_batch(np.zeros(5 * 128, np.uint8), sc, [6]))
    _refused(lambda: e.msm_batch_resident(sc, [3, 3, 5]))
    _refused(lambda: e.kzg_commit(sc, [11]))
    _refused(lambda: e.msm_small(pts, np.zeros(19, np.uint8), 2))  # 10 x u16 needs 20 B
    _refused(lambda: e.msm_small(pts, sc, 3))                     # widths 1, 2, 4 only


def test_short_device_buffers_refused():
    e = _engine()
    dp, ds = _FakeDevice(100 * 64), _FakeDevice(100 * 32)
    _refused(lambda: e.msm_batch_device(dp, ds, [50, 51]))
    _refused(lambda: e.msm_batch_device(dp, _FakeDevice(99 * 32), [100]))
    _refused(lambda: e.grumpkin_msm_batch_device(dp, ds, [101]))
    _refused(lambda: e.g2_msm_batch_device(_FakeDevice(99 * 128), ds, [100]))
    _refused(lambda: e.msm_batch_resident(ds, [101]))
    _refused(lambda: e.kzg_commit(ds, [60, 41]))


def test_counts_out_of_range_refused():
    e = _engine()
    pts, sc = np.zeros(64, np.uint8), np.zeros(32, np.uint8)
    _refused(lambda: e.msm_batch(pts, sc, [-1]))
    _refused(lambda: e.msm_batch(pts, sc, [1 << 32]))
