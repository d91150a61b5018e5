"""GPU test of zkmsm_msm_batch_resident_device_noalloc, the entry point that
allocates nothing (its buffers come from a fixed per-device array): through
the C ABI, two calls in a row (the second smaller) give the same bytes as
zkmsm_msm_batch_resident_device, the engine's own device memory does not
grow during them, and a call beyond the array's limits is refused."""
import ctypes

import numpy as np
import pytest

import zkmetal_amd as zk
from oracle import bn254 as B
from oracle import ref
from helpers import result_affine

pytestmark = pytest.mark.gpu

SEED = 0x0A11
N_SRS = 1 << 17


def _call(fn, eng, dsc, counts):
    c = np.ascontiguousarray(np.asarray(counts, dtype=np.uint32))
    out = np.zeros(len(counts) * 96, dtype=np.uint8)
    rc = fn(eng.handle, dsc.ptr, ctypes.c_void_p(c.ctypes.data), len(counts),
            ctypes.c_void_p(out.ctypes.data))
    return rc, [out[i * 96:(i + 1) * 96].tobytes() for i in range(len(counts))]


def test_noalloc_entry_matches_the_allocating_one():
    L = zk.lib()
    eng = zk.MetalMSM()
    first = [20000, 30000, 7, 0, 1]
    second = [1000, 5]
    dsc = zk.DeviceBuffer(32 * N_SRS)
    try:
        print("noalloc: srs 2^17, calls", first, second, flush=True)
        eng.srs_generate(SEED, N_SRS)
        sc = zk.gen_scalars(SEED + 1, N_SRS, 1)
        dsc.upload(sc)
        held = eng.workspace_bytes()
        rc1, na1 = _call(L.zkmsm_msm_batch_resident_device_noalloc, eng, dsc, first)
        rc2, na2 = _call(L.zkmsm_msm_batch_resident_device_noalloc, eng, dsc, second)
        assert rc1 == zk.SUCCESS and rc2 == zk.SUCCESS
        assert eng.workspace_bytes() == held  # no engine buffer grew
        # beyond the limits: refused, nothing run
        rc3, _ = _call(L.zkmsm_msm_batch_resident_device_noalloc, eng, dsc, [70000])
        assert rc3 == zk.ERR_INVALID_INPUT
        rc4, _ = _call(L.zkmsm_msm_batch_resident_device_noalloc, eng, dsc, [1] * 65)
        assert rc4 == zk.ERR_INVALID_INPUT
        # the allocating entry point on the same inputs
        rca, al1 = _call(L.zkmsm_msm_batch_resident_device, eng, dsc, first)
        rcb, al2 = _call(L.zkmsm_msm_batch_resident_device, eng, dsc, second)
        assert rca == zk.SUCCESS and rcb == zk.SUCCESS
        assert na1 == al1 and na2 == al2
        assert na1[3] == bytes(96)
        ks = zk.gen_scalars(SEED, N_SRS, 0).tobytes()
        off = first[0]
        want = B.mul(B.G, ref.scalar_dot(sc[off:off + first[1]].tobytes(), ks[:32 * first[1]], first[1]))
        assert result_affine(na1[1]) == want
        # the noalloc result after other work is unchanged (the array keeps
        # the last call's contents; every call clears what must start at zero)
        rc5, na5 = _call(L.zkmsm_msm_batch_resident_device_noalloc, eng, dsc, second)
        assert rc5 == zk.SUCCESS and na5 == na2
    finally:
        dsc.free()
        eng.close()
