"""GPU tests of two edge paths of the drop-in ABI:

* non-canonical base coordinates: a Montgomery coordinate given as
  x + k p (any 256-bit representative of the same residue) is the same
  point -- the reference's inputs are only required to be below 2^256
  (docs/metal-tuning-guide.md:267-282 discusses non-canonical inputs); the
  MSM through zkmetal_bn254_msm equals the closed form of the canonical bases;
* ZKMETAL_ERR_ALLOC_FAILED: the reference defines the code
  (Sources/zkMetal-ffi/zkMetal_ffi.swift:14) but never returns it; here an
  impossible device allocation and an engine held under a workspace cap
  return it from the host side, and the engine stays usable afterwards."""
import random

import numpy as np
import pytest

import zkmetal_amd as zk
from oracle import bn254 as B
from oracle import ref
from helpers import result_affine

pytestmark = pytest.mark.gpu


def _closed(ks, sc, n):
    return B.mul(B.G, ref.scalar_dot(sc.tobytes(), ks[:32 * n], n))


def test_non_canonical_base_coordinates():
    n, seed = 3000, 0xC0DE
    gen = zk.MetalMSM()
    try:
        gen.srs_generate(seed, n)
        bases = gen.srs_download(n)
    finally:
        gen.close()
    rnd = random.Random(seed)
    pts = bytearray(bases.tobytes())
    changed = 0
    for i in range(n):
        for off in (0, 32):  # x, then y
            v = int.from_bytes(pts[64 * i + off:64 * i + off + 32], "little")
            if v == 0:
                continue  # keep (0,0) identities as they are
            k = rnd.randrange(0, 6)
            if v + k * B.P < 1 << 256:
                pts[64 * i + off:64 * i + off + 32] = (v + k * B.P).to_bytes(32, "little")
                changed += k > 0
    assert changed > n // 2
    sc = zk.gen_scalars(seed + 1, n, 0)
    ks = zk.gen_scalars(seed, n, 0).tobytes()
    eng = zk.MetalMSM()  # no resident SRS: the bases go through the table build
    try:
        got = eng.msm(bytes(pts), sc)
        assert result_affine(got) == _closed(ks, sc, n)
        assert got == eng.msm(bases.tobytes(), sc)  # same bytes as the canonical input
    finally:
        eng.close()


def test_impossible_device_allocation_is_alloc_failed():
    with pytest.raises(zk.MSMError) as ei:
        zk.DeviceBuffer(1 << 60)
    assert ei.value.status == zk.ERR_ALLOC_FAILED


def test_workspace_cap_gives_alloc_failed_and_the_engine_recovers():
    n, seed = 1 << 14, 0xA11C
    eng = zk.MetalMSM()
    try:
        eng.set_workspace_limit(1 << 20)
        with pytest.raises(zk.MSMError) as ei:
            eng.srs_generate(seed, n)  # the 4 MiB table does not fit
        assert ei.value.status == zk.ERR_ALLOC_FAILED
        assert eng.workspace_bytes() <= 1 << 20
        eng.set_workspace_limit(0)
        eng.srs_generate(seed, n)
        sc = zk.gen_scalars(seed + 1, n, 1)
        want = _closed(zk.gen_scalars(seed, n, 0).tobytes(), sc, n)
        assert result_affine(eng.msm_batch_resident(sc, [n])[0]) == want
        # a cap below what the call's workspace needs: refused, then fine again
        eng.set_workspace_limit(eng.workspace_bytes() + (64 << 10))
        sc2 = zk.gen_scalars(seed + 2, n, 0)
        ds = zk.DeviceBuffer(32 * 8 * n)
        try:
            ds.upload(np.concatenate([sc2] * 8))
            with pytest.raises(zk.MSMError) as ei:
                eng.msm_batch_resident(ds, [n] * 8)  # eight MSMs need more buckets
            assert ei.value.status == zk.ERR_ALLOC_FAILED
            eng.set_workspace_limit(0)
            r8 = eng.msm_batch_resident(ds, [n] * 8)
        finally:
            ds.free()
        want2 = _closed(zk.gen_scalars(seed, n, 0).tobytes(), sc2, n)
        assert all(result_affine(r) == want2 for r in r8)
        assert result_affine(eng.msm_batch_resident(sc, [n])[0]) == want
    finally:
        eng.close()
