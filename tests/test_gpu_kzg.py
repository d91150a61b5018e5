"""GPU tests of KZG commitments with Montgomery-form Fr coefficients
(zkmsm_kzg_commit, KZGEngine.swift:131-179): the on-device conversion must
give exactly the commitment of the standard-form coefficients."""
import random

import numpy as np
import pytest

import zkmetal_amd as zk
from helpers import closed_form, result_affine, scalars_to_ints
from oracle import bn254 as B

pytestmark = pytest.mark.gpu

SEED = 0x4B5A
N_SRS = 1 << 14
R256 = (1 << 256) % B.R_ORDER


def to_mont_bytes(xs, lift=None):
    """x -> x * 2^256 mod r as 32 LE bytes; lift[i] adds that many r (still
    < 2^256) to test non-canonical Montgomery inputs"""
    out = bytearray()
    for i, x in enumerate(xs):
        m = x * R256 % B.R_ORDER
        if lift is not None:
            m += lift[i] * B.R_ORDER
            assert m < 2**256
        out += m.to_bytes(32, "little")
    return bytes(out)


@pytest.fixture(scope="module")
def srs():
    e = zk.MetalMSM()
    e.srs_generate(SEED, N_SRS)
    yield e
    e.close()


def test_commit_equals_standard_msm(srs):
    rnd = random.Random(1)
    counts = [N_SRS, 1, 300, 4097, 0, 9]
    polys = [[rnd.randrange(B.R_ORDER) for _ in range(n)] for n in counts]
    for p in polys[:2]:
        if p:
            p[0] = 0  # zero coefficient
    polys[2][:3] = [1, B.R_ORDER - 1, 2]
    flat = [x for p in polys for x in p]
    lift = [rnd.randrange(0, 5) for _ in flat]  # non-canonical Montgomery encodings
    got = srs.kzg_commit(to_mont_bytes(flat, lift), counts)
    std = b"".join(x.to_bytes(32, "little") for x in flat)
    want = srs.msm_batch_resident(np.frombuffer(std, dtype=np.uint8), counts)
    assert got == want
    ks = scalars_to_ints(zk.gen_scalars(SEED, N_SRS, 0))
    for p, res in zip(polys, got):
        if p:
            assert result_affine(res) == closed_form(ks[:len(p)], p)
        else:
            assert res == bytes(96)


def test_commit_device_coeffs(srs):
    rnd = random.Random(2)
    counts = [5000, 5000]
    flat = [rnd.randrange(B.R_ORDER) for _ in range(sum(counts))]
    mont = np.frombuffer(to_mont_bytes(flat), dtype=np.uint8)
    d = zk.DeviceBuffer(mont.nbytes)
    d.upload(mont)
    got = srs.kzg_commit(d, counts)
    d.free()
    assert got == srs.kzg_commit(mont, counts)
    ks = scalars_to_ints(zk.gen_scalars(SEED, N_SRS, 0))
    assert result_affine(got[1]) == closed_form(ks[:5000], flat[5000:])
