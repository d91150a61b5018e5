"""GPU parity tests of the BN254 G1 MSM through the C ABI.

Bar: the canonical affine result equals the oracle's bit for bit -- the
golden fixtures (pure-Python oracle), the C restatement of the reference CPU
Pippenger, and the closed form (sum s_i k_i mod r) * G for bases k_i * G,
which holds at every size.  Edge cases follow the reference's own tests
(Sources/zkbench/correctness_tests.swift:1701-1840, MSMTests.swift)."""
import math
import random

import numpy as np
import pytest

import zkmetal_amd as zk
from helpers import (affine_bytes_to_points, assert_normalised, closed_form, expected_affine,
                     golden_msm_fixtures, result_affine, scalars_to_ints)
from oracle import bn254 as B
from oracle import ref

pytestmark = pytest.mark.gpu

SRS_SEED = 0x5125EED
SRS_N = 1 << 16


@pytest.fixture(scope="module")
def srs(engine):
    """Device-generated table P_i = k_i G; returns (host bases bytes, k_i ints)."""
    engine.srs_generate(SRS_SEED, SRS_N)
    bases = engine.srs_download(SRS_N)
    ks = scalars_to_ints(zk.gen_scalars(SRS_SEED, SRS_N, 0))
    return bases, ks


@pytest.mark.parametrize("fx", golden_msm_fixtures(), ids=lambda f: f["name"])
def test_golden_fixture(engine, fx):
    r = engine.msm(bytes.fromhex(fx["points"]), bytes.fromhex(fx["scalars"]))
    assert_normalised(r)
    assert result_affine(r) == expected_affine(fx)


def test_golden_fixtures_in_one_batch(engine):
    fxs = golden_msm_fixtures()
    pts = b"".join(bytes.fromhex(f["points"]) for f in fxs)
    sc = b"".join(bytes.fromhex(f["scalars"]) for f in fxs)
    counts = [f["n"] for f in fxs]
    # a zero-count entry in the middle writes 96 zero bytes (zkMetal_ffi.swift:1260-1264)
    counts.insert(3, 0)
    res = engine.msm_batch(pts, sc, counts)
    assert res[3] == bytes(96)
    res = res[:3] + res[4:]
    for f, r in zip(fxs, res):
        assert result_affine(r) == expected_affine(f), f["name"]


def test_invalid_inputs(engine):
    with pytest.raises(zk.MSMError) as ei:
        engine.msm(b"", b"", n=0)
    assert ei.value.status == zk.ERR_INVALID_INPUT
    with pytest.raises(zk.MSMError):
        engine.msm_batch(b"", b"", [])
    with pytest.raises(zk.MSMError):
        engine.set_window_bits(17)


def test_srs_generate_matches_cpu(srs):
    bases, ks = srs
    pts = affine_bytes_to_points(bases[:64 * 300])
    for i in list(range(200)) + [299]:
        assert pts[i] == B.mul(B.G, ks[i]), i


def test_small_scalar_variants(engine, srs):
    bases, ks = srs
    n = 777
    rnd = np.random.default_rng(4)
    for width, dt in ((1, np.uint8), (2, np.uint16), (4, np.uint32)):
        s = rnd.integers(0, np.iinfo(dt).max, size=n, endpoint=True, dtype=dt)
        r = engine.msm_small(bases[:64 * n], s.tobytes(), width)
        assert result_affine(r) == closed_form(ks[:n], [int(x) for x in s]), width


@pytest.mark.parametrize("n", [1, 5, 63, 1000, 4097, 20000])
def test_random_vs_closed_form(engine, srs, n):
    bases, ks = srs
    for dist in (0, 1, 2):
        sc = zk.gen_scalars(1000 + n + dist, n, dist)
        r = engine.msm(bases[:64 * n], sc)
        assert result_affine(r) == closed_form(ks[:n], scalars_to_ints(sc)), (n, dist)


def test_every_window_width(engine, srs):
    bases, ks = srs
    n = 3000
    sc = zk.gen_scalars(77, n, 0)
    want = closed_form(ks[:n], scalars_to_ints(sc))
    try:
        for c in range(2, 17):
            engine.set_window_bits(c)
            assert result_affine(engine.msm(bases[:64 * n], sc)) == want, c
    finally:
        engine.set_window_bits(0)


def test_2_16_against_c_oracle(engine, srs):
    bases, ks = srs
    n = 1 << 16
    sc = zk.gen_scalars(0x16, n, 0)
    r = engine.msm(bases[:64 * n], sc)
    jac = ref.pippenger(bytes(bases[:64 * n]), sc.tobytes(), n, 0, 8)
    aff = ref.jacobian_to_affine_bytes(jac)
    c_pt = (B.from_mont(int.from_bytes(aff[:32], "little")), B.from_mont(int.from_bytes(aff[32:], "little")))
    assert result_affine(r) == c_pt
    assert c_pt == closed_form(ks[:n], scalars_to_ints(sc))


def test_repeated_and_cancelling_bases(engine, srs):
    bases, ks = srs
    n = 4096
    # every base appears 4 times with scalars 1..4 plus a cancelling copy
    idx = [i % 512 for i in range(n)]
    pts = b"".join(bytes(bases[64 * i:64 * i + 64]) for i in idx)
    ss = [(i // 512) % 4 + 1 for i in range(n)]
    sc = b"".join(B.scalar_to_bytes(s) for s in ss)
    want = closed_form([ks[i] for i in idx], ss)
    assert result_affine(engine.msm(pts, sc)) == want
    ss2 = [B.R_ORDER - s if i >= n // 2 else s for i, s in enumerate([7] * n)]
    pts2 = pts[:64 * (n // 2)] * 2
    sc2 = b"".join(B.scalar_to_bytes(s) for s in ss2)
    assert result_affine(engine.msm(pts2, sc2)) is None


def test_identity_bases_scattered(engine, srs):
    """(0,0) bases anywhere: a whole 64-point bitmap word, single ones on word
    boundaries and at the last index; in one MSM and in a batch whose parts
    start at offsets that are not multiples of 64."""
    bases, ks = srs
    n = 5000
    rnd = random.Random(77)
    ident = set(range(128, 192)) | {0, 63, 64, 65, 1000, n - 1}
    ident |= {rnd.randrange(n) for _ in range(200)}
    pts = bytearray(bases[:64 * n])
    for i in ident:
        pts[64 * i:64 * i + 64] = bytes(64)
    sc = zk.gen_scalars(0x1D, n, 0)
    ints = scalars_to_ints(sc)
    kk = [0 if i in ident else ks[i] for i in range(n)]
    assert result_affine(engine.msm(bytes(pts), sc)) == closed_form(kk, ints)
    counts = [37, 1000, n - 1037]
    res = engine.msm_batch(bytes(pts), sc.tobytes(), counts)
    off = 0
    for j, c in enumerate(counts):
        assert result_affine(res[j]) == closed_form(kk[off:off + c], ints[off:off + c]), j
        off += c


def test_resident_batch_mixed_sizes(engine, srs):
    bases, ks = srs
    rnd = random.Random(12)
    counts = [int(2 ** rnd.uniform(4, 16)) for _ in range(40)] + [1, SRS_N, 0, 2]
    sc_all = zk.gen_scalars(0xBA7C4, sum(counts), 1)
    res = engine.msm_batch_resident(sc_all, counts)
    off = 0
    ints = scalars_to_ints(sc_all)
    for j, c in enumerate(counts):
        if c == 0:
            assert res[j] == bytes(96)
            continue
        assert result_affine(res[j]) == closed_form(ks[:c], ints[off:off + c]), (j, c)
        off += c
    # alone == in the mix, byte for byte (outputs are normalised)
    off = 0
    for j, c in enumerate(counts[:6]):
        alone = engine.msm_batch_resident(sc_all[off:off + c], [c])[0]
        assert alone == res[j]
        off += c


def test_device_resident_scalars(engine, srs):
    bases, ks = srs
    counts = [SRS_N, 1000, 12345]
    sc = zk.gen_scalars(99, sum(counts), 0)
    buf = zk.DeviceBuffer(sc.nbytes)
    buf.upload(sc)
    try:
        res_dev = engine.msm_batch_resident(buf, counts)
    finally:
        buf.free()
    res_host = engine.msm_batch_resident(sc, counts)
    assert res_dev == res_host
    ints = scalars_to_ints(sc)
    assert result_affine(res_dev[2]) == closed_form(ks[:12345], ints[SRS_N + 1000:])


def test_deterministic(engine, srs):
    bases, _ = srs
    sc = zk.gen_scalars(5, 30000, 0)
    a = engine.msm(bases[:64 * 30000], sc)
    b = engine.msm(bases[:64 * 30000], sc)
    assert a == b


def test_2_20_plain_closed_form(engine):
    n = 1 << 20
    e2 = zk.MetalMSM()
    try:
        e2.srs_generate(0x2020, n)
        bases = e2.srs_download(n)
    finally:
        e2.close()
    ks = scalars_to_ints(zk.gen_scalars(0x2020, n, 0))
    sc = zk.gen_scalars(0x2021, n, 0)
    r = engine.msm(bases, sc)
    assert result_affine(r) == closed_form(ks, scalars_to_ints(sc))
