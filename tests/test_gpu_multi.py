"""GPU parity tests of the multi-MSM entry points through the C ABI.

zkmsm_multi_msm: k scalar sets over one base set (Sources/zkMetal/MSM/
MultiMSM.swift:21-60 multiMSM).  zkmsm_inner_product_msm: one MSM per
partition of the points (MultiMSM.swift:90-124 innerProductMSM, including
its validation of the boundaries).  Bar: canonical affine results equal the
closed form (bases k_i * G), and each result is byte-identical to the same
MSM computed alone."""
import pytest

import zkmetal_amd as zk
from helpers import closed_form, result_affine, scalars_to_ints

pytestmark = pytest.mark.gpu

SEED = 0x3317


@pytest.fixture(scope="module")
def bases():
    e = zk.MetalMSM()
    try:
        e.srs_generate(SEED, 6000)
        pts = e.srs_download(6000)
    finally:
        e.close()
    return pts, scalars_to_ints(zk.gen_scalars(SEED, 6000, 0))


def test_multi_msm_matches_closed_form_and_single_calls(engine, bases):
    pts, ks = bases
    n, k = 5000, 6
    sets = zk.gen_scalars(0x31, n * k, 0)
    res = engine.multi_msm(pts[:64 * n], sets)
    assert len(res) == k
    ints = scalars_to_ints(sets)
    for j in range(k):
        assert result_affine(res[j]) == closed_form(ks[:n], ints[j * n:(j + 1) * n]), j
    for j in (0, k - 1):
        alone = engine.msm(pts[:64 * n], sets[j * n:(j + 1) * n])
        assert alone == res[j]


def test_multi_msm_one_set_and_mixed_scalars(engine, bases):
    pts, ks = bases
    n = 777
    sets = zk.gen_scalars(0x32, n * 3, 1)  # zeros, small values, 1 and r - 1 mixed in
    res = engine.multi_msm(pts[:64 * n], sets)
    ints = scalars_to_ints(sets)
    for j in range(3):
        assert result_affine(res[j]) == closed_form(ks[:n], ints[j * n:(j + 1) * n])
    one = engine.multi_msm(pts[:64 * n], sets[:n])
    assert one[0] == res[0]


def test_multi_msm_rejects_bad_shapes(engine, bases):
    pts, _ = bases
    with pytest.raises(zk.MSMError):
        engine.multi_msm(pts[:64 * 10], zk.gen_scalars(1, 15, 0))
    with pytest.raises(zk.MSMError):
        engine.multi_msm(b"", zk.gen_scalars(1, 4, 0))


def test_inner_product_partitions(engine, bases):
    pts, ks = bases
    n = 6000
    sc = zk.gen_scalars(0x33, n, 0)
    ints = scalars_to_ints(sc)
    parts = [0, 0, 1000, 1000, 2500, 2501, n]
    res = engine.inner_product_msm(pts[:64 * n], sc, parts)
    assert len(res) == len(parts) - 1
    for j in range(len(parts) - 1):
        lo, hi = parts[j], parts[j + 1]
        if lo == hi:
            assert res[j] == bytes(96)
        else:
            assert result_affine(res[j]) == closed_form(ks[lo:hi], ints[lo:hi]), j
    # the IPA halves form (MultiMSM.swift:127-144)
    left, right = engine.inner_product_msm(pts[:64 * n], sc, [0, n // 2, n])
    whole = engine.msm(pts[:64 * n], sc)
    from oracle import bn254 as B
    assert B.add(result_affine(left), result_affine(right)) == result_affine(whole)


@pytest.mark.parametrize("parts", [[1, 100], [0, 50], [0, 60, 40, 100], [0]])
def test_inner_product_rejects_bad_boundaries(engine, bases, parts):
    pts, _ = bases
    sc = zk.gen_scalars(0x34, 100, 0)
    with pytest.raises(zk.MSMError) as ei:
        engine.inner_product_msm(pts[:64 * 100], sc, parts)
    assert ei.value.status == zk.ERR_INVALID_INPUT
