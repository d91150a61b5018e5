"""GPU tests of the resident-SRS detection for host-point calls
(zkmsm_set_srs_detect): a call whose point slices are byte-for-byte prefixes
of the resident SRS runs on the resident table, and its results are exactly
those of the ordinary path (zkMetal_ffi.swift:1241-1298 semantics; closed
form for bases k_i * G)."""
import pytest

import zkmetal_amd as zk
from helpers import closed_form, result_affine, scalars_to_ints

pytestmark = pytest.mark.gpu

SEED = 0x5D7


@pytest.fixture(scope="module")
def srs_pair():
    """An engine whose SRS was uploaded from host bytes, and those bytes."""
    gen = zk.MetalMSM()
    try:
        gen.srs_generate(SEED, 4000)
        pts = gen.srs_download(4000)
    finally:
        gen.close()
    e = zk.MetalMSM()
    e.srs_upload(pts)
    yield e, pts, scalars_to_ints(zk.gen_scalars(SEED, 4000, 0))
    e.close()


def test_single_and_batch_prefixes_hit(srs_pair):
    e, pts, ks = srs_pair
    sc = zk.gen_scalars(0x71, 4000, 0)
    ints = scalars_to_ints(sc)
    h0 = e.srs_detect_hits()
    r = e.msm(pts[:64 * 3000], sc[:3000])
    assert e.srs_detect_hits() == h0 + 1
    assert result_affine(r) == closed_form(ks[:3000], ints[:3000])
    counts = [100, 4000, 0, 57]
    cat_pts = b"".join(bytes(pts[:64 * c]) for c in counts)
    cat_sc = zk.gen_scalars(0x72, sum(counts), 1)
    res = e.msm_batch(cat_pts, cat_sc.tobytes(), counts)
    assert e.srs_detect_hits() == h0 + 2
    e.set_srs_detect(False)
    try:
        plain = e.msm_batch(cat_pts, cat_sc.tobytes(), counts)
    finally:
        e.set_srs_detect(True)
    assert e.srs_detect_hits() == h0 + 2
    assert res == plain
    assert res[2] == bytes(96)
    ints2 = scalars_to_ints(cat_sc)
    off = 0
    for j, c in enumerate(counts):
        if c:
            assert result_affine(res[j]) == closed_form(ks[:c], ints2[off:off + c]), j
        off += c


def test_near_misses_take_the_ordinary_path(srs_pair):
    e, pts, ks = srs_pair
    sc = zk.gen_scalars(0x73, 2000, 0)
    ints = scalars_to_ints(sc)
    h0 = e.srs_detect_hits()
    # one base in the middle replaced by another base of the set
    mod = bytearray(bytes(pts[:64 * 2000]))
    mod[64 * 1500:64 * 1501] = bytes(pts[64 * 3999:64 * 4000])
    kk = list(ks[:2000])
    kk[1500] = ks[3999]
    assert result_affine(e.msm(bytes(mod), sc)) == closed_form(kk, ints)
    # a slice that does not start at base 0
    r = e.msm(pts[64:64 * 2001], sc)
    assert result_affine(r) == closed_form(ks[1:2001], ints)
    assert e.srs_detect_hits() == h0


def test_generated_srs_matches_its_download():
    e = zk.MetalMSM()
    try:
        e.srs_generate(SEED + 1, 3000)
        pts = e.srs_download(3000)
        ks = scalars_to_ints(zk.gen_scalars(SEED + 1, 3000, 0))
        sc = zk.gen_scalars(0x74, 2500, 0)
        h0 = e.srs_detect_hits()
        r = e.msm(pts[:64 * 2500], sc)
        assert e.srs_detect_hits() == h0 + 1
        assert result_affine(r) == closed_form(ks[:2500], scalars_to_ints(sc))
        # longer than the SRS: the ordinary path
        sc2 = zk.gen_scalars(0x75, 3001, 0)
        longer = e.msm_batch(bytes(pts) + bytes(pts[:64]), sc2.tobytes(), [3001])
        assert e.srs_detect_hits() == h0 + 1
        assert result_affine(longer[0]) == closed_form(ks + ks[:1], scalars_to_ints(sc2))
    finally:
        e.close()
