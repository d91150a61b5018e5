"""The window combine and normalisation on the device (k_tail_windows /
k_tail_combine) against the host tail: the same bytes for every case --
single MSMs at every window width, batches of mixed sizes with empty and
one-point MSMs, all-zero scalars (the identity), Grumpkin batches, and graph
replays -- and the closed form for a sample (the host tail restates
MSMEngine.swift:899-904; both must give the canonical affine result)."""
import numpy as np
import pytest

import zkmetal_amd as zk
from oracle import bn254 as B
from oracle import ref
from helpers import result_affine

pytestmark = pytest.mark.gpu

SEED = 0x7A11
N_SRS = 1 << 15


@pytest.fixture(scope="module")
def eng():
    e = zk.MetalMSM()
    e.srs_generate(SEED, N_SRS)
    yield e
    e.set_tail_mode(e.TAIL_AUTO)
    e.close()


def _both(e, fn):
    e.set_tail_mode(e.TAIL_HOST)
    h = fn()
    assert e.last_timing()["device_tail_ms"] == 0
    e.set_tail_mode(e.TAIL_DEVICE)
    d = fn()
    e.set_tail_mode(e.TAIL_AUTO)
    return h, d


def test_single_msms_every_window_width(eng):
    sc = zk.gen_scalars(SEED + 1, 5000, 0)
    for c in range(2, 17):
        eng.set_window_bits(c)
        try:
            h, d = _both(eng, lambda: eng.msm_batch_resident(sc, [5000]))
        finally:
            eng.set_window_bits(0)
        assert h == d, f"c={c}"
    ks = zk.gen_scalars(SEED, 5000, 0).tobytes()
    assert result_affine(d[0]) == B.mul(B.G, ref.scalar_dot(sc.tobytes(), ks, 5000))


def test_mixed_batch_with_empty_one_point_and_zero_msms(eng):
    rnd = np.random.default_rng(3)
    counts = [int(x) for x in rnd.integers(1, N_SRS, size=40)] + [1, 0, 2, N_SRS, 0, 7]
    sc = zk.gen_scalars(SEED + 2, sum(counts), 1)
    sc[:counts[0]] = 0  # the first MSM's scalars all zero: the identity
    h, d = _both(eng, lambda: eng.msm_batch_resident(sc, counts))
    assert h == d
    assert result_affine(d[0]) is None
    assert d[41] == bytes(96)  # zero count: 96 zero bytes
    one = B.to_mont(1).to_bytes(32, "little")
    assert d[0] == one + one + bytes(32)  # identity as (1, 1, 0)


def test_graph_replays_end_with_final_results(eng):
    counts = [3000, 1 << 14, 17, 1]
    sc = zk.gen_scalars(SEED + 3, sum(counts), 0)
    ds = zk.DeviceBuffer(sc.nbytes)
    ds.upload(sc)
    try:
        eng.set_tail_mode(eng.TAIL_HOST)
        want = eng.msm_batch_resident(ds, counts)
        eng.set_tail_mode(eng.TAIL_AUTO)
        eng.set_graph_mode(True)
        try:
            got = [eng.msm_batch_resident(ds, counts) for _ in range(3)]
        finally:
            eng.set_graph_mode(False)
        assert all(g == want for g in got)
    finally:
        ds.free()


def test_grumpkin_batch(eng):
    counts = [1500, 1, 4000, 64]
    n = sum(counts)
    dp = zk.DeviceBuffer(64 * n)
    sc = zk.gen_scalars(SEED + 4, n, 2)
    ds = zk.DeviceBuffer(sc.nbytes)
    try:
        eng.grumpkin_points_generate(SEED + 5, n, dp)
        ds.upload(sc)
        h, d = _both(eng, lambda: eng.grumpkin_msm_batch_device(dp, ds, counts))
        assert h == d
    finally:
        dp.free()
        ds.free()
