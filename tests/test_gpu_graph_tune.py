"""GPU tests of graph mode (captured launch sequence replayed per call shape)
and of the window-width autotuner: both change how the work is launched or
cut into windows, never the result."""
import pytest

import zkmetal_amd as zk
from zkmetal_amd import tune

pytestmark = pytest.mark.gpu

SEED = 0x6A4F
N_SRS = 1 << 15


@pytest.fixture(scope="module")
def srs():
    e = zk.MetalMSM()
    e.srs_generate(SEED, N_SRS)
    yield e
    e.close()


def test_graph_replays_match_eager(srs):
    counts = [N_SRS, 5000, 1, 77, 1 << 12]
    scal = zk.gen_scalars(SEED + 1, sum(counts), 1)
    other = zk.gen_scalars(SEED + 2, 3000, 0)
    d = zk.DeviceBuffer(scal.nbytes)
    d.upload(scal)
    eager = srs.msm_batch_resident(d, counts)
    eager_other = srs.msm_batch_resident(other, [3000])
    srs.set_graph_mode(True)
    try:
        c0 = srs.graph_captures()
        for rep in range(4):
            assert srs.msm_batch_resident(d, counts) == eager, rep
            # eager-shaped work between replays: a different call shape
            # captures its own graph, then the first shape captures again
            assert srs.msm_batch_resident(other, [3000]) == eager_other, rep
        assert srs.graph_captures() - c0 == 8
        # same shape back to back: replays without capturing
        c1 = srs.graph_captures()
        for rep in range(3):
            assert srs.msm_batch_resident(d, counts) == eager, rep
        assert srs.graph_captures() - c1 == 1
        t = srs.last_timing()
        assert t["device_ms"] > 0 and t["accumulate_ms"] == 0
    finally:
        srs.set_graph_mode(False)
        d.free()
    assert srs.msm_batch_resident(other, [3000]) == eager_other


def test_auto_mode_replays_a_repeated_shape(srs):
    counts = [N_SRS, 20000, 1, 300]
    scal = zk.gen_scalars(SEED + 5, sum(counts), 1)
    other = zk.gen_scalars(SEED + 6, 3000, 0)
    d = zk.DeviceBuffer(scal.nbytes)
    d.upload(scal)
    try:
        srs.set_graph_mode(srs.GRAPH_OFF)
        eager = srs.msm_batch_resident(d, counts)
        eager_other = srs.msm_batch_resident(other, [3000])
        srs.set_graph_mode(srs.GRAPH_AUTO)
        c0 = srs.graph_captures()
        # first call eager, the second (same shape and buffers) captures, the
        # rest replay
        assert [srs.msm_batch_resident(d, counts) for _ in range(4)] == [eager] * 4
        assert srs.graph_captures() - c0 == 1
        assert srs.last_timing()["accumulate_ms"] == 0  # a replay: device total only
        # a one-off shape runs eagerly and captures nothing; the captured shape
        # then replays again
        assert srs.msm_batch_resident(other, [3000]) == eager_other
        assert srs.last_timing()["accumulate_ms"] > 0
        assert srs.msm_batch_resident(d, counts) == eager
        assert srs.graph_captures() - c0 == 1
    finally:
        srs.set_graph_mode(srs.GRAPH_AUTO)
        d.free()


def test_tuner_tables_keep_results(srs):
    sc = zk.gen_scalars(SEED + 3, 3000, 2)
    before = srs.msm_batch_resident(sc, [3000])
    before_batch = srs.msm_batch_resident(zk.gen_scalars(SEED + 4, 4 * 1000, 0), [1000] * 4)
    # the sweep runs on its own engine (own seeded table); its tables then
    # go to this one
    res = tune.tune(None, min_logn=10, max_logn=12, reps=1, spread=1, batch_points=1 << 13)
    for key in ("plain", "batch"):
        assert set(res[key]) == {10, 11, 12}
        assert all(4 <= c <= 16 for c in res[key].values())
    plain, batch = tune.tables(res)
    assert plain[12] == res["plain"][12] and batch[10] == res["batch"][10]
    tune.apply(srs, res)
    assert srs.msm_batch_resident(sc, [3000]) == before
    assert srs.msm_batch_resident(zk.gen_scalars(SEED + 4, 4 * 1000, 0), [1000] * 4) == before_batch
    srs.set_window_table([0] * 33)
    srs.set_window_table([0] * 33, batch=True)
