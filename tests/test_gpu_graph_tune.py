"""GPU tests of graph mode (captured launch sequence replayed per call shape)
and of the autotuner (window widths and the per-size-class launch schedule:
pipeline parts, buckets per reduction lane): all change how the work is
launched or cut into windows, never the result."""
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
    res = tune.tune(None, min_logn=10, max_logn=12, reps=1, spread=1, batch_points=1 << 13,
                    sched_batch_logn=(13, 14))
    for key in ("plain", "batch"):
        assert set(res[key]) == {10, 11, 12}
        assert all(4 <= c <= 16 for c in res[key].values())
    plain, batch = tune.tables(res)
    assert plain[12] == res["plain"][12] and batch[10] == res["batch"][10]
    for key in ("plain", "batch"):
        for cls, (p, g) in res["schedule"][key].items():
            assert 0 <= p <= 8 and g in (0, 4, 8, 16, 32, 64) and (p or g), (key, cls)
    assert res["schedule_ms"]["plain"] and res["schedule_ms"]["batch"]
    tune.apply(srs, res)
    assert srs.msm_batch_resident(sc, [3000]) == before
    assert srs.msm_batch_resident(zk.gen_scalars(SEED + 4, 4 * 1000, 0), [1000] * 4) == before_batch
    srs.set_window_table([0] * 33)
    srs.set_window_table([0] * 33, batch=True)
    srs.set_schedule_table([0] * 33, [0] * 33)
    srs.set_schedule_table([0] * 33, [0] * 33, batch=True)


def _class_of_last_call(eng):
    return (int(eng.last_timing()["entries"]) - 1).bit_length()


@pytest.mark.parametrize("batch", [False, True])
def test_schedule_tables_set_parts_and_reduction_tile(srs, batch):
    counts = [N_SRS, 3000, 1, 20000] if batch else [N_SRS]
    scal = zk.gen_scalars(SEED + 7, sum(counts), 0)
    d = zk.DeviceBuffer(scal.nbytes)
    d.upload(scal)
    srs.set_graph_mode(srs.GRAPH_OFF)
    try:
        srs.set_schedule_table([0] * 33, [0] * 33, batch=batch)
        want = srs.msm_batch_resident(d, counts)
        cls = _class_of_last_call(srs)
        builtin_per = srs.last_timing()["buckets_per_group"]
        assert builtin_per >= 8
        for parts, per in ((3, 4), (1, 64), (5, 16), (2, 0), (0, 32)):
            pt, pr = [0] * 33, [0] * 33
            pt[cls], pr[cls] = parts, per
            srs.set_schedule_table(pt, pr, batch=batch)
            assert srs.msm_batch_resident(d, counts) == want, (parts, per)
            t = srs.last_timing()
            if parts:
                # the plan has at least 3 segments here; with more parts than
                # segments it makes one part per segment
                assert 1 <= t["parts"] <= parts, (parts, t["parts"])
                if parts <= 3:
                    assert t["parts"] == parts, (parts, t["parts"])
            if per:
                assert t["buckets_per_group"] == per
        # the other table (one MSM vs several) does not apply to this call
        srs.set_schedule_table([0] * 33, [0] * 33, batch=batch)
        pt = [0] * 33
        pt[cls] = 1
        srs.set_schedule_table(pt, [4] * 33, batch=not batch)
        assert srs.msm_batch_resident(d, counts) == want
        assert srs.last_timing()["buckets_per_group"] == builtin_per
        srs.set_schedule_table([0] * 33, [0] * 33, batch=not batch)
        # out-of-range values are refused and leave the table as it was
        for bad_parts, bad_per in ((9, 0), (0, 3), (0, 2), (0, 128), (0, 12)):
            pt, pr = [0] * 33, [0] * 33
            pt[cls], pr[cls] = bad_parts, bad_per
            with pytest.raises(zk.MSMError):
                srs.set_schedule_table(pt, pr, batch=batch)
        with pytest.raises(zk.MSMError):
            srs.set_schedule_table([0] * 33, [0] * 32, batch=batch)
    finally:
        srs.set_schedule_table([0] * 33, [0] * 33, batch=batch)
        srs.set_graph_mode(srs.GRAPH_AUTO)
        d.free()


def test_extreme_schedule_at_2_20_matches_the_closed_form():
    """Eight parts and 64 buckets per reduction lane on one 2^20-point MSM
    (device bases, table built per call) give the closed form."""
    from helpers import result_affine
    from oracle import bn254 as B
    from oracle import ref

    n = 1 << 20
    e = zk.MetalMSM()
    try:
        e.set_graph_mode(e.GRAPH_OFF)
        e.srs_generate(SEED + 11, n)
        bases = e.srs_download(n)
        sc = zk.gen_scalars(SEED + 12, n, 0)
        dp, ds = zk.DeviceBuffer(bases.nbytes), zk.DeviceBuffer(sc.nbytes)
        dp.upload(bases)
        ds.upload(sc)
        try:
            for parts, per in ((8, 64), (1, 4)):
                e.set_schedule_table([parts] * 33, [per] * 33)
                got = result_affine(e.msm_batch_device(dp, ds, [n])[0])
                t = e.last_timing()
                assert t["buckets_per_group"] == per and t["parts"] == parts
                ks = zk.gen_scalars(SEED + 11, n, 0).tobytes()
                assert got == B.mul(B.G, ref.scalar_dot(sc.tobytes(), ks, n)), (parts, per)
        finally:
            dp.free()
            ds.free()
    finally:
        e.close()
