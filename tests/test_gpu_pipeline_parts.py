"""The part pipeline (bucket sort of part p + 1 on a second stream while part
p accumulates; one merge of every part's crossing partials and one bucket
reduction after the last accumulation) gives the same bytes for every number
of parts: single MSMs, batches of mixed sizes, a skewed batch (heavy buckets
through the long-run merge), Grumpkin and G2, eager and captured as a graph;
and the closed form."""
import numpy as np
import pytest

import zkmetal_amd as zk
from oracle import bn254 as B
from oracle import ref
from helpers import result_affine

pytestmark = pytest.mark.gpu

SEED = 0x9A27


@pytest.fixture(scope="module")
def eng():
    e = zk.MetalMSM()
    e.srs_generate(SEED, 1 << 18)
    yield e
    e.set_pipeline_parts(0)
    e.close()


def _each_parts(e, fn, parts=(1, 2, 3, 4, 8)):
    out = []
    for p in parts:
        e.set_pipeline_parts(p)
        out.append(fn())
    e.set_pipeline_parts(0)
    return out


def test_single_msm_any_parts(eng):
    n = 1 << 18
    sc = zk.gen_scalars(SEED + 1, n, 0)
    ds = zk.DeviceBuffer(sc.nbytes)
    ds.upload(sc)
    try:
        print("single 2^18 over parts 1..8", flush=True)
        res = _each_parts(eng, lambda: eng.msm_batch_resident(ds, [n]))
        assert all(r == res[0] for r in res)
        ks = zk.gen_scalars(SEED, n, 0).tobytes()
        assert result_affine(res[0][0]) == B.mul(B.G, ref.scalar_dot(sc.tobytes(), ks, n))
    finally:
        ds.free()


def test_mixed_and_skewed_batches_any_parts(eng):
    rnd = np.random.default_rng(SEED)
    counts = [int(2 ** u) for u in rnd.uniform(8, 18, size=24)] + [1, 3]
    for dist in (0, 1):
        sc = zk.gen_scalars(SEED + 2 + dist, sum(counts), dist)
        print(f"batch of {len(counts)} (dist {dist}) over parts", flush=True)
        res = _each_parts(eng, lambda: eng.msm_batch_resident(sc, counts))
        assert all(r == res[0] for r in res)


def test_graph_replay_with_parts(eng):
    counts = [1 << 17, 5000, 1 << 16]
    sc = zk.gen_scalars(SEED + 5, sum(counts), 0)
    ds = zk.DeviceBuffer(sc.nbytes)
    ds.upload(sc)
    try:
        eng.set_pipeline_parts(1)
        want = eng.msm_batch_resident(ds, counts)
        eng.set_pipeline_parts(4)
        eng.set_graph_mode(True)
        got = [eng.msm_batch_resident(ds, counts) for _ in range(3)]
        assert all(g == want for g in got)
    finally:
        eng.set_graph_mode(False)
        eng.set_pipeline_parts(0)
        ds.free()


def test_grumpkin_and_g2_any_parts(eng):
    n = 1 << 15
    sc = zk.gen_scalars(SEED + 6, n, 0)
    ds = zk.DeviceBuffer(sc.nbytes)
    ds.upload(sc)
    dp = zk.DeviceBuffer(64 * n)
    dq = zk.DeviceBuffer(128 * n)
    try:
        eng.grumpkin_points_generate(SEED + 7, n, dp)
        eng.g2_points_generate(SEED + 8, n, dq)
        rg = _each_parts(eng, lambda: eng.grumpkin_msm_batch_device(dp, ds, [n]), parts=(1, 2, 5))
        assert all(r == rg[0] for r in rg)
        r2 = _each_parts(eng, lambda: eng.g2_msm_batch_device(dq, ds, [n]), parts=(1, 4))
        assert all(r == r2[0] for r in r2)
    finally:
        ds.free()
        dp.free()
        dq.free()
