"""GPU parity at the shapes of the three system workloads, each result
against the closed form (sum s_i k_i mod r) * G of the resident bases
k_i * G (the C oracle's exact dot product, then one Python scalar
multiplication):

* batched: 64 MSMs of sizes floor(2^U), U ~ U[12, 20] (seeded), over
  prefixes of one resident 2^20 table, in one call -- once eagerly and twice
  as a captured hipGraph (capture, then replay);
* skewed: 32 MSMs of 2^18 witness-like scalars (40% zero, 30% below 2^16,
  10% 1 or r - 1, 20% uniform) in one call;
* served: a seeded Poisson stream (200 requests/s) of 40 requests with sizes
  log-uniform in [2^10, 2^22] and witness-like scalars through the request
  server on a resident 2^23 SRS; each result equals the closed form and the
  same MSM computed alone.

The reference cross-checks its GPU MSM against its CPU path at large n too
(Sources/zkbench/cpu_bench.swift:72-81)."""
import time

import numpy as np
import pytest

import zkmetal_amd as zk
from oracle import bn254 as B
from oracle import ref
from helpers import result_affine

pytestmark = pytest.mark.gpu


def _closed_forms(ks, sc, counts):
    """expected affine point of each MSM j over the first counts[j] bases"""
    sb = sc.tobytes()
    want, off = [], 0
    for n in counts:
        want.append(B.mul(B.G, ref.scalar_dot(sb[32 * off:32 * (off + n)], ks[:32 * n], n)))
        off += n
    return want


def test_batched_64_msms_over_a_2_20_table_eager_and_graph():
    seed = 0xB47C
    rnd = np.random.default_rng(seed)
    counts = [int(2 ** u) for u in rnd.uniform(12, 20, size=64)]
    print(f"batched: 64 MSMs, {sum(counts)} points, sizes {min(counts)}..{max(counts)}", flush=True)
    eng = zk.MetalMSM()
    ds = zk.DeviceBuffer(32 * sum(counts))
    try:
        eng.srs_generate(seed, 1 << 20)
        sc = zk.gen_scalars(seed + 2, sum(counts), 0)
        ds.upload(sc)
        want = _closed_forms(zk.gen_scalars(seed, 1 << 20, 0).tobytes(), sc, counts)
        eager = eng.msm_batch_resident(ds, counts)
        assert [result_affine(r) for r in eager] == want
        eng.set_graph_mode(True)
        caps = eng.graph_captures()
        graph = [eng.msm_batch_resident(ds, counts) for _ in range(2)]
        assert eng.graph_captures() == caps + 1  # the second call replays
        assert graph[0] == eager and graph[1] == eager
    finally:
        eng.set_graph_mode(False)
        ds.free()
        eng.close()


def test_skewed_32_msms_of_2_18_witness_like_scalars():
    seed = 0x5C3D
    counts = [1 << 18] * 32
    print("skewed: 32 x 2^18 witness-like", flush=True)
    eng = zk.MetalMSM()
    ds = zk.DeviceBuffer(32 * sum(counts))
    try:
        eng.srs_generate(seed, 1 << 18)
        sc = zk.gen_scalars(seed + 1, sum(counts), 1)
        ds.upload(sc)
        want = _closed_forms(zk.gen_scalars(seed, 1 << 18, 0).tobytes(), sc, counts)
        got = eng.msm_batch_resident(ds, counts)
        assert [result_affine(r) for r in got] == want
    finally:
        ds.free()
        eng.close()


def test_served_poisson_stream_on_a_2_23_srs():
    seed = 0x5E4F
    n_srs = 1 << 23
    rnd = np.random.default_rng(seed)
    nreq = 40
    sizes = [int(2 ** u) for u in rnd.uniform(10, 22, size=nreq)]
    gaps = rnd.exponential(1.0 / 200.0, size=nreq)
    print(f"served: {nreq} requests, {sum(sizes)} points, 2^23 SRS", flush=True)
    eng = zk.MetalMSM()
    try:
        eng.srs_generate(seed, n_srs)
        ks = zk.gen_scalars(seed, n_srs, 0).tobytes()
        reqs = [zk.gen_scalars(seed * 100 + i, n, 1) for i, n in enumerate(sizes)]
        srv = zk.Server(eng, slots=2, max_batch=16, hbm_budget_bytes=8 << 30)
        try:
            tickets = []
            for g, sc in zip(gaps, reqs):
                time.sleep(float(g))
                tickets.append(srv.submit(sc))
            served = [srv.wait(t)[0] for t in tickets]
            st = srv.stats()
        finally:
            srv.close()
        assert st["completed"] == nreq and st["failed"] == 0
        for i, (sc, r) in enumerate(zip(reqs, served)):
            assert result_affine(r) == _closed_forms(ks, sc, [len(sc)])[0], f"request {i}"
            assert eng.msm_batch_resident(sc, [len(sc)])[0] == r, f"request {i} alone"
    finally:
        eng.close()
