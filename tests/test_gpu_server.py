"""GPU tests of the request server (zkmsm_server_*): every result equals the
same MSM computed alone, bit for bit, whatever batch it was served in;
admission respects the HBM budget; concurrent submitters are safe."""
import random
import threading
import time

import numpy as np
import pytest

import zkmetal_amd as zk
from helpers import closed_form, result_affine, scalars_to_ints

pytestmark = pytest.mark.gpu

SEED = 0x5E7E
N_SRS = 1 << 16


@pytest.fixture(scope="module")
def srs_engine():
    e = zk.MetalMSM()
    e.srs_generate(SEED, N_SRS)
    yield e
    e.close()


def _requests(seed, count, lo, hi, dist):
    rnd = random.Random(seed)
    sizes = [int(2 ** rnd.uniform(lo, hi)) for _ in range(count)]
    return [zk.gen_scalars(seed * 1000 + i, n, dist) for i, n in enumerate(sizes)]


def test_results_equal_alone_and_in_the_mix(srs_engine):
    reqs = _requests(1, 40, 4, 16, 1)
    budget = 3 * zk.server_footprint(1 << 15)  # forces admission to split the stream
    srv = zk.Server(srs_engine, slots=2, max_batch=8, hbm_budget_bytes=budget)
    rnd = random.Random(2)
    tickets = []
    for sc in reqs:
        tickets.append(srv.submit(sc))
        time.sleep(rnd.random() * 0.002)
    results = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["batches"] >= 2 and st["mean_batch"] >= 1
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["hbm_in_use_bytes"] == 0
    assert st["hbm_peak_bytes"] <= max(budget, max(zk.server_footprint(len(s)) for s in reqs))
    for sc, (res, lat) in zip(reqs, results):
        alone = srs_engine.msm_batch_resident(sc, [len(sc)])[0]
        assert res == alone
        assert lat > 0
    ks = scalars_to_ints(zk.gen_scalars(SEED, N_SRS, 0))
    for sc, (res, _) in list(zip(reqs, results))[:6]:
        assert result_affine(res) == closed_form(ks[:len(sc)], scalars_to_ints(sc))


def test_concurrent_submitters_and_poll(srs_engine):
    reqs = _requests(3, 24, 8, 14, 0)
    srv = zk.Server(srs_engine, slots=3, max_batch=4)
    out = [None] * len(reqs)

    def client(idx):
        for i in idx:
            t = srv.submit(reqs[i])
            while True:
                r = srv.poll(t)
                if r is not None:
                    out[i] = r
                    break
                time.sleep(0.0005)

    th = [threading.Thread(target=client, args=(list(range(k, len(reqs), 4)),)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    srv.close()
    for sc, r in zip(reqs, out):
        assert r == srs_engine.msm_batch_resident(sc, [len(sc)])[0]


def test_invalid_requests(srs_engine):
    srv = zk.Server(srs_engine, slots=1)
    with pytest.raises(zk.MSMError):
        srv.submit(b"")
    with pytest.raises(zk.MSMError):
        srv.submit(np.zeros((N_SRS + 1, 32), dtype=np.uint8))
    with pytest.raises(zk.MSMError):
        srv.wait(123456)
    srv.close()


def test_srs_replacement_refused_while_a_server_reads_it():
    """A server reads its owner's resident table: while it exists, the calls
    that would replace that table are refused and the server keeps serving
    the original bases; after it is closed they work again."""
    e = zk.MetalMSM()
    try:
        e.srs_generate(SEED, 1 << 12)
        bases = e.srs_download(1 << 12)
        sc = zk.gen_scalars(0x11FE, 1000, 0)
        want = e.msm_batch_resident(sc, [1000])[0]
        srv = zk.Server(e, slots=1)
        try:
            for fn in (lambda: e.srs_generate(SEED + 1, 1 << 13),
                       lambda: e.srs_upload(bases[:64 * 100]),
                       lambda: e.srs_load_ptau(data=b"ptau" + bytes(64))):
                with pytest.raises(zk.MSMError) as ei:
                    fn()
                assert ei.value.status == zk.ERR_INVALID_INPUT
            r, _ = srv.wait(srv.submit(sc))
            assert r == want
        finally:
            srv.close()
        e.srs_generate(SEED + 1, 1 << 12)  # no server: allowed again
        assert e.msm_batch_resident(sc, [1000])[0] != want
    finally:
        e.close()


def test_engine_handle_released_before_its_server():
    """Destroying the SRS engine's handle while a server is attached defers
    the release: the server still computes over the table."""
    e = zk.MetalMSM()
    e.srs_generate(SEED, 1 << 12)
    sc = zk.gen_scalars(0x11FF, 3000, 1)
    want = e.msm_batch_resident(sc, [3000])[0]
    srv = zk.Server(e, slots=2)
    e.close()
    try:
        r, _ = srv.wait(srv.submit(sc))
        assert r == want
    finally:
        srv.close()
