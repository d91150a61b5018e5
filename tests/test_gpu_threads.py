"""Engines used from several host threads at once (ctypes releases the GIL
for every library call, so the calls really overlap): each engine serialises
its own calls, engines on one device run side by side, and the process-wide
engine behind the *_auto entry points is shared.  Every result equals the
closed form (sum s_i k_i mod r) * G computed serially beforehand."""
import ctypes
import threading

import pytest

import zkmetal_amd as zk
from helpers import closed_form, result_affine, scalars_to_ints

pytestmark = pytest.mark.gpu

SEED = 0x7417
N = 1 << 14


@pytest.fixture(scope="module")
def table():
    e = zk.MetalMSM()
    try:
        e.srs_generate(SEED, N)
        bases = e.srs_download(N)
    finally:
        e.close()
    return bases, scalars_to_ints(zk.gen_scalars(SEED, N, 0))


def _jobs(ks, seed, count):
    """count (n, scalars, expected point) jobs of assorted sizes"""
    out = []
    for j in range(count):
        n = [1, 77, 1000, 5000, N][j % 5]
        sc = zk.gen_scalars(seed + j, n, j % 3)
        out.append((n, sc, closed_form(ks[:n], scalars_to_ints(sc))))
    return out


def _run_threads(targets):
    errors = []

    def wrap(fn):
        try:
            fn()
        except Exception as ex:  # noqa: BLE001 -- reported below
            errors.append(repr(ex))

    th = [threading.Thread(target=wrap, args=(t,)) for t in targets]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=240)
    assert not any(t.is_alive() for t in th), "a thread did not finish"
    assert not errors, errors


def test_engines_in_parallel_threads(table):
    bases, ks = table
    engines = [zk.MetalMSM() for _ in range(3)]
    try:
        for e in engines[:2]:
            e.srs_upload(bases)
        jobs = [_jobs(ks, SEED + 100 * t, 10) for t in range(4)]
        got = [[None] * 10 for _ in range(4)]

        def host_calls(t, e):
            def fn():
                for i, (n, sc, _) in enumerate(jobs[t]):
                    got[t][i] = result_affine(e.msm(bases[:64 * n], sc))
            return fn

        def resident_calls(t, e):
            def fn():
                for i, (n, sc, _) in enumerate(jobs[t]):
                    got[t][i] = result_affine(e.msm_batch_resident(sc, [n])[0])
            return fn

        # threads 0 and 1 share engine 0 (its mutex serialises them); engine 1
        # serves resident calls, engine 2 host calls with uploaded bases
        _run_threads([host_calls(0, engines[0]), resident_calls(1, engines[0]),
                      resident_calls(2, engines[1]), host_calls(3, engines[2])])
        for t in range(4):
            for i, (n, _, want) in enumerate(jobs[t]):
                assert got[t][i] == want, (t, i, n)
    finally:
        for e in engines:
            e.close()


def test_auto_engine_shared_by_threads(table):
    bases, ks = table
    L = zk.lib()
    jobs = [_jobs(ks, SEED + 1000 + 100 * t, 5) for t in range(3)]
    got = [[None] * 5 for _ in range(3)]

    def auto_calls(t):
        def fn():
            for i, (n, sc, _) in enumerate(jobs[t]):
                out = (ctypes.c_uint8 * 96)()
                base = ctypes.addressof(out)
                pts = bases[:64 * n]
                rc = L.zkmetal_bn254_msm_auto(ctypes.c_void_p(pts.ctypes.data), ctypes.c_void_p(sc.ctypes.data),
                                              n, base, base + 32, base + 64)
                assert rc == zk.SUCCESS, rc
                got[t][i] = result_affine(bytes(out))
        return fn

    _run_threads([auto_calls(t) for t in range(3)])
    for t in range(3):
        for i, (n, _, want) in enumerate(jobs[t]):
            assert got[t][i] == want, (t, i, n)
