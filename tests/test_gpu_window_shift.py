"""Window-shifted resident tables (zkmsm_set_window_shift): an engine whose
SRS holds the copies 2^(16 w) P_i and runs MSMs with one bucket set per MSM
gives the same bytes as an engine with the plain table, on every path that
reads the resident table -- single MSMs of small, ragged (windows padded to
whole sort tiles) and full size, batches that mix shifted and plain MSMs
(the auto mode's cost estimate), every pipeline-part split of one MSM
(windows spread over several segments), the host and device tails and
graph replays, Montgomery-form coefficients (KZG), the allocation-free entry
point and the request server -- and a sample of each against the closed
form (sum s_i k_i mod r) * G.  Identity bases (0, 0) in an uploaded SRS get
no shifted copies and no digits.  The reference reduces every window
separately (Sources/zkMetal/MSM/MSMEngine.swift:899-904): the shifted plan
must not change a single output byte."""
import ctypes

import numpy as np
import pytest

import zkmetal_amd as zk
from oracle import bn254 as B
from oracle import ref
from helpers import result_affine

pytestmark = pytest.mark.gpu

SEED = 0x5A1F
N_SRS = 1 << 18


@pytest.fixture(scope="module")
def pair():
    """(window-shifted engine running every MSM shifted, plain engine)"""
    sh, pl = zk.MetalMSM(), zk.MetalMSM()
    sh.set_window_shift(sh.SHIFT_ALL)
    pl.set_window_shift(pl.SHIFT_OFF)
    try:
        sh.srs_generate(SEED, N_SRS)
        pl.srs_generate(SEED, N_SRS)
        assert sh.srs_window_shift() == 16 and pl.srs_window_shift() == 0
        yield sh, pl
    finally:
        sh.close()
        pl.close()


@pytest.fixture(scope="module")
def ks():
    return zk.gen_scalars(SEED, N_SRS, 0).tobytes()


def _closed(ks, sc, n):
    return B.mul(B.G, ref.scalar_dot(sc.tobytes(), ks[:32 * n], n))


def test_single_msms_small_ragged_and_full(pair, ks):
    sh, pl = pair
    for n in (1 << 13, 1 << 14, (1 << 14) + 37, 100003, N_SRS):
        print(f"single n={n}", flush=True)
        sc = zk.gen_scalars(SEED + n, n, 1)
        a = sh.msm_batch_resident(sc, [n])
        b = pl.msm_batch_resident(sc, [n])
        assert a == b, f"n={n}"
    assert result_affine(a[0]) == _closed(ks, sc, N_SRS)


def test_batch_mixing_shifted_and_plain_msms(pair, ks):
    sh, pl = pair
    # the cost estimate (auto) keeps the small MSMs plain and shifts the rest
    counts = [5, 1 << 12, 70000, 1, 3000, 1 << 16, 0, 123457, 9000]
    sc = zk.gen_scalars(SEED + 1, sum(counts), 1)
    sc[:5] = 0
    b = pl.msm_batch_resident(sc, counts)
    sh.set_window_shift(sh.SHIFT_AUTO)
    try:
        a = sh.msm_batch_resident(sc, counts)
        assert a == b
        sh.set_tail_mode(sh.TAIL_DEVICE)  # device combine: MSMs of >= 4096 points shift
        assert sh.msm_batch_resident(sc, counts) == b
    finally:
        sh.set_tail_mode(sh.TAIL_AUTO)
        sh.set_window_shift(sh.SHIFT_ALL)
    assert sh.msm_batch_resident(sc, counts) == b
    off = sum(counts[:7])
    assert result_affine(a[7]) == _closed(ks, sc[off:off + counts[7]], counts[7])
    assert result_affine(a[0]) is None and a[6] == bytes(96)


def test_every_part_split_of_one_msm(pair):
    sh, pl = pair
    n = N_SRS
    sc = zk.gen_scalars(SEED + 2, n, 0)
    want = pl.msm_batch_resident(sc, [n])
    try:
        for p in (1, 2, 3, 4, 8):
            sh.set_pipeline_parts(p)
            assert sh.msm_batch_resident(sc, [n]) == want, f"parts={p}"
    finally:
        sh.set_pipeline_parts(0)


def test_device_tail_and_graph_replays(pair):
    sh, pl = pair
    counts = [1 << 16, 40000, 3, 1 << 15]
    sc = zk.gen_scalars(SEED + 3, sum(counts), 0)
    ds = zk.DeviceBuffer(sc.nbytes)
    ds.upload(sc)
    try:
        want = pl.msm_batch_resident(ds, counts)
        for mode in (sh.TAIL_HOST, sh.TAIL_DEVICE):
            sh.set_tail_mode(mode)
            assert sh.msm_batch_resident(ds, counts) == want
        sh.set_tail_mode(sh.TAIL_AUTO)
        sh.set_graph_mode(True)
        try:
            assert all(sh.msm_batch_resident(ds, counts) == want for _ in range(3))
        finally:
            sh.set_graph_mode(False)
    finally:
        sh.set_tail_mode(sh.TAIL_AUTO)
        ds.free()


def test_montgomery_form_coefficients(pair):
    sh, pl = pair
    counts = [50000, 1 << 14]
    co = zk.gen_scalars(SEED + 4, sum(counts), 0)  # any 256-bit words: taken as Montgomery form
    assert sh.kzg_commit(co, counts) == pl.kzg_commit(co, counts)


def test_allocation_free_entry_point(pair):
    sh, pl = pair
    counts = [30000, 20000, 7]
    sc = zk.gen_scalars(SEED + 5, sum(counts), 1)
    ds = zk.DeviceBuffer(sc.nbytes)
    ds.upload(sc)
    try:
        c = np.asarray(counts, dtype=np.uint32)
        out = np.zeros(len(counts) * 96, dtype=np.uint8)
        rc = zk.lib().zkmsm_msm_batch_resident_device_noalloc(
            sh.handle, ds.ptr, ctypes.c_void_p(c.ctypes.data), len(counts),
            ctypes.c_void_p(out.ctypes.data))
        assert rc == zk.SUCCESS
        got = [out[i * 96:(i + 1) * 96].tobytes() for i in range(len(counts))]
        assert got == pl.msm_batch_resident(ds, counts)
    finally:
        ds.free()


def test_request_server_on_a_shifted_table(pair):
    sh, pl = pair
    reqs = [zk.gen_scalars(SEED + 10 + i, n, 1) for i, n in enumerate([1 << 17, 5000, 60000, 1 << 14])]
    srv = zk.Server(sh, slots=2, max_batch=4)
    try:
        got = [srv.wait(t)[0] for t in [srv.submit(r) for r in reqs]]
    finally:
        srv.close()
    for r, g in zip(reqs, got):
        assert g == pl.msm_batch_resident(r, [len(r)])[0]


def test_forced_window_width_runs_plain(pair):
    sh, pl = pair
    n = 1 << 16
    sc = zk.gen_scalars(SEED + 6, n, 0)
    for e in (sh, pl):
        e.set_window_bits(16)
    try:
        assert sh.msm_batch_resident(sc, [n]) == pl.msm_batch_resident(sc, [n])
    finally:
        for e in (sh, pl):
            e.set_window_bits(0)


def test_uploaded_srs_with_identity_bases():
    n, seed = 40000, 0x1D0
    gen = zk.MetalMSM()
    try:
        gen.srs_generate(seed, n)
        pts = bytearray(gen.srs_download(n).tobytes())
    finally:
        gen.close()
    ids = [0, 1, 777, 2047, 2048, n - 1]
    for i in ids:
        pts[64 * i:64 * i + 64] = bytes(64)
    sc = zk.gen_scalars(seed + 1, n + 30000, 0)
    kk = bytearray(zk.gen_scalars(seed, n, 0).tobytes())
    for i in ids:
        kk[32 * i:32 * i + 32] = bytes(32)  # (0, 0) is the identity: k_i = 0
    res = []
    for on in (True, False):
        e = zk.MetalMSM()
        try:
            e.set_window_shift(e.SHIFT_ALL if on else e.SHIFT_OFF)
            e.srs_upload(bytes(pts))
            assert e.srs_window_shift() == (16 if on else 0)
            res.append(e.msm_batch_resident(sc, [n, 30000]))
        finally:
            e.close()
    assert res[0] == res[1]
    assert result_affine(res[0][0]) == _closed(bytes(kk), sc[:n], n)
