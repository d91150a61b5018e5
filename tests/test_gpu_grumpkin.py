"""GPU parity tests of the Grumpkin MSM (zkmsm_grumpkin_msm[_batch]) and of
the Fr-field 29-bit arithmetic it runs on, against the Python oracle
(oracle/grumpkin.py).

Grumpkin: y^2 = x^3 - 17 over BN254 Fr, order BN254 p
(Sources/zkMetal/Curve/GrumpkinCurve.swift, MSM/GrumpkinMSMEngine.swift:246).
Bar: the canonical affine result equals the closed form
(sum s_i k_i mod p) * G for bases k_i * G, bit for bit, at every size; edge
cases as for BN254 G1 (identity bases, zero / >= p scalars, cancellation,
repeated bases, zero counts in a batch, every window width)."""
import random

import pytest

import zkmetal_amd as zk
from oracle import grumpkin as G

pytestmark = pytest.mark.gpu


def fe(x):
    return int(x).to_bytes(32, "little")


def rd(buf, i):
    return int.from_bytes(bytes(buf[32 * i:32 * i + 32]), "little")


def sc_bytes(ss):
    return b"".join(fe(s) for s in ss)


@pytest.fixture(scope="module")
def bases():
    """P_i = (i + 1) G for i < 5000 (one addition each) and their logs."""
    pts, acc = [], None
    for _ in range(5000):
        acc = G.add(acc, G.G)
        pts.append(acc)
    return pts, list(range(1, 5001))


def pbytes(pts):
    return b"".join(G.point_to_bytes(p) for p in pts)


def test_fr_field_mul_sqr(engine):
    rnd = random.Random(21)
    edge = [0, 1, 2, G.R - 1, G.R - 2, G.MONT_R, G.R >> 1]
    a = edge + [rnd.randrange(G.R) for _ in range(1000)]
    b = [rnd.randrange(G.R) for _ in a]
    A = b"".join(fe(G.to_mont(x)) for x in a)
    Bb = b"".join(fe(G.to_mont(x)) for x in b)
    mul = engine.test_fp(8, A, Bb)
    sq = engine.test_fp(9, A, A)
    for i, (x, y) in enumerate(zip(a, b)):
        assert rd(mul, i) == G.to_mont(x * y % G.R), i
        assert rd(sq, i) == G.to_mont(x * x % G.R), i


def _xyzz(pt, z):
    if pt is None:
        return bytes(128)
    zz, zzz = z * z % G.R, z * z * z % G.R
    return b"".join(fe(G.to_mont(v)) for v in (pt[0] * zz % G.R, pt[1] * zzz % G.R, zz, zzz))


def _xyzz_affine(buf, i):
    X, Y, ZZ, ZZZ = (G.from_mont(rd(buf, 4 * i + k)) for k in range(4))
    if ZZ == 0:
        return None
    return (X * pow(ZZ, -1, G.R) % G.R, Y * pow(ZZZ, -1, G.R) % G.R)


def test_grumpkin_point_ops(engine):
    rnd = random.Random(22)
    pts = [G.mul(G.G, rnd.randrange(1, G.ORDER)) for _ in range(12)]
    cases = [(pts[i], pts[i + 1]) for i in range(len(pts) - 1)]
    P = pts[0]
    cases += [(P, P), (P, G.neg(P)), (P, None), (None, P), (None, None), (G.G, G.G)]
    Pb = b"".join(_xyzz(p, rnd.randrange(1, G.R)) for p, _ in cases)
    Qb = b"".join(_xyzz(q, rnd.randrange(1, G.R)) for _, q in cases)
    Qa = b"".join(_xyzz(q, 1) for _, q in cases)
    add = engine.test_xyzz(6, Pb, Qb)
    dbl = engine.test_xyzz(7, Pb, Qb)
    madd = engine.test_xyzz(8, Pb, Qa)
    for i, (p, q) in enumerate(cases):
        assert _xyzz_affine(add, i) == G.add(p, q), i
        assert _xyzz_affine(dbl, i) == G.add(p, p), i
        assert _xyzz_affine(madd, i) == G.add(p, q), i


@pytest.mark.parametrize("n", [1, 2, 7, 100, 2049, 5000])
def test_msm_closed_form(engine, bases, n):
    pts, ks = bases
    rnd = random.Random(n)
    ss = [rnd.randrange(G.ORDER) for _ in range(n)]
    r = engine.grumpkin_msm(pbytes(pts[:n]), sc_bytes(ss))
    assert G.jacobian_bytes_to_affine(r) == G.msm_closed_form(ks[:n], ss)
    assert r[64:96] in (fe(G.to_mont(1)), bytes(32))  # normalised


def test_msm_edge_scalars_and_identity_bases(engine, bases):
    pts, ks = bases
    n = 700
    rnd = random.Random(23)
    ss = [rnd.randrange(2**256) for _ in range(n)]
    ss[:8] = [0, 1, G.ORDER - 1, G.ORDER, G.ORDER + 1, 2**256 - 1, 2**254, G.R]
    use = list(pts[:n])
    kk = list(ks[:n])
    for i in (3, 64, 65, 400, n - 1):
        use[i] = None
        kk[i] = 0
    r = engine.grumpkin_msm(pbytes(use), sc_bytes(ss))
    assert G.jacobian_bytes_to_affine(r) == G.msm_closed_form(kk, ss)


def test_msm_cancellation_and_repeats(engine, bases):
    pts, ks = bases
    # every base 4 times with scalars 1..4 -> 10 P each (correctness_tests.swift:1756-1768)
    idx = [i % 300 for i in range(1200)]
    ss = [(i // 300) % 4 + 1 for i in range(1200)]
    r = engine.grumpkin_msm(pbytes([pts[i] for i in idx]), sc_bytes(ss))
    assert G.jacobian_bytes_to_affine(r) == G.msm_closed_form([ks[i] for i in idx], ss)
    # s P + (p - s) P = O
    half = [pts[i] for i in range(500)]
    ss2 = [7] * 500 + [G.ORDER - 7] * 500
    r2 = engine.grumpkin_msm(pbytes(half + half), sc_bytes(ss2))
    assert G.jacobian_bytes_to_affine(r2) is None
    assert r2 == fe(G.to_mont(1)) * 2 + bytes(32)


def test_every_window_width(engine, bases):
    pts, ks = bases
    n = 300
    rnd = random.Random(24)
    ss = [rnd.randrange(G.ORDER) for _ in range(n)]
    want = G.msm_closed_form(ks[:n], ss)
    try:
        for c in range(4, 17):
            engine.set_window_bits(c)
            r = engine.grumpkin_msm(pbytes(pts[:n]), sc_bytes(ss))
            assert G.jacobian_bytes_to_affine(r) == want, c
        # 255-bit scalars need c >= 4 (at most 64 windows)
        engine.set_window_bits(3)
        with pytest.raises(zk.MSMError):
            engine.grumpkin_msm(pbytes(pts[:n]), sc_bytes(ss))
    finally:
        engine.set_window_bits(0)


def test_batch_with_zero_count_and_alone_equals_in_batch(engine, bases):
    pts, ks = bases
    counts = [10, 0, 2000, 1, 777]
    rnd = random.Random(25)
    ss = [rnd.randrange(G.ORDER) for _ in range(sum(counts))]
    cat_pts, off = [], 0
    for c in counts:
        cat_pts += pts[:c]
    res = engine.grumpkin_msm_batch(pbytes(cat_pts), sc_bytes(ss), counts)
    assert res[1] == bytes(96)
    off = 0
    for j, c in enumerate(counts):
        if c:
            assert G.jacobian_bytes_to_affine(res[j]) == G.msm_closed_form(ks[:c], ss[off:off + c]), j
            alone = engine.grumpkin_msm(pbytes(pts[:c]), sc_bytes(ss[off:off + c]))
            assert alone == res[j]
        off += c


def test_bn254_and_grumpkin_interleaved_on_one_engine(engine, bases):
    """Both curves on one engine, alternately (shared workspace)."""
    from oracle import bn254 as Bn
    from helpers import closed_form, result_affine, scalars_to_ints

    pts, ks = bases
    e2 = zk.MetalMSM()
    try:
        e2.srs_generate(0x61, 1000)
        bn_pts = e2.srs_download(1000)
    finally:
        e2.close()
    bn_ks = scalars_to_ints(zk.gen_scalars(0x61, 1000, 0))
    sc = zk.gen_scalars(0x62, 1000, 0)
    ints = scalars_to_ints(sc)
    for _ in range(2):
        r1 = engine.msm(bn_pts, sc)
        r2 = engine.grumpkin_msm(pbytes(pts[:1000]), sc.tobytes())
        assert result_affine(r1) == closed_form(bn_ks, ints)
        assert G.jacobian_bytes_to_affine(r2) == G.msm_closed_form(ks[:1000], ints)
    assert Bn.on_curve(result_affine(r1))


def test_invalid_inputs(engine):
    with pytest.raises(zk.MSMError):
        engine.grumpkin_msm(b"", b"")
    with pytest.raises(zk.MSMError):
        engine.grumpkin_msm(bytes(64), bytes(64))


def _gen_device(engine, seed, n):
    d = zk.DeviceBuffer(64 * n)
    engine.grumpkin_points_generate(seed, n, d)
    return d


def test_points_generate_matches_scalar_multiples(engine):
    from helpers import scalars_to_ints
    n = 3000
    d = _gen_device(engine, 0x71, n)
    try:
        pts = d.download()
    finally:
        d.free()
    ks = scalars_to_ints(zk.gen_scalars(0x71, n, 0))
    rnd = random.Random(26)
    for i in list(range(20)) + rnd.sample(range(n), 30):
        b = pts[64 * i:64 * i + 64]
        want = G.mul(G.G, ks[i])
        assert b == G.point_to_bytes(want), i
        assert G.on_curve(want)


def test_device_entry_equals_host_entry_and_c_oracle(engine):
    """2^16 bases on the device: device-buffer MSM == host-buffer MSM, and
    both equal the C restatement of grumpkin_pippenger_msm."""
    from oracle import ref
    n = 1 << 16
    d = _gen_device(engine, 0x72, n)
    sc = zk.gen_scalars(0x73, n, 0)
    ds = zk.DeviceBuffer(32 * n)
    try:
        ds.upload(sc)
        pts = d.download()
        dev = engine.grumpkin_msm_batch_device(d, ds, [n])[0]
    finally:
        d.free()
        ds.free()
    host = engine.grumpkin_msm(pts, sc.tobytes())
    assert dev == host
    cpu = ref.pippenger(pts, sc.tobytes(), n, curve="grumpkin")
    assert G.jacobian_bytes_to_affine(dev) == G.jacobian_bytes_to_affine(cpu)


def test_device_entry_closed_form_at_2_20_and_batch(engine):
    from oracle import ref
    n = 1 << 20
    seed = 0x74
    d = _gen_device(engine, seed, n)
    ks = zk.gen_scalars(seed, n, 0)
    sc = zk.gen_scalars(0x75, n, 0)
    ds = zk.DeviceBuffer(32 * n)
    try:
        ds.upload(sc)
        r = engine.grumpkin_msm_batch_device(d, ds, [n])[0]
        # a batch over the same buffers: MSM j covers points/scalars of its slice
        counts = [1000, 0, n - 1000]
        rb = engine.grumpkin_msm_batch_device(d, ds, counts)
    finally:
        d.free()
        ds.free()
    want = G.mul(G.G, ref.scalar_dot(sc.tobytes(), ks.tobytes(), n, curve="grumpkin"))
    assert G.jacobian_bytes_to_affine(r) == want
    assert rb[1] == bytes(96)
    lo = G.mul(G.G, ref.scalar_dot(sc[:1000].tobytes(), ks[:1000].tobytes(), 1000, curve="grumpkin"))
    assert G.jacobian_bytes_to_affine(rb[0]) == lo
    assert G.add(G.jacobian_bytes_to_affine(rb[0]), G.jacobian_bytes_to_affine(rb[2])) == want


@pytest.mark.parametrize("parts,per", [(3, 4), (1, 64), (8, 16)])
def test_schedule_tables_keep_results(bases, parts, per):
    """Pipeline parts and the bucket-reduction tile change how a Grumpkin call
    is launched, never its result (closed form, single and batched)."""
    pts, ks = bases
    counts = [3000, 1, 2000]
    rnd = random.Random(26 + parts)
    ss = [rnd.randrange(G.ORDER) for _ in range(sum(counts))]
    cat = []
    for c in counts:
        cat += pts[:c]
    eng = zk.MetalMSM()
    try:
        eng.set_graph_mode(eng.GRAPH_OFF)
        for batch in (False, True):
            eng.set_schedule_table([parts] * 33, [per] * 33, batch=batch)
        res = eng.grumpkin_msm_batch(pbytes(cat), sc_bytes(ss), counts)
        assert eng.last_timing()["buckets_per_group"] == per
        one = eng.grumpkin_msm(pbytes(pts[:counts[0]]), sc_bytes(ss[:counts[0]]))
        assert eng.last_timing()["buckets_per_group"] == per
        assert one == res[0]
        off = 0
        for j, c in enumerate(counts):
            assert G.jacobian_bytes_to_affine(res[j]) == G.msm_closed_form(ks[:c], ss[off:off + c]), j
            off += c
    finally:
        eng.close()
