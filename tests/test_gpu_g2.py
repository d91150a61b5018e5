"""GPU parity tests of the BN254 G2 MSM (zkmsm_bn254_g2_msm[_batch][_device])
and of the Fp2 / G2 arithmetic it runs on, against the Python oracle
(oracle/bn254_g2.py).

The reference stores no expected G2 MSM outputs; its own checks
(Sources/zkbench/g2_msm_bench.swift:44-125) compare the GPU with a CPU sum
over bases i * G2: 16 points with scalars 1..16, 256 points with the harness
LCG scalars, 128 points at window width 4.  Those cases are here, with the
expected value from the closed form (sum s_i k_i mod r) * G2, plus edge
scalars, identity bases, repeats, cancellation, every window width, batches
with a zero count, on-device generated bases and 2^16 / 2^18 sizes, and a
skewed-scalar case against the C G2 Pippenger (oracle/bn254_ref.c)."""
import random

import pytest

import zkmetal_amd as zk
from oracle import bn254_g2 as G

pytestmark = pytest.mark.gpu

R = G.R_ORDER


def fe(x):
    return int(x).to_bytes(32, "little")


def sc_bytes(ss):
    return b"".join(fe(s) for s in ss)


def pbytes(pts):
    return b"".join(G.point_to_bytes(p) for p in pts)


@pytest.fixture(scope="module")
def bases():
    """P_i = (i + 1) G2 for i < 1200 and their logs."""
    pts, acc = [], None
    for _ in range(1200):
        acc = G.add(acc, G.G2)
        pts.append(acc)
    return pts, list(range(1, 1201))


def lcg_scalars(n, seed=0xDEADBEEFCAFEBABE):
    """the reference harness's scalars (g2_msm_bench.swift:30-40)"""
    out, rng = [], seed
    for _ in range(n):
        v = 0
        for j in range(8):
            rng = (rng * 6364136223846793005 + 1442695040888963407) % 2**64
            v |= (rng >> 32) << (32 * j)
        out.append(v)
    return out


# ---- primitives: 256-byte host-form XYZZ points (c0 record, c1 record)

def _host_xyzz(pt, z):
    """XYZZ representative (X, Y, ZZ, ZZZ) of an affine point with ZZ = z^2"""
    if pt is None:
        return bytes(256)
    zz = G.mul2(z, z)
    zzz = G.mul2(zz, z)
    X, Y = G.mul2(pt[0], zz), G.mul2(pt[1], zzz)
    c = (X, Y, zz, zzz)
    return b"".join(fe(G.to_mont(v[0])) for v in c) + b"".join(fe(G.to_mont(v[1])) for v in c)


def _aff(pt):
    return _host_xyzz(pt, (1, 0))


def _rd(buf, i, k):
    """component k (0..7: X0 Y0 ZZ0 ZZZ0 X1 Y1 ZZ1 ZZZ1) of point i"""
    o = 256 * i + 32 * k
    return G.from_mont(int.from_bytes(bytes(buf[o:o + 32]), "little"))


def _xyzz_affine(buf, i):
    X = (_rd(buf, i, 0), _rd(buf, i, 4))
    Y = (_rd(buf, i, 1), _rd(buf, i, 5))
    ZZ = (_rd(buf, i, 2), _rd(buf, i, 6))
    ZZZ = (_rd(buf, i, 3), _rd(buf, i, 7))
    if ZZ == G.ZERO2:
        return None
    return (G.mul2(X, G.inv2(ZZ)), G.mul2(Y, G.inv2(ZZZ)))


def _rz(rnd):
    return (rnd.randrange(1, G.P), rnd.randrange(G.P))


def test_fp2_field_ops(engine):
    rnd = random.Random(31)
    xs = [(0, 0), (1, 0), (0, 1), (G.P - 1, G.P - 1)] + [(rnd.randrange(G.P), rnd.randrange(G.P)) for _ in range(300)]
    ys = [(rnd.randrange(G.P), rnd.randrange(G.P)) for _ in xs]

    def pack(vals):
        return b"".join(fe(G.to_mont(v[0])) + bytes(96) + fe(G.to_mont(v[1])) + bytes(96) for v in vals)

    A, Bv = pack(xs), pack(ys)
    mul = engine.test_g2(4, A, Bv)
    sqr = engine.test_g2(5, A, A)
    inv = engine.test_g2(6, A, A)
    for i, (x, y) in enumerate(zip(xs, ys)):
        assert (_rd(mul, i, 0), _rd(mul, i, 4)) == G.mul2(x, y), i
        assert (_rd(sqr, i, 0), _rd(sqr, i, 4)) == G.mul2(x, x), i
        want = G.ZERO2 if x == G.ZERO2 else G.inv2(x)
        assert (_rd(inv, i, 0), _rd(inv, i, 4)) == want, i


def test_g2_point_ops(engine):
    rnd = random.Random(32)
    pts = [G.mul(G.G2, rnd.randrange(1, R)) for _ in range(10)]
    P0 = pts[0]
    cases = [(pts[i], pts[i + 1]) for i in range(9)]
    cases += [(P0, P0), (P0, G.neg(P0)), (P0, None), (None, P0), (None, None), (G.G2, G.G2)]
    Pb = b"".join(_host_xyzz(p, _rz(rnd)) for p, _ in cases)
    Qb = b"".join(_host_xyzz(q, _rz(rnd)) for _, q in cases)
    Qa = b"".join(_aff(q) for _, q in cases)
    add = engine.test_g2(0, Pb, Qb)
    dbl = engine.test_g2(1, Pb, Qb)
    madd = engine.test_g2(2, Pb, Qa)
    for i, (p, q) in enumerate(cases):
        assert _xyzz_affine(add, i) == G.add(p, q), i
        assert _xyzz_affine(dbl, i) == G.add(p, p), i
        assert _xyzz_affine(madd, i) == G.add(p, q), i
    # the bucket-loop addition (q never the identity there)
    run_cases = [(p, q) for p, q in cases if q is not None]
    Pr = b"".join(_host_xyzz(p, _rz(rnd)) for p, _ in run_cases)
    Qr = b"".join(_aff(q) for _, q in run_cases)
    run = engine.test_g2(3, Pr, Qr)
    for i, (p, q) in enumerate(run_cases):
        assert _xyzz_affine(run, i) == G.add(p, q), i


def _check(r, want):
    assert G.jacobian_bytes_to_affine(r) == want
    one = fe(G.to_mont(1))
    assert r[128:192] in (one + bytes(32), bytes(64))  # normalised


@pytest.mark.parametrize("n", [1, 2, 7, 65, 300, 1200])
def test_msm_closed_form(engine, bases, n):
    pts, ks = bases
    rnd = random.Random(n)
    ss = [rnd.randrange(R) for _ in range(n)]
    _check(engine.g2_msm(pbytes(pts[:n]), sc_bytes(ss)), G.msm_closed_form(ks[:n], ss))


def test_reference_harness_cases(engine, bases):
    pts, ks = bases
    # 16 points, scalars 1..16
    _check(engine.g2_msm(pbytes(pts[:16]), sc_bytes(range(1, 17))), G.msm_closed_form(ks[:16], range(1, 17)))
    # 256 points, the harness LCG scalars (full 256-bit, reduced mod r)
    ss = lcg_scalars(256)
    _check(engine.g2_msm(pbytes(pts[:256]), sc_bytes(ss)), G.msm_closed_form(ks[:256], ss))
    # 128 points at window width 4
    try:
        engine.set_window_bits(4)
        _check(engine.g2_msm(pbytes(pts[:128]), sc_bytes(ss[:128])), G.msm_closed_form(ks[:128], ss[:128]))
    finally:
        engine.set_window_bits(0)


def test_edge_scalars_identity_bases_repeats_cancellation(engine, bases):
    pts, ks = bases
    n = 400
    rnd = random.Random(33)
    ss = [rnd.randrange(2**256) for _ in range(n)]
    ss[:8] = [0, 1, R - 1, R, R + 1, 2**256 - 1, 2**253, G.P]
    use, kk = list(pts[:n]), list(ks[:n])
    for i in (5, 64, 65, n - 1):
        use[i], kk[i] = None, 0
    _check(engine.g2_msm(pbytes(use), sc_bytes(ss)), G.msm_closed_form(kk, ss))
    # every base 4 times with scalars 1..4 -> 10 P each
    idx = [i % 100 for i in range(400)]
    s4 = [(i // 100) % 4 + 1 for i in range(400)]
    _check(engine.g2_msm(pbytes([pts[i] for i in idx]), sc_bytes(s4)),
           G.msm_closed_form([ks[i] for i in idx], s4))
    # s P + (r - s) P = O
    half = pts[:150]
    r0 = engine.g2_msm(pbytes(half + half), sc_bytes([11] * 150 + [R - 11] * 150))
    assert G.jacobian_bytes_to_affine(r0) is None
    one = fe(G.to_mont(1))
    assert r0 == one + bytes(32) + one + bytes(32) + bytes(64)


def test_every_window_width(engine, bases):
    pts, ks = bases
    n = 200
    rnd = random.Random(34)
    ss = [rnd.randrange(R) for _ in range(n)]
    want = G.msm_closed_form(ks[:n], ss)
    try:
        for c in range(4, 17):
            engine.set_window_bits(c)
            assert G.jacobian_bytes_to_affine(engine.g2_msm(pbytes(pts[:n]), sc_bytes(ss))) == want, c
        engine.set_window_bits(3)  # 255-bit scalars need c >= 4
        with pytest.raises(zk.MSMError):
            engine.g2_msm(pbytes(pts[:n]), sc_bytes(ss))
    finally:
        engine.set_window_bits(0)


def test_batch_with_zero_count_and_alone_equals_in_batch(engine, bases):
    pts, ks = bases
    counts = [10, 0, 700, 1, 300]
    rnd = random.Random(35)
    ss = [rnd.randrange(R) for _ in range(sum(counts))]
    cat = []
    for c in counts:
        cat += pts[:c]
    res = engine.g2_msm_batch(pbytes(cat), sc_bytes(ss), counts)
    assert res[1] == bytes(192)
    off = 0
    for j, c in enumerate(counts):
        if c:
            assert G.jacobian_bytes_to_affine(res[j]) == G.msm_closed_form(ks[:c], ss[off:off + c]), j
            assert engine.g2_msm(pbytes(pts[:c]), sc_bytes(ss[off:off + c])) == res[j]
        off += c


def _gen_device(engine, seed, n):
    d = zk.DeviceBuffer(128 * n)
    engine.g2_points_generate(seed, n, d)
    return d


def test_points_generate_matches_scalar_multiples(engine):
    from helpers import scalars_to_ints
    n = 2000
    d = _gen_device(engine, 0x81, n)
    try:
        pts = d.download()
    finally:
        d.free()
    ks = scalars_to_ints(zk.gen_scalars(0x81, n, 0))
    rnd = random.Random(36)
    for i in list(range(6)) + rnd.sample(range(n), 10):
        assert pts[128 * i:128 * i + 128] == G.point_to_bytes(G.mul(G.G2, ks[i])), i


def test_device_entry_equals_host_entry_at_2_16_and_closed_form_at_2_18(engine):
    from oracle import ref
    n = 1 << 16
    d = _gen_device(engine, 0x82, n)
    sc = zk.gen_scalars(0x83, n, 0)
    ds = zk.DeviceBuffer(32 * n)
    try:
        ds.upload(sc)
        dev = engine.g2_msm_batch_device(d, ds, [n])[0]
        host = engine.g2_msm(d.download(), sc.tobytes())
    finally:
        d.free()
        ds.free()
    assert dev == host
    ks = zk.gen_scalars(0x82, n, 0)
    assert G.jacobian_bytes_to_affine(dev) == G.mul(G.G2, ref.scalar_dot(sc.tobytes(), ks.tobytes(), n))
    n = 1 << 18
    d = _gen_device(engine, 0x84, n)
    sc = zk.gen_scalars(0x85, n, 0)
    ds = zk.DeviceBuffer(32 * n)
    try:
        ds.upload(sc)
        r = engine.g2_msm_batch_device(d, ds, [n])[0]
    finally:
        d.free()
        ds.free()
    ks = zk.gen_scalars(0x84, n, 0)
    assert G.jacobian_bytes_to_affine(r) == G.mul(G.G2, ref.scalar_dot(sc.tobytes(), ks.tobytes(), n))


def test_g1_grumpkin_g2_interleaved_on_one_engine(engine, bases):
    from oracle import grumpkin as Gk
    from helpers import closed_form, result_affine, scalars_to_ints
    pts, ks = bases
    e2 = zk.MetalMSM()
    try:
        e2.srs_generate(0x86, 300)
        g1 = e2.srs_download(300)
    finally:
        e2.close()
    g1_ks = scalars_to_ints(zk.gen_scalars(0x86, 300, 0))
    gk_pts, acc = [], None
    for _ in range(300):
        acc = Gk.add(acc, Gk.G)
        gk_pts.append(acc)
    sc = zk.gen_scalars(0x87, 300, 0)
    ints = scalars_to_ints(sc)
    for _ in range(2):
        a = engine.msm(g1, sc)
        b = engine.g2_msm(pbytes(pts[:300]), sc.tobytes())
        c = engine.grumpkin_msm(b"".join(Gk.point_to_bytes(p) for p in gk_pts), sc.tobytes())
        assert result_affine(a) == closed_form(g1_ks, ints)
        assert G.jacobian_bytes_to_affine(b) == G.msm_closed_form(ks[:300], ints)
        assert Gk.jacobian_bytes_to_affine(c) == Gk.msm_closed_form(range(1, 301), ints)


def test_invalid_inputs(engine):
    with pytest.raises(zk.MSMError):
        engine.g2_msm(b"", b"")
    with pytest.raises(zk.MSMError):
        engine.g2_msm(bytes(128), bytes(64))


def test_skewed_scalars_against_the_c_pippenger_at_2_14(engine):
    """Every scalar from a pool of 5 values (a few buckets per window hold
    thousands of entries each): the G2 MSM against the C restatement of the
    G2 Pippenger on the downloaded bases."""
    import numpy as np
    from oracle import ref
    n = 1 << 14
    d = _gen_device(engine, 0x88, n)
    try:
        pts = d.download(128 * n)
    finally:
        d.free()
    rnd = random.Random(37)
    pool = [rnd.randrange(R) for _ in range(4)] + [R - 1]
    ss = [pool[rnd.randrange(5)] for _ in range(n)]
    sc = np.frombuffer(sc_bytes(ss), dtype=np.uint8)
    got = engine.g2_msm(pts, sc.tobytes())
    want = ref.g2_pippenger(pts, sc.tobytes(), n, 0, 8)
    assert G.jacobian_bytes_to_affine(got) == G.jacobian_bytes_to_affine(want)


@pytest.mark.parametrize("parts,per", [(3, 4), (1, 64)])
def test_schedule_tables_keep_results(bases, parts, per):
    """Pipeline parts and the bucket-reduction tile change how a G2 call is
    launched, never its result (closed form, single and batched)."""
    pts, ks = bases
    counts = [700, 1, 1100]
    rnd = random.Random(36 + parts)
    ss = [rnd.randrange(R) for _ in range(sum(counts))]
    cat = []
    for c in counts:
        cat += pts[:c]
    eng = zk.MetalMSM()
    try:
        eng.set_graph_mode(eng.GRAPH_OFF)
        for batch in (False, True):
            eng.set_schedule_table([parts] * 33, [per] * 33, batch=batch)
        res = eng.g2_msm_batch(pbytes(cat), sc_bytes(ss), counts)
        assert eng.last_timing()["buckets_per_group"] == per
        one = eng.g2_msm(pbytes(pts[:counts[2]]), sc_bytes(ss[counts[0] + counts[1]:]))
        assert one == res[2]
        off = 0
        for j, c in enumerate(counts):
            assert G.jacobian_bytes_to_affine(res[j]) == G.msm_closed_form(ks[:c], ss[off:off + c]), j
            off += c
    finally:
        eng.close()
