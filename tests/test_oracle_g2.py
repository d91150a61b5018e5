"""CPU pinning of the BN254 G2 oracle (oracle/bn254_g2.py) against the
reference's generator and the group structure (no GPU)."""
import random

from oracle import bn254_g2 as G


def test_fp2_field():
    rnd = random.Random(1)
    u = (0, 1)
    assert G.mul2(u, u) == G.f2(-1)
    for _ in range(30):
        a = (rnd.randrange(G.P), rnd.randrange(G.P))
        b = (rnd.randrange(G.P), rnd.randrange(G.P))
        c = (rnd.randrange(G.P), rnd.randrange(G.P))
        assert G.mul2(G.mul2(a, b), c) == G.mul2(a, G.mul2(b, c))
        assert G.mul2(a, G.add2(b, c)) == G.add2(G.mul2(a, b), G.mul2(a, c))
        assert G.mul2(a, G.inv2(a)) == G.ONE2
    # B2 * (9 + u) = 3
    assert G.mul2(G.B2, G.XI) == (3, 0)


def test_generator_on_twist_with_order_r():
    assert G.on_curve(G.G2)
    assert G.mul(G.G2, G.R_ORDER) is None
    assert G.add(G.G2, G.G2) is not None
    # a planted mistake in the twist coefficient is caught
    x, y = G.G2
    assert G.sub2(G.mul2(y, y), G.add2(G.mul2(G.mul2(x, x), x), G.add2(G.B2, G.ONE2))) != G.ZERO2
    # and so is a wrong generator coordinate
    assert not G.on_curve((x, G.add2(y, G.ONE2)))


def test_group_law():
    rnd = random.Random(2)
    pts = [G.mul(G.G2, rnd.randrange(1, G.R_ORDER)) for _ in range(4)]
    for p in pts:
        assert G.on_curve(p)
    a, b, c = pts[:3]
    assert G.add(G.add(a, b), c) == G.add(a, G.add(b, c))
    assert G.add(a, b) == G.add(b, a)
    assert G.add(a, G.neg(a)) is None
    acc = None
    for _ in range(7):
        acc = G.add(acc, G.G2)
    assert acc == G.mul(G.G2, 7)
    k1, k2 = rnd.randrange(G.R_ORDER), rnd.randrange(G.R_ORDER)
    assert G.add(G.mul(G.G2, k1), G.mul(G.G2, k2)) == G.mul(G.G2, k1 + k2)
    assert G.mul(G.G2, G.R_ORDER - 1) == G.neg(G.G2)


def test_closed_form_equals_brute_force():
    rnd = random.Random(3)
    ks = [rnd.randrange(1, G.R_ORDER) for _ in range(4)]
    ss = [rnd.randrange(0, 2**256) for _ in range(4)]
    brute = None
    for k, s in zip(ks, ss):
        brute = G.add(brute, G.mul(G.mul(G.G2, k), s))
    assert brute == G.msm_closed_form(ks, ss)
    # the reference bench's small case: points (i+1) G2, scalars i+1
    # (g2_msm_bench.swift:44-69): sum (i+1)^2 = 1496 for 16 points
    assert G.msm_closed_form(range(1, 17), range(1, 17)) == G.mul(G.G2, 1496)


def test_byte_forms_round_trip():
    p = G.mul(G.G2, 99)
    b = G.point_to_bytes(p)
    assert G.bytes_to_point(b) == p
    one = G.fe(G.to_mont(1))
    zero = bytes(32)
    jac = b + one + zero
    assert G.jacobian_bytes_to_affine(jac) == p
    assert G.jacobian_bytes_to_affine(one + zero + one + zero + zero + zero) is None
    assert G.bytes_to_point(bytes(128)) is None


# ---- the C restatement of the G2 Pippenger (oracle/bn254_ref.c, liboracle.so)

def _jac(pt, z):
    """Affine point -> 192-byte Jacobian bytes with Z = z (z an Fp2 element)."""
    if pt is None:
        one = G.fe(G.to_mont(1))
        return one + bytes(32) + one + bytes(32) + bytes(64)
    (x, y) = pt
    z2 = G.mul2(z, z)
    X, Y = G.mul2(x, z2), G.mul2(y, G.mul2(z2, z))
    return b"".join(G.fe(G.to_mont(v)) for v in (*X, *Y, *z))


def _aff(jac192):
    from oracle import ref
    return G.bytes_to_point(ref.g2_jacobian_to_affine_bytes(jac192))


def test_c_g2_point_ops_match_the_python_group_law():
    from oracle import ref
    rnd = random.Random(11)
    for _ in range(6):
        k1, k2 = rnd.randrange(1, G.R_ORDER), rnd.randrange(1, G.R_ORDER)
        p, q = G.mul(G.G2, k1), G.mul(G.G2, k2)
        z1 = (rnd.randrange(1, G.P), rnd.randrange(G.P))
        z2 = (rnd.randrange(1, G.P), rnd.randrange(G.P))
        jp, jq = _jac(p, z1), _jac(q, z2)
        # Jacobian -> affine undoes the random Z
        assert _aff(jp) == p
        assert _aff(ref.g2_point_op("oracle_g2_jac_add", jp, jq)) == G.add(p, q)
        assert _aff(ref.g2_point_op("oracle_g2_jac_dbl", jp)) == G.add(p, p)
        assert _aff(ref.g2_point_op("oracle_g2_jac_add_affine", jp, G.point_to_bytes(q))) == G.add(p, q)
        # the exception cases: P + P through the adders, P + (-P), identities
        jp2 = _jac(p, z2)
        assert _aff(ref.g2_point_op("oracle_g2_jac_add", jp, jp2)) == G.add(p, p)
        assert _aff(ref.g2_point_op("oracle_g2_jac_add_affine", jp, G.point_to_bytes(p))) == G.add(p, p)
        assert _aff(ref.g2_point_op("oracle_g2_jac_add", jp, _jac(G.neg(p), z2))) is None
        assert _aff(ref.g2_point_op("oracle_g2_jac_add_affine", jp, G.point_to_bytes(G.neg(p)))) is None
        ident = _jac(None, None)
        assert _aff(ref.g2_point_op("oracle_g2_jac_add", ident, jq)) == q
        assert _aff(ref.g2_point_op("oracle_g2_jac_add_affine", ident, G.point_to_bytes(q))) == q
        assert _aff(ref.g2_point_op("oracle_g2_jac_add_affine", jp, bytes(128))) == p
        assert _aff(ref.g2_point_op("oracle_g2_jac_dbl", ident)) is None


def _bases(rnd, n):
    ks = [rnd.randrange(1, G.R_ORDER) for _ in range(n)]
    return ks, b"".join(G.point_to_bytes(G.mul(G.G2, k)) for k in ks)


def _scalars_bytes(ss):
    return b"".join(int(s).to_bytes(32, "little") for s in ss)


def test_c_g2_pippenger_matches_the_closed_form():
    from oracle import ref
    rnd = random.Random(12)
    n = 48
    ks, pts = _bases(rnd, n)
    ss = [rnd.randrange(2**256) for _ in range(n)]
    ss[0], ss[1], ss[2], ss[3] = 0, G.R_ORDER - 1, G.R_ORDER, 2**256 - 1
    want = G.msm_closed_form(ks, ss)
    for c in (0, 3, 5, 11):
        got = ref.g2_pippenger(pts, _scalars_bytes(ss), n, window_bits=c, threads=4)
        assert _aff(got) == want, c
    # the same point many times (every bucket add meets a doubling) and a
    # base with its negation (bucket sums cancel to the identity)
    p = G.point_to_bytes(G.mul(G.G2, ks[0]))
    m = G.point_to_bytes(G.neg(G.mul(G.G2, ks[0])))
    got = ref.g2_pippenger(p * 8, _scalars_bytes([5] * 8), 8, window_bits=4)
    assert _aff(got) == G.mul(G.G2, 40 * ks[0])
    got = ref.g2_pippenger(p + m, _scalars_bytes([77, 77]), 2, window_bits=4)
    assert _aff(got) is None
    # identity bases and an empty input
    got = ref.g2_pippenger(bytes(128) * 3, _scalars_bytes([1, 2, 3]), 3)
    assert _aff(got) is None
    assert _aff(ref.g2_pippenger(b"", b"", 0)) is None


def test_c_g2_pippenger_equals_brute_force_on_a_tiny_case():
    from oracle import ref
    rnd = random.Random(13)
    ks, pts = _bases(rnd, 3)
    ss = [rnd.randrange(2**256) for _ in range(3)]
    brute = None
    for k, s in zip(ks, ss):
        brute = G.add(brute, G.mul(G.mul(G.G2, k), s))
    assert _aff(ref.g2_pippenger(pts, _scalars_bytes(ss), 3, window_bits=7)) == brute
    # a planted mistake (a scalar bit dropped) is seen
    ss_bad = [ss[0] ^ (1 << 200)] + ss[1:]
    assert _aff(ref.g2_pippenger(pts, _scalars_bytes(ss_bad), 3, window_bits=7)) != brute
