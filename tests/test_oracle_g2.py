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
