"""CPU pinning of the Grumpkin oracle (oracle/grumpkin.py) against the
reference's constants and the group structure (no GPU)."""
import random

from oracle import bn254 as B
from oracle import grumpkin as G


def test_field_constants_match_the_reference():
    # Sources/NeonFieldOps/grumpkin_ops.c:17-25
    assert B.int_from_limbs(G.REF_P_LIMBS64, bits=64) == G.R
    assert G.REF_INV64 == (-pow(G.R, -1, 2**64)) % 2**64
    assert B.int_from_limbs(G.REF_ONE_LIMBS64, bits=64) == pow(2, 256, G.R)


def test_generator_and_order():
    # GrumpkinCurve.swift:215-222: G = (1, sqrt(-16)), the quoted y
    assert G.on_curve(G.G)
    assert (G.GEN_Y * G.GEN_Y + 16) % G.R == 0
    assert G.mul(G.G, G.ORDER) is None  # order p (prime), G != O
    assert G.add(G.G, G.G) is not None and G.add(G.G, G.neg(G.G)) is None
    # a planted mistake in b is caught by the on-curve check
    x, y = G.G
    assert (y * y - x * x * x - (G.B + 1)) % G.R != 0


def test_group_law_consistency():
    rnd = random.Random(3)
    pts = [G.mul(G.G, rnd.randrange(1, G.ORDER)) for _ in range(6)]
    for p in pts:
        assert G.on_curve(p)
    a, b, c = pts[:3]
    assert G.add(G.add(a, b), c) == G.add(a, G.add(b, c))
    assert G.add(a, b) == G.add(b, a)
    k1, k2 = rnd.randrange(G.ORDER), rnd.randrange(G.ORDER)
    assert G.add(G.mul(G.G, k1), G.mul(G.G, k2)) == G.mul(G.G, k1 + k2)
    # 5 G by repeated addition
    acc = None
    for _ in range(5):
        acc = G.add(acc, G.G)
    assert acc == G.mul(G.G, 5)


def test_closed_form_equals_brute_force():
    rnd = random.Random(4)
    ks = [rnd.randrange(1, G.ORDER) for _ in range(5)]
    ss = [rnd.randrange(0, 2**256) for _ in range(5)]
    brute = None
    for k, s in zip(ks, ss):
        brute = G.add(brute, G.mul(G.mul(G.G, k), s))
    assert brute == G.msm_closed_form(ks, ss)


def test_byte_forms_round_trip():
    p = G.mul(G.G, 12345)
    b = G.point_to_bytes(p)
    x = G.from_mont(int.from_bytes(b[:32], "little"))
    y = G.from_mont(int.from_bytes(b[32:], "little"))
    assert (x, y) == p
    one = G.to_mont(1).to_bytes(32, "little")
    assert G.jacobian_bytes_to_affine(b + one) == p
    assert G.jacobian_bytes_to_affine(one + one + bytes(32)) is None


# ---- the C restatement of grumpkin_pippenger_msm (oracle/liboracle_grumpkin.so)

def _c_affine(r96):
    from oracle import ref
    aff = ref.jacobian_to_affine_bytes(r96, curve="grumpkin")
    x = G.from_mont(int.from_bytes(aff[:32], "little"))
    y = G.from_mont(int.from_bytes(aff[32:], "little"))
    return None if aff == bytes(64) else (x, y)


def _c_msm(pts, ss, c=0):
    from oracle import ref
    pb = b"".join(G.point_to_bytes(p) for p in pts)
    sb = b"".join(int(s).to_bytes(32, "little") for s in ss)
    return _c_affine(ref.pippenger(pb, sb, len(pts), window_bits=c, curve="grumpkin"))


def test_c_field_is_fr():
    from oracle import ref
    rnd = random.Random(5)
    for _ in range(50):
        a, b = rnd.randrange(G.R), rnd.randrange(G.R)
        got = ref.fp_op("oracle_fp_mul", G.to_mont(a), G.to_mont(b), curve="grumpkin")
        assert got == G.to_mont(a * b % G.R)
        assert ref.fp_op("oracle_fp_sqr", G.to_mont(a), curve="grumpkin") == G.to_mont(a * a % G.R)
        assert ref.fp_op("oracle_fp_sub", a, b, curve="grumpkin") == (a - b) % G.R
    # R mod r and R^2 mod r agree with the reference's literal one (grumpkin_ops.c:22-25)
    assert ref.fp_op("oracle_to_mont", 1, curve="grumpkin") == B.int_from_limbs(G.REF_ONE_LIMBS64, bits=64)


def test_c_pippenger_matches_closed_form_and_edges():
    rnd = random.Random(6)
    ks = [rnd.randrange(1, G.ORDER) for _ in range(40)]
    pts = [G.mul(G.G, k) for k in ks]
    ss = [rnd.randrange(2**256) for _ in ks]
    ss[:6] = [0, 1, G.ORDER - 1, G.ORDER, 2**256 - 1, G.R]
    pts[7], ks[7] = None, 0  # (0,0) base = identity, skipped
    want = G.msm_closed_form(ks, ss)
    for c in (0, 3, 5, 8):
        assert _c_msm(pts, ss, c) == want, c
    # repeated bases with scalars 1..4 -> 10 P (correctness_tests.swift:1756-1768)
    P = pts[0]
    assert _c_msm([P] * 4, [1, 2, 3, 4]) == G.mul(P, 10)
    # cancellation: s P + (p - s) P = O
    assert _c_msm([P, P], [9, G.ORDER - 9]) is None
    # a planted mistake (one base replaced by its negation) is caught
    bad = list(pts)
    bad[10] = G.neg(bad[10])
    assert _c_msm(bad, ss) != want


def test_c_scalar_dot_is_mod_group_order():
    from oracle import ref
    rnd = random.Random(7)
    ss = [rnd.randrange(2**256) for _ in range(20)]
    ks = [rnd.randrange(G.ORDER) for _ in range(20)]
    sb = b"".join(s.to_bytes(32, "little") for s in ss)
    kb = b"".join(k.to_bytes(32, "little") for k in ks)
    assert ref.scalar_dot(sb, kb, 20, curve="grumpkin") == sum(s * k for s, k in zip(ss, ks)) % G.ORDER
