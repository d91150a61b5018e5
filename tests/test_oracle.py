"""CPU tests that pin the oracle before anything is checked against it.

oracle/bn254.py (pure Python) is pinned by the reference's EIP-196 vectors,
constants recomputed from p, and algebraic facts; oracle/bn254_ref.c (the C
restatement of the reference CPU Pippenger) is pinned against the Python
side, the golden fixtures, brute force and the closed form.  Each check is
also shown to reject a planted mistake.
"""
import json
import os
import random

import numpy as np
import pytest

from helpers import GOLDEN, closed_form, expected_affine, golden_msm_fixtures
from oracle import bn254 as B
from oracle import ref


def test_constants_recomputed_from_p():
    # bn254_msm.c:24-32, BN254Fp.swift:11-29
    assert B.MONT_R == 0x0E0A77C19A07DF2F666EA36F7879462C0A78EB28F5C70B3DD35D438DC58F0D9D
    assert B.MONT_R2 == 0x06D89F71CAB8351F47AB1EFF0A417FF6B5E71911D44501FBF32CFC5B538AFA89
    assert B.MONT_INV64 == 0x87D20782E4866389
    assert (B.P * B.MONT_INV64) % 2**64 == 2**64 - 1


def test_eip196_vectors():
    with open(os.path.join(GOLDEN, "eip196.json")) as fh:
        v = json.load(fh)
    g = (int(v["G"][0], 16), int(v["G"][1], 16))
    assert B.on_curve(g)
    assert B.add(g, g) == (int(v["twoG"][0], 16), int(v["twoG"][1], 16))
    assert B.mul(g, 2) == B.add(g, g)
    assert B.neg(g) == (1, int(v["negG_y"], 16))
    assert B.mul(g, 0) is None
    assert B.mul(g, B.R_ORDER) is None
    assert B.mul(g, B.R_ORDER + 5) == B.mul(g, 5)
    # a planted mistake (wrong y) is caught
    assert not B.on_curve((g[0], g[1] + 1))


def test_endomorphism_constants():
    assert pow(B.BETA, 3, B.P) == 1 and B.BETA != 1
    assert (B.BETA * B.BETA + B.BETA + 1) % B.P == 0
    assert (B.LAMBDA * B.LAMBDA + B.LAMBDA + 1) % B.R_ORDER == 0
    # phi(P) == lambda P on a point other than G
    P = B.mul(B.G, 123456789)
    assert (B.BETA * P[0] % B.P, P[1]) == B.mul(P, B.LAMBDA)
    other = (B.R_ORDER - 1 - B.LAMBDA) % B.R_ORDER  # the other cube root of unity
    assert (B.BETA * P[0] % B.P, P[1]) != B.mul(P, other)


def _glv_cases():
    rnd = random.Random(7)
    cases = [0, 1, 2, 3, B.R_ORDER - 1, B.R_ORDER - 2, B.R_ORDER, B.R_ORDER + 1, 2**127, 2**128,
             B.R_ORDER // 2, B.R_ORDER // 2 + 1, B.LAMBDA, 2**256 - 1]
    cases += [rnd.randrange(2**256) for _ in range(300)]
    cases += [rnd.randrange(B.R_ORDER) for _ in range(300)]
    return cases


def test_glv_python_invariants_and_bounds():
    for k in _glv_cases():
        k1, k2, n1, n2 = B.glv_decompose(k)
        s1 = -k1 if n1 else k1
        s2 = -k2 if n2 else k2
        assert (s1 + B.LAMBDA * s2 - k) % B.R_ORDER == 0
        assert k1 < 2**127 and k2 < 2**127
    # planted mistake: dropping the k2 sign breaks the identity for some k
    bad = 0
    for k in _glv_cases()[:50]:
        k1, k2, n1, n2 = B.glv_decompose(k)
        s1 = -k1 if n1 else k1
        if n2 and (s1 + B.LAMBDA * k2 - k) % B.R_ORDER != 0:
            bad += 1
    assert bad > 0


def test_glv_c_matches_python():
    for k in _glv_cases():
        assert ref.glv_decompose(k) == B.glv_decompose(k)


def test_signed_digits():
    rnd = random.Random(3)
    for c in (2, 5, 8, 11, 13, 16):
        nwin = (128 + c - 1) // c
        for _ in range(200):
            k = rnd.randrange(2**127)
            d, carry = B.signed_digits(k, c, nwin)
            assert carry == 0
            assert sum(x << (c * w) for w, x in enumerate(d)) == k
            assert all(-(1 << (c - 1)) < x <= (1 << (c - 1)) for x in d)
            assert ref.signed_digits(k, c, nwin) == (d, carry)
    # a value at 2^(c*nwin-1) would need the dropped carry
    d, carry = B.signed_digits((1 << 127) | ((1 << 127) - 1), 16, 8)
    assert carry == 1


def test_c_field_ops_match_python():
    rnd = random.Random(11)
    for _ in range(300):
        a, b = rnd.randrange(B.P), rnd.randrange(B.P)
        am, bm = B.to_mont(a), B.to_mont(b)
        assert ref.fp_op("oracle_fp_mul", am, bm) == B.to_mont(a * b % B.P)
        assert ref.fp_op("oracle_fp_add", am, bm) == B.to_mont((a + b) % B.P)
        assert ref.fp_op("oracle_fp_sub", am, bm) == B.to_mont((a - b) % B.P)
        if a:
            assert ref.fp_op("oracle_fp_inv", am) == B.to_mont(pow(a, -1, B.P))
        assert ref.fp_op("oracle_to_mont", a) == am
        assert ref.fp_op("oracle_from_mont", am) == a


def test_c_squaring_and_multiply_edges():
    """The C squaring (its own product schedule) and the unrolled CIOS product
    against Python integers, on random values and on the values that stress
    the carries: 0, 1, p-1, p-2, all-ones words below p, powers of two."""
    rnd = random.Random(12)
    edges = [0, 1, 2, B.P - 1, B.P - 2, (B.P - 1) // 2, (1 << 253), (1 << 254) - 1 - ((1 << 254) - B.P),
             int("ffffffffffffffff", 16), int("ffffffffffffffff" * 3, 16)]
    vals = [v % B.P for v in edges] + [rnd.randrange(B.P) for _ in range(300)]
    for a in vals:
        am = B.to_mont(a)
        assert ref.fp_op("oracle_fp_sqr", am) == B.to_mont(a * a % B.P)
        b = rnd.choice(vals)
        assert ref.fp_op("oracle_fp_mul", am, B.to_mont(b)) == B.to_mont(a * b % B.P)


def _c_msm_affine(fx, window_bits=0, threads=0):
    pts = bytes.fromhex(fx["points"])
    sc = bytes.fromhex(fx["scalars"])
    jac = ref.pippenger(pts, sc, fx["n"], window_bits, threads)
    aff = ref.jacobian_to_affine_bytes(jac)
    x = int.from_bytes(aff[:32], "little")
    y = int.from_bytes(aff[32:], "little")
    if x == 0 and y == 0:
        return None
    return (B.from_mont(x), B.from_mont(y))


@pytest.mark.parametrize("fx", golden_msm_fixtures(), ids=lambda f: f["name"])
def test_c_oracle_matches_golden(fx):
    assert _c_msm_affine(fx) == expected_affine(fx)


def test_c_oracle_window_widths_agree():
    fx = [f for f in golden_msm_fixtures() if f["name"] == "uniform_n64"][0]
    want = expected_affine(fx)
    for c in (3, 5, 8, 11, 13, 16):
        assert _c_msm_affine(fx, window_bits=c, threads=3) == want


def test_closed_form_rejects_planted_mistake():
    ks = [B.uniform_fr(1, i) for i in range(8)]
    ss = [B.uniform_fr(2, i) for i in range(8)]
    pts = [B.mul(B.G, k) for k in ks]
    assert closed_form(ks, ss) == B.msm_bruteforce(pts, ss)
    pts_bad = list(pts)
    pts_bad[3] = B.add(pts_bad[3], B.G)
    assert closed_form(ks, ss) != B.msm_bruteforce(pts_bad, ss)


def test_c_oracle_rejects_planted_mistake():
    fx = [f for f in golden_msm_fixtures() if f["name"] == "uniform_n16"][0]
    pts = bytearray.fromhex(fx["points"])
    pts[64 * 5 + 40] ^= 1  # corrupt one base's y coordinate
    jac = ref.pippenger(bytes(pts), bytes.fromhex(fx["scalars"]), fx["n"])
    aff = ref.jacobian_to_affine_bytes(jac)
    got = (B.from_mont(int.from_bytes(aff[:32], "little")), B.from_mont(int.from_bytes(aff[32:], "little")))
    assert got != expected_affine(fx)


def test_c_oracle_scalar_mul_and_dot():
    g = B.affine_to_bytes(B.G)
    for k in (1, 2, 5, B.R_ORDER - 1, 2**200 + 12345):
        jac = ref.scalar_mul_affine(g, k)
        aff = ref.jacobian_to_affine_bytes(jac)
        got = (B.from_mont(int.from_bytes(aff[:32], "little")), B.from_mont(int.from_bytes(aff[32:], "little")))
        assert got == B.mul(B.G, k)
    ss = [B.gen_scalar(9, i, B.DIST_FULL256) for i in range(40)]
    ks = [B.uniform_fr(10, i) for i in range(40)]
    sb = b"".join(B.scalar_to_bytes(s) for s in ss)
    kb = b"".join(B.scalar_to_bytes(k) for k in ks)
    assert ref.scalar_dot(sb, kb, 40) == sum(s * k for s, k in zip(ss, ks)) % B.R_ORDER
    # the exact 576-bit sum's carries: 1000 products of 2^256 - 1 with itself,
    # and the empty sum
    top = B.scalar_to_bytes(2**256 - 1) * 1000
    assert ref.scalar_dot(top, top, 1000) == 1000 * (2**256 - 1) ** 2 % B.R_ORDER
    assert ref.scalar_dot(b"", b"", 0) == 0
