"""GPU unit tests of the device field, XYZZ point and GLV code (through the
C ABI probes zkmsm_test_fp / zkmsm_test_xyzz / zkmsm_test_glv) against the
Python big-integer oracle."""
import random

import numpy as np
import pytest

from oracle import bn254 as B

pytestmark = pytest.mark.gpu


def fe(x):
    return int(x).to_bytes(32, "little")


def rd(buf, i):
    return int.from_bytes(bytes(buf[32 * i:32 * i + 32]), "little")


EDGE = [0, 1, 2, B.P - 1, B.P - 2, B.MONT_R % B.P, (1 << 255) % B.P, B.P >> 1]


def _inputs(n, seed):
    rnd = random.Random(seed)
    vals = EDGE + [rnd.randrange(B.P) for _ in range(n - len(EDGE))]
    return vals, [rnd.randrange(B.P) for _ in range(n)]


def test_fp_mul_add_sub(engine):
    a, b = _inputs(2048, 1)
    A = b"".join(fe(B.to_mont(x)) for x in a)
    Bb = b"".join(fe(B.to_mont(x)) for x in b)
    mul = engine.test_fp(0, A, Bb)
    add = engine.test_fp(1, A, Bb)
    sub = engine.test_fp(2, A, Bb)
    for i, (x, y) in enumerate(zip(a, b)):
        assert rd(mul, i) == B.to_mont(x * y % B.P), i
        assert rd(add, i) == B.to_mont((x + y) % B.P), i
        assert rd(sub, i) == B.to_mont((x - y) % B.P), i


def test_fq29_mul_matches_montgomery(engine):
    """9 x 29-bit limb multiply (R = 2^261), converted in and out of the
    8 x 32-bit form: must equal the 2^256-Montgomery product exactly."""
    a, b = _inputs(2048, 11)
    A = b"".join(fe(B.to_mont(x)) for x in a)
    Bb = b"".join(fe(B.to_mont(x)) for x in b)
    out = engine.test_fp(6, A, Bb)
    sq = engine.test_fp(7, A, A)
    for i, (x, y) in enumerate(zip(a, b)):
        assert rd(out, i) == B.to_mont(x * y % B.P), i
        assert rd(sq, i) == B.to_mont(x * x % B.P), i


def test_fp_inv_neg_canon(engine):
    a, _ = _inputs(512, 2)
    a = [x for x in a if x]
    A = b"".join(fe(B.to_mont(x)) for x in a)
    inv = engine.test_fp(3, A, A)
    neg = engine.test_fp(4, A, A)
    for i, x in enumerate(a):
        assert rd(inv, i) == B.to_mont(pow(x, -1, B.P))
        assert rd(neg, i) == B.to_mont((-x) % B.P)
    z = engine.test_fp(4, fe(0), fe(0))
    assert rd(z, 0) == 0
    # non-canonical encodings x + k p < 2^256 reduce to x
    raw = [x + k * B.P for x in (0, 1, B.P - 1, 12345) for k in range(6) if x + k * B.P < 2**256]
    C = engine.test_fp(5, b"".join(fe(v) for v in raw), b"".join(fe(v) for v in raw))
    for i, v in enumerate(raw):
        assert rd(C, i) == v % B.P


def _xyzz_bytes(pt, z):
    """XYZZ (Montgomery) of affine pt with projective factor z; None -> identity."""
    if pt is None:
        return bytes(128)
    x, y = pt
    zz, zzz = z * z % B.P, z * z * z % B.P
    return b"".join(fe(B.to_mont(v)) for v in (x * zz % B.P, y * zzz % B.P, zz, zzz))


def _xyzz_affine(buf, i):
    X, Y, ZZ, ZZZ = (B.from_mont(rd(buf, 4 * i + k)) for k in range(4))
    if ZZ == 0:
        return None
    return (X * pow(ZZ, -1, B.P) % B.P, Y * pow(ZZZ, -1, B.P) % B.P)


@pytest.mark.parametrize("base", [0, 3], ids=["limb29", "limb32"])
def test_xyzz_add_dbl_madd(engine, base):
    """base 0: the 29-bit-limb arithmetic of the MSM kernels; base 3: the
    32-bit-limb arithmetic of the base generator."""
    rnd = random.Random(5)
    pts = [B.mul(B.G, rnd.randrange(1, B.R_ORDER)) for _ in range(24)]
    cases = []
    for i in range(len(pts) - 1):
        cases.append((pts[i], pts[i + 1]))
    P = pts[0]
    cases += [(P, P), (P, B.neg(P)), (P, None), (None, P), (None, None), (B.G, B.G)]
    Pb = b"".join(_xyzz_bytes(p, rnd.randrange(1, B.P)) for p, _ in cases)
    Qb = b"".join(_xyzz_bytes(q, rnd.randrange(1, B.P)) for _, q in cases)
    add = engine.test_xyzz(base + 0, Pb, Qb)
    dbl = engine.test_xyzz(base + 1, Pb, Qb)
    # madd uses q's X, Y as an affine point: pass Z = 1 and (0,0) for None
    Qa = b"".join(_xyzz_bytes(q, 1) if q is not None else bytes(128) for _, q in cases)
    madd = engine.test_xyzz(base + 2, Pb, Qa)
    for i, (p, q) in enumerate(cases):
        assert _xyzz_affine(add, i) == B.add(p, q), i
        assert _xyzz_affine(dbl, i) == B.add(p, p), i
        assert _xyzz_affine(madd, i) == B.add(p, q), i


def test_glv_split_matches_oracle(engine):
    rnd = random.Random(9)
    ks = [0, 1, B.R_ORDER - 1, B.R_ORDER, B.R_ORDER + 7, 2**256 - 1, 2**127, B.LAMBDA]
    ks += [rnd.randrange(2**256) for _ in range(1000)]
    out = engine.test_glv(b"".join(fe(k) for k in ks))
    for i, k in enumerate(ks):
        rec = bytes(out[48 * i:48 * i + 48])
        k1 = int.from_bytes(rec[:16], "little")
        k2 = int.from_bytes(rec[16:32], "little")
        n1 = int.from_bytes(rec[32:36], "little")
        n2 = int.from_bytes(rec[36:40], "little")
        assert (k1, k2, bool(n1), bool(n2)) == B.glv_decompose(k), hex(k)
