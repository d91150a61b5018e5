"""CPU model of the BN254 G2 kernel arithmetic (zkmetal_amd/csrc/curve_g2.hpp):
Fp2 = Fp[u]/(u^2+1) on the 29-bit Fp model of test_fq29_model, operation for
operation, with the same conditional subtractions.

What it pins: every Fp2 value the kernels hold between operations has both
components below 2p (the invariant the identity tests rely on), no column or
limb step overflows (the Fp model asserts those), over random and extreme
inputs; and the point operations against the oracle's affine group law on
the twist (oracle/bn254_g2.py)."""
import random

from oracle import bn254_g2 as G2
from test_fq29_model import (P, R, fq_add, fq_csub, fq_mul, fq_mul_add, fq_mul_sub, fq_sqr, fq_sub,
                             iz2, limbs, lt, sub_lazy, val, worst)


def red4(a):
    lt(a, 4)
    return fq_csub(2, a)


def red8(a):
    lt(a, 8)
    return fq_csub(2, fq_csub(4, a))


def chk(a):
    lt(a[0], 2)
    lt(a[1], 2)
    return a


def mul2(a, b):
    chk(a), chk(b)
    t0 = fq_mul(a[0], b[0])
    t1 = fq_mul(a[1], b[1])
    t2 = fq_mul(fq_add(a[0], a[1]), fq_add(b[0], b[1]))
    return chk((red4(fq_sub(2, t0, t1)), red8(fq_sub(3, t2, fq_add(t0, t1)))))


def sqr2(a):
    chk(a)
    c0 = fq_mul(fq_add(a[0], a[1]), fq_sub(2, a[0], a[1]))
    t = fq_mul(a[0], a[1])
    return chk((c0, red4(fq_add(t, t))))


def sub2(a, b):
    chk(a), chk(b)
    return chk((red4(fq_sub(2, a[0], b[0])), red4(fq_sub(2, a[1], b[1]))))


def dbl2(a):
    return chk((red4(fq_add(a[0], a[0])), red4(fq_add(a[1], a[1]))))


def triple2(a):
    return chk((red8(fq_add(fq_add(a[0], a[0]), a[0])), red8(fq_add(fq_add(a[1], a[1]), a[1]))))


def iz(a):
    return iz2(a[0]) and iz2(a[1])


ONE2 = (limbs(R % P), limbs(0))
ZERO2 = (limbs(0), limbs(0))


def dbl_core(X, Y):
    U = dbl2(Y)
    V = sqr2(U)
    W = mul2(U, V)
    S = mul2(X, V)
    M = triple2(sqr2(X))
    X3 = sub2(sub2(sqr2(M), S), S)
    Y3 = sub2(mul2(M, sub2(S, X3)), mul2(W, Y))
    return X3, Y3, V, W


def dbl(p):
    X3, Y3, V, W = dbl_core(p[0], p[1])
    return (X3, Y3, mul2(V, p[2]), mul2(W, p[3]))


def madd_run(acc, acc_id, x2, y2):
    X, Y, ZZ, ZZZ = acc
    Pp = sub2(mul2(x2, ZZ), X)
    Rr = sub2(mul2(y2, ZZZ), Y)
    PP = sqr2(Pp)
    PPP = mul2(Pp, PP)
    Q = mul2(X, PP)
    rX = sub2(sub2(sub2(sqr2(Rr), PPP), Q), Q)
    rY = sub2(mul2(Rr, sub2(Q, rX)), mul2(Y, PPP))
    r = (rX, rY, mul2(ZZ, PP), mul2(ZZZ, PPP))
    cancel = False
    if not acc_id and iz(Pp):
        d = iz(Rr)
        cancel = not d
        if d:
            r = dbl_core(x2, y2)
    q = (x2, y2, ONE2, ONE2)
    return (q if acc_id else r), cancel


# ---- the bucket loop's addition (G2Arith::madd_run): schoolbook Fp2 products
# with merged reductions, c0 = (a0 b0 + a1 (2p - b1))/R, c1 = (a0 b1 + a1 b0)/R,
# subtractions folded into the high columns, one or two conditional
# subtractions on the stored outputs only.

def nl(k, v):
    """k*p - v limb-wise (the borrow form), v normalised < k p"""
    return sub_lazy(k, limbs(0), v)


def lz(a, b):
    r = [a[i] + b[i] for i in range(9)]
    assert all(x < (1 << 30) for x in r)
    return r


def madd_run_lazy(acc, acc_id, x2, y2):
    X, Y, ZZ, ZZZ = acc
    for c in acc:
        chk(c)
    Pp = (fq_mul_add(x2[0], ZZ[0], x2[1], nl(2, ZZ[1]), 2, X[0]), fq_mul_add(x2[0], ZZ[1], x2[1], ZZ[0], 2, X[1]))
    Rr = (fq_mul_add(y2[0], ZZZ[0], y2[1], nl(2, ZZZ[1]), 2, Y[0]), fq_mul_add(y2[0], ZZZ[1], y2[1], ZZZ[0], 2, Y[1]))
    for v in Pp + Rr:
        lt(v, 3.1)
    # fq_mul_sub<0>: one operand normalised, the other's limbs < 1.5 * 2^30
    PP = (fq_mul_sub(0, fq_add(Pp[0], Pp[1]), sub_lazy(4, Pp[0], Pp[1]), Pp[0]),
          fq_mul_sub(0, lz(Pp[0], Pp[0]), Pp[1], Pp[0]))
    lt(PP[0], 1.3), lt(PP[1], 1.3)
    PPP = (fq_mul_add(Pp[0], PP[0], Pp[1], nl(2, PP[1])), fq_mul_add(Pp[0], PP[1], Pp[1], PP[0]))
    Q = (fq_mul_add(X[0], PP[0], X[1], nl(2, PP[1])), fq_mul_add(X[0], PP[1], X[1], PP[0]))
    for v in PPP + Q:
        lt(v, 1.1)
    T = tuple([2 * Q[j][i] + PPP[j][i] for i in range(9)] for j in range(2))
    lt(T[0], 4), lt(T[1], 4)
    X3 = (fq_csub(2, fq_csub(4, fq_mul_sub(4, fq_add(Rr[0], Rr[1]), sub_lazy(4, Rr[0], Rr[1]), T[0], borrow=3))),
          fq_csub(2, fq_csub(4, fq_mul_sub(4, lz(Rr[0], Rr[0]), Rr[1], T[1], borrow=3))))
    chk(X3)
    M2 = (fq_mul_add(Y[0], PPP[0], Y[1], nl(2, PPP[1])), fq_mul_add(Y[0], PPP[1], Y[1], PPP[0]))
    TT = (fq_sub(2, Q[0], X3[0]), fq_sub(2, Q[1], X3[1]))
    lt(TT[0], 3.1), lt(TT[1], 3.1)
    Y3 = (fq_csub(2, fq_mul_add(Rr[0], TT[0], Rr[1], nl(4, TT[1]), 2, M2[0])),
          fq_csub(2, fq_mul_add(Rr[0], TT[1], Rr[1], TT[0], 2, M2[1])))
    chk(Y3)
    ZZ3 = (fq_mul_add(ZZ[0], PP[0], ZZ[1], nl(2, PP[1])), fq_mul_add(ZZ[0], PP[1], ZZ[1], PP[0]))
    ZZZ3 = (fq_mul_add(ZZZ[0], PPP[0], ZZZ[1], nl(2, PPP[1])), fq_mul_add(ZZZ[0], PPP[1], ZZZ[1], PPP[0]))
    r = (X3, Y3, chk(ZZ3), chk(ZZZ3))
    cancel = False
    if not acc_id and iz(PP):
        d = iz((fq_csub(2, Rr[0]), fq_csub(2, Rr[1])))
        cancel = not d
        if d:
            X3d, Y3d, V, W = dbl_core(x2, y2)
            r = (X3d, Y3d, V, W)
    if acc_id:
        r = (x2, y2, ONE2, ONE2)
    return r, cancel


def m2(a, b, k=0, e=None, kb=2):
    """mul2m<K, KB>: schoolbook Fp2 product, merged reductions, -e + k p"""
    return (fq_mul_add(a[0], b[0], a[1], nl(kb, b[1]), k, e and e[0]),
            fq_mul_add(a[0], b[1], a[1], b[0], k, e and e[1]))


def add_lazy(acc, q):
    """G2Arith::add on merged reductions: both XYZZ, components < 2p"""
    for c in acc + q:
        chk(c)
    p_id, q_id = iz(acc[2]), iz(q[2])
    U1 = m2(acc[0], q[2])
    S1 = m2(acc[1], q[3])
    Pp = m2(q[0], acc[2], 2, U1)
    Rr = m2(q[1], acc[3], 2, S1)
    for v in Pp + Rr:
        lt(v, 3.1)
    PP = (fq_mul_sub(0, fq_add(Pp[0], Pp[1]), sub_lazy(4, Pp[0], Pp[1]), Pp[0]),
          fq_mul_sub(0, lz(Pp[0], Pp[0]), Pp[1], Pp[0]))
    PPP = m2(Pp, PP)
    Q = m2(U1, PP)
    T = tuple([2 * Q[j][i] + PPP[j][i] for i in range(9)] for j in range(2))
    lt(T[0], 4), lt(T[1], 4)
    X3 = (fq_csub(2, fq_csub(4, fq_mul_sub(4, fq_add(Rr[0], Rr[1]), sub_lazy(4, Rr[0], Rr[1]), T[0], borrow=3))),
          fq_csub(2, fq_csub(4, fq_mul_sub(4, lz(Rr[0], Rr[0]), Rr[1], T[1], borrow=3))))
    M2 = m2(S1, PPP)
    TT = (fq_sub(2, Q[0], X3[0]), fq_sub(2, Q[1], X3[1]))
    Y3 = m2(Rr, TT, 2, M2, kb=4)
    Y3 = (fq_csub(2, Y3[0]), fq_csub(2, Y3[1]))
    r = (chk(X3), chk(Y3), chk(m2(m2(acc[2], q[2]), PP)), chk(m2(m2(acc[3], q[3]), PPP)))
    d = iz(PP) and iz((fq_csub(2, Rr[0]), fq_csub(2, Rr[1]))) and not p_id and not q_id
    if d:
        r = dbl(acc)
    if p_id:
        r = q
    return acc if q_id else r


def add(acc, q):
    p_id, q_id = iz(acc[2]), iz(q[2])
    U1 = mul2(acc[0], q[2])
    U2 = mul2(q[0], acc[2])
    S1 = mul2(acc[1], q[3])
    S2 = mul2(q[1], acc[3])
    Pp = sub2(U2, U1)
    Rr = sub2(S2, S1)
    d = iz(Pp) and iz(Rr) and not p_id and not q_id
    PP = sqr2(Pp)
    PPP = mul2(Pp, PP)
    Q = mul2(U1, PP)
    rX = sub2(sub2(sub2(sqr2(Rr), PPP), Q), Q)
    rY = sub2(mul2(Rr, sub2(Q, rX)), mul2(S1, PPP))
    r = (rX, rY, mul2(mul2(acc[2], q[2]), PP), mul2(mul2(acc[3], q[3]), PPP))
    if d:
        r = dbl(acc)
    if p_id:
        r = q
    return acc if q_id else r


def to2(a):
    return (limbs(a[0] * R % P), limbs(a[1] * R % P))


def from2(a):
    ri = pow(R, -1, P)
    return (val(a[0]) * ri % P, val(a[1]) * ri % P)


def affine(pt):
    X, Y, ZZ, ZZZ = (from2(c) for c in pt)
    if ZZ == (0, 0):
        return None
    return (G2.mul2(X, G2.inv2(ZZ)), G2.mul2(Y, G2.inv2(ZZZ)))


def xyzz_of(pt, rnd):
    """a random XYZZ representative of an affine point, with components
    pushed to the top of their bound (value + p where that stays < 2p)"""
    if pt is None:
        return (ZERO2, ZERO2, ZERO2, ZERO2)
    z = (rnd.randrange(1, P), rnd.randrange(P))
    zz = G2.mul2(z, z)
    zzz = G2.mul2(zz, z)
    coords = (G2.mul2(pt[0], zz), G2.mul2(pt[1], zzz), zz, zzz)

    def lift(c):
        v = c * R % P
        return limbs(v + P if v < P - 1 and rnd.random() < 0.5 else v)

    return tuple((lift(c[0]), lift(c[1])) for c in coords)


def test_fp2_ops_extremes_and_values():
    rnd = random.Random(11)
    edge = [limbs(0), limbs(1), limbs(P - 1), limbs(P), worst(2), limbs(2 * P - 2)]
    vals = [(a, b) for a in edge for b in edge] + [
        (limbs(rnd.randrange(2 * P)), limbs(rnd.randrange(2 * P))) for _ in range(200)]
    ri = pow(R, -1, P)
    for i in range(len(vals)):
        a, b = vals[i], vals[(i * 7 + 3) % len(vals)]
        A = (val(a[0]) * ri % P, val(a[1]) * ri % P)
        Bv = (val(b[0]) * ri % P, val(b[1]) * ri % P)
        assert from2(mul2(a, b)) == G2.mul2(A, Bv)
        assert from2(sqr2(a)) == G2.mul2(A, A)
        assert from2(sub2(a, b)) == G2.sub2(A, Bv)
        assert from2(dbl2(a)) == G2.add2(A, A)
        assert from2(triple2(a)) == G2.add2(G2.add2(A, A), A)


def test_g2_point_ops_against_the_oracle():
    rnd = random.Random(12)
    pts = [G2.mul(G2.G2, rnd.randrange(1, G2.R_ORDER)) for _ in range(10)]
    P0 = pts[0]
    cases = [(pts[i], pts[i + 1]) for i in range(9)]
    cases += [(P0, P0), (P0, G2.neg(P0)), (P0, None), (None, P0), (None, None)]
    for a, b in cases:
        A, Bq = xyzz_of(a, rnd), xyzz_of(b, rnd)
        assert affine(add(A, Bq)) == G2.add(a, b)
        assert affine(add_lazy(A, Bq)) == G2.add(a, b)
        assert affine(dbl(A)) == G2.add(a, a)
    # the run addition: a run over a few points with a repeat and a cancellation
    # (both forms; the table holds y or p - y, canonical or p itself)
    run = [pts[1], pts[2], pts[2], G2.neg(G2.add(pts[1], G2.add(pts[2], pts[2]))), pts[3], pts[3]]
    for step in (madd_run, madd_run_lazy):
        acc, acc_id, want = None, True, None
        acc = (ZERO2, ZERO2, ONE2, ONE2)
        for q in run:
            qa = to2(q[0]), to2(q[1])
            acc, acc_id = step(acc, acc_id, qa[0], qa[1])
            want = G2.add(want, q)
            got = None if acc_id else affine(acc)
            assert got == want


def test_g2_run_addition_at_extremes():
    """the bucket loop's addition from accumulators at the top of their
    bound (every component just below 2p where its residue allows), long
    runs, and table operands equal to p"""
    rnd = random.Random(13)
    for _ in range(12):
        a = G2.mul(G2.G2, rnd.randrange(1, G2.R_ORDER))
        q = G2.mul(G2.G2, rnd.randrange(1, G2.R_ORDER))
        A = xyzz_of(a, rnd)
        qa = to2(q[0]), to2(q[1])
        out, f = madd_run_lazy(A, False, qa[0], qa[1])
        assert not f and affine(out) == G2.add(a, q)
        cur, want = out, G2.add(a, q)
        for _ in range(6):
            q2 = G2.mul(G2.G2, rnd.randrange(1, G2.R_ORDER))
            cur, f = madd_run_lazy(cur, False, to2(q2[0]), to2(q2[1]))
            want = G2.add(want, q2)
            assert not f and affine(cur) == want
        # the general addition chained on its own outputs
        s, sw = cur, want
        for _ in range(4):
            s = add_lazy(s, A)
            sw = G2.add(sw, a)
            assert affine(s) == sw
        assert affine(add_lazy(s, s)) == G2.add(sw, sw)
    # a y component of 0 stored as p - 0 = p
    y = (limbs(P), limbs(5))
    assert val(sub2(y, (limbs(0), limbs(5)))[0]) % P == 0
    # the table's p - y form of y = 0 components is p itself (< 2p)
    ny = (limbs(P), limbs(P - 5))
    assert iz(sub2(ny, (limbs(0), limbs(P - 5))))
