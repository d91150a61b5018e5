"""CPU model of the 29-bit-limb field and XYZZ arithmetic used by the MSM
kernels (zkmetal_amd/csrc/field29.hpp, curve29.hpp), limb for limb.

What it pins: the design invariants the GPU cannot report -- no column of the
Montgomery product exceeds 64 bits, no signed limb step leaves int32, and
every value stays below its stated bound (products < 1.5p, stored X, Y < 4p,
ZZ, ZZZ < 2p) -- over random and extreme inputs, and the point results
against the oracle's affine group law.  The GPU parity tests check the
kernels themselves."""
import random

from oracle import bn254 as B

P = B.P
MASK = (1 << 29) - 1
R = 1 << 261
PINV = (-pow(P, -1, 1 << 29)) % (1 << 29)
U64 = 1 << 64


def limbs(x):
    return [(x >> (29 * i)) & MASK for i in range(9)]


def val(a):
    return sum(v << (29 * i) for i, v in enumerate(a))


PL = limbs(P)


def kp(k):
    return limbs(k * P)


def fq_mul(a, b):
    assert all(v < (1 << 30) for v in a + b)
    assert val(a) * val(b) < P * R // 2
    m = [0] * 9
    t = [0] * 9
    acc = 0

    def mac(x, y):
        nonlocal acc
        acc += x * y
        assert acc < U64, "column overflow"

    for k in range(9):
        for j in range(k):
            mac(a[j], b[k - j])
            mac(m[j], PL[k - j])
        mac(a[k], b[0])
        m[k] = ((acc & 0xFFFFFFFF) * PINV) & MASK
        mac(m[k], PL[0])
        assert acc & MASK == 0
        acc >>= 29
    for k in range(9, 17):
        for j in range(k - 8, 9):
            mac(a[j], b[k - j])
            mac(m[j], PL[k - j])
        t[k - 9] = acc & MASK
        acc >>= 29
    t[8] = acc
    assert t[8] < (1 << 32)
    r = val(t)
    assert r % P == val(a) * val(b) * pow(R, -1, P) % P
    assert r < val(a) * val(b) // R + P + 1
    return t


def fq_sqr(a):
    """the squaring variant: doubled cross products once, then the squares"""
    assert all(v < (1 << 29) for v in a)
    assert val(a) * val(a) < P * R // 2
    a2 = [2 * v for v in a]
    m = [0] * 9
    t = [0] * 9
    acc = 0

    def mac(x, y):
        nonlocal acc
        acc += x * y
        assert acc < U64, "column overflow"

    for k in range(17):
        lo = max(0, k - 8)
        for i in range(lo, 9):
            if 2 * i < k and k - i <= 8:
                mac(a2[i], a[k - i])
        if k % 2 == 0 and k // 2 <= 8:
            mac(a[k // 2], a[k // 2])
        for j in range(lo, min(k, 9)):
            mac(m[j], PL[k - j])
        if k < 9:
            m[k] = ((acc & 0xFFFFFFFF) * PINV) & MASK
            mac(m[k], PL[0])
            assert acc & MASK == 0
            acc >>= 29
        else:
            t[k - 9] = acc & MASK
            acc >>= 29
    t[8] = acc
    assert t == fq_mul(a, a)
    return t


def kpr(k, b=1):
    """k*p with b * 2^29 borrowed into limbs 0..7 (q_kpr<K, B>)"""
    c = kp(k)
    r = [c[0] + b * (1 << 29)] + [c[i] + b * (1 << 29) - b for i in range(1, 8)] + [c[8] - b]
    assert val(r) == k * P
    return r


def sub_lazy(k, q, c, b=1):
    """fq_sub_lazy<K, B>: q - c + k*p limb by limb, no carries"""
    K = kpr(k, b)
    r = [q[i] + K[i] - c[i] for i in range(9)]
    assert all(0 <= v < (1 << 32) for v in r)
    assert val(r) == val(q) - val(c) + k * P
    return r


def fq_mul_sub(k, a, b, c, sqr=False, borrow=1):
    """fq_mul_sub<K, B> / fq_sqr_sub<K, B>: a*b/R - c + k*p, the limbs of
    k*p - c entering the high columns (k = 0: the plain product)."""
    if sqr:
        assert a is b and all(v < (1 << 29) for v in a)
    # limbs < 2^30, or one operand normalised and the other < 1.5 * 2^30
    assert (all(v < (1 << 30) for v in a + b)
            or (all(v < (1 << 29) for v in a) and all(v < 3 << 29 for v in b)))
    assert val(a) * val(b) < P * R  # result < 2p
    d = [0] * 9
    if k:
        K = kpr(k, borrow)
        assert all(v <= borrow * MASK for v in c)
        d = [K[i] - c[i] for i in range(9)]
        assert all(0 <= x < (1 << 32) for x in d), "negative limb of k*p - c"
    m = [0] * 9
    t = [0] * 9
    acc = 0

    def mac(x, y):
        nonlocal acc
        acc += x * y
        assert acc < U64, "column overflow"

    for kk in range(9):
        for j in range(kk):
            mac(a[j], b[kk - j])
            mac(m[j], PL[kk - j])
        mac(a[kk], b[0])
        m[kk] = ((acc & 0xFFFFFFFF) * PINV) & MASK
        mac(m[kk], PL[0])
        assert acc & MASK == 0
        acc >>= 29
    for kk in range(9, 17):
        acc += d[kk - 9]
        for j in range(kk - 8, 9):
            mac(a[j], b[kk - j])
            mac(m[j], PL[kk - j])
        t[kk - 9] = acc & MASK
        acc >>= 29
    t[8] = acc + d[8]
    assert t[8] < (1 << 29)
    r = val(t)
    cv = val(c) if k else 0
    prod = val(a) * val(b)
    assert (r + cv - k * P) % P == prod * pow(R, -1, P) % P
    assert r < prod // R + P + 1 - cv + k * P
    return t


def fq_mul_add(a, b, c, d, k=0, e=None, borrow=1):
    """fq_mul_add_sub<K, B>: (a*b + c*d)/R - e + k*p with one Montgomery
    reduction (k = 0: fq_mul_add, no subtraction), the limbs of k*p - e
    entering the high columns (the kernel splits each column over two
    accumulators; their sum is this one).
    a, c normalised; b limbs < 1.5 * 2^30, d limbs < 2^30 (column sums below
    9 * 1.5 * 2^59 + 9 * 2^59 + 9 * 2^58 = 2^63.75)."""
    assert all(v < (1 << 29) for v in a + c)
    assert all(v < (3 << 29) for v in b) and all(v < (1 << 30) for v in d)
    s = val(a) * val(b) + val(c) * val(d)
    assert s < P * R // 2
    dd = [0] * 9
    if k:
        K = kpr(k, borrow)
        assert all(v <= borrow * MASK for v in e)
        dd = [K[i] - e[i] for i in range(9)]
        assert all(0 <= x < (1 << 32) for x in dd), "negative limb of k*p - e"
    m = [0] * 9
    t = [0] * 9
    acc = 0

    def mac(x, y):
        nonlocal acc
        acc += x * y
        assert acc < U64, "column overflow"

    for kk in range(9):
        for j in range(kk):
            mac(a[j], b[kk - j])
            mac(c[j], d[kk - j])
            mac(m[j], PL[kk - j])
        mac(a[kk], b[0])
        mac(c[kk], d[0])
        m[kk] = ((acc & 0xFFFFFFFF) * PINV) & MASK
        mac(m[kk], PL[0])
        assert acc & MASK == 0
        acc >>= 29
    for kk in range(9, 17):
        acc += dd[kk - 9]
        for j in range(kk - 8, 9):
            mac(a[j], b[kk - j])
            mac(c[j], d[kk - j])
            mac(m[j], PL[kk - j])
        t[kk - 9] = acc & MASK
        acc >>= 29
    t[8] = acc + dd[8]
    assert t[8] < (1 << 29)
    r = val(t)
    ev = val(e) if k else 0
    assert (r + ev - k * P) % P == s * pow(R, -1, P) % P
    assert r < s // R + P + 1 - ev + k * P
    return t


def fq_add(a, b):
    r, c = [], 0
    for i in range(9):
        s = a[i] + b[i] + c
        assert s < (1 << 32)
        r.append(s & MASK)
        c = s >> 29
    assert c == 0
    return r


def fq_sub(k, a, b):
    assert val(b) < k * P
    K = kp(k)
    r, c = [], 0
    for i in range(9):
        s = (a[i] + K[i]) - b[i] + c
        assert -(1 << 31) <= s < (1 << 31) and a[i] + K[i] < (1 << 32)
        r.append(s & MASK)
        c = s >> 29
    assert c == 0
    assert val(r) == val(a) - val(b) + k * P
    return r


def fq_csub(k, a):
    K = kp(k)
    d, c = [], 0
    for i in range(9):
        s = a[i] - K[i] + c
        d.append(s & MASK)
        c = s >> 29
    r = a if c < 0 else d
    assert val(r) == (val(a) if val(a) < k * P else val(a) - k * P)
    return r


def iz2(a):
    assert val(a) < 2 * P
    return val(a) in (0, P)


ONE = limbs(R % P)


def lt(a, bound):
    assert val(a) < bound * P, (val(a) / P, bound)


def dbl_core(X, Y):
    U = fq_add(Y, Y)
    lt(U, 8)
    V = fq_sqr(U)
    W = fq_mul(U, V)
    S = fq_mul(X, V)
    xx = fq_sqr(X)
    M = fq_add(fq_add(xx, xx), xx)
    lt(M, 4.5)
    X3 = fq_csub(2, fq_sub(3, fq_sqr(M), fq_add(S, S)))
    lt(X3, 2.5)
    Y3 = fq_sub(2, fq_mul(M, fq_sub(3, S, X3)), fq_mul(W, Y))
    lt(Y3, 3.5)
    for v in (V, W):
        lt(v, 1.5)
    return X3, Y3, V, W


def madd(acc, x2, y2, q_id):
    X1, Y1, ZZ1, ZZZ1 = acc
    lt(X1, 4), lt(Y1, 4), lt(ZZ1, 2), lt(ZZZ1, 2)
    p_id = iz2(ZZ1)
    U2 = fq_mul(x2, ZZ1)
    S2 = fq_mul(y2, ZZZ1)
    Pp = fq_sub(4, U2, X1)
    Rr = fq_sub(4, S2, Y1)
    PP = fq_sqr(Pp)
    PPP = fq_mul(Pp, PP)
    Q = fq_mul(X1, PP)
    RR = fq_sqr(Rr)
    dbl = iz2(PP) and iz2(RR) and not p_id and not q_id
    X3 = fq_csub(3, fq_sub(5, RR, fq_add(fq_add(Q, Q), PPP)))
    lt(X3, 3.5)
    Y3 = fq_sub(2, fq_mul(Rr, fq_sub(4, Q, X3)), fq_mul(Y1, PPP))
    r = (X3, Y3, fq_mul(ZZ1, PP), fq_mul(ZZZ1, PPP))
    if dbl:
        r = dbl_core(x2, y2)
    if p_id:
        r = (x2, y2, ONE, ONE)
    if q_id:
        r = acc
    return r


RUN_START = (limbs(0), limbs(0), ONE, ONE)


def madd_run(acc, acc_id, x2, y2):
    """qxyzz_madd_run: the accumulation-loop addition, identity as a flag,
    exceptional cases screened on limb 0 of PP.  Products in pairs (P, R),
    (PP, RR), (PPP, Q), (ZZ3, ZZZ3) -- pairing is scheduling only -- and
    Y3 = R (Q - X3 + 7p) + Y (2p - PPP) under one reduction."""
    X1, Y1, ZZ1, ZZZ1 = acc
    lt(X1, 6.5), lt(Y1, 4), lt(ZZ1, 2), lt(ZZZ1, 2)
    Pp = fq_mul_sub(7, x2, ZZ1, X1)
    Rr = fq_mul_sub(5, y2, ZZZ1, Y1)
    lt(Pp, 8.5), lt(Rr, 6.5)
    PP = fq_sqr(Pp)
    RR = fq_sqr(Rr)
    lt(PP, 1.6), lt(RR, 1.6)
    PPP = fq_mul(Pp, PP)
    Q = fq_mul(X1, PP)
    T = [2 * Q[i] + PPP[i] for i in range(9)]  # lazy: no carries
    lt(T, 4.5)
    X3 = fq_sub(5, RR, T)
    lt(X3, 6.5)
    Y3 = fq_mul_add(Rr, sub_lazy(7, Q, X3), Y1, sub_lazy(2, limbs(0), PPP))
    lt(Y3, 1.5)
    r = (X3, Y3, fq_mul(ZZ1, PP), fq_mul(ZZZ1, PPP))
    cancel = False
    if PP[0] in (0, PL[0]):  # the one-limb screen
        z = not acc_id and iz2(PP)
        d = z and iz2(RR)
        cancel = z and not d
        if d:
            r = dbl_core(x2, y2)
    else:
        assert not iz2(PP), "screen missed PP = 0 mod p"
    if acc_id:
        r = (x2, y2, ONE, ONE)
    return r, cancel


def add(acc, q):
    X1, Y1, ZZ1, ZZZ1 = acc
    X2, Y2, ZZ2, ZZZ2 = q
    """qxyzz_add: subtractions folded into the products (U2 - U1 + 2p,
    S2 - S1 + 2p), Y3's two products under one reduction"""
    # stored points come straight from the accumulation loop: X < 6.5p
    lt(X1, 6.5), lt(Y1, 4), lt(ZZ1, 2), lt(ZZZ1, 2)
    lt(X2, 6.5), lt(Y2, 4), lt(ZZ2, 2), lt(ZZZ2, 2)
    p_id, q_id = iz2(ZZ1), iz2(ZZ2)
    U1 = fq_mul(X1, ZZ2)
    S1 = fq_mul(Y1, ZZZ2)
    Pp = fq_mul_sub(2, X2, ZZ1, U1)
    Rr = fq_mul_sub(2, Y2, ZZZ1, S1)
    lt(Pp, 3.5), lt(Rr, 3.5)
    PP = fq_sqr(Pp)
    RR = fq_sqr(Rr)
    PPP = fq_mul(Pp, PP)
    Q = fq_mul(U1, PP)
    dbl = iz2(PP) and iz2(RR) and not p_id and not q_id
    T = [2 * Q[i] + PPP[i] for i in range(9)]
    X3 = fq_csub(3, fq_sub(5, RR, T))
    lt(X3, 4)
    Y3 = fq_mul_add(Rr, sub_lazy(4, Q, X3), S1, sub_lazy(2, limbs(0), PPP))
    lt(Y3, 1.5)
    r = (X3, Y3, fq_mul(fq_mul(ZZ1, ZZ2), PP), fq_mul(fq_mul(ZZZ1, ZZZ2), PPP))
    if dbl:
        X3, Y3, V, W = dbl_core(X1, Y1)
        r = (X3, Y3, fq_mul(V, ZZ1), fq_mul(W, ZZZ1))
    if p_id:
        r = q
    if q_id:
        r = acc
    return r


def to_int(x):  # field element -> internal limbs of x * R mod p, canonical
    return limbs(x * R % P)


def from_int(a):
    return val(a) * pow(R, -1, P) % P


def affine(pt):
    X, Y, ZZ, ZZZ = (from_int(c) for c in pt)
    if ZZ == 0:
        return None
    return (X * pow(ZZ, -1, P) % P, Y * pow(ZZZ, -1, P) % P)


def worst(k):
    """a value just below k*p with normalised limbs (k may be a half)"""
    return limbs(int(2 * k) * P // 2 - 1)


def test_mul_extremes():
    for a in (worst(8), worst(5), limbs(0), limbs(1), worst(1), worst(2)):
        for b in (worst(8), worst(4), limbs(P - 1), limbs(1)):
            if val(a) * val(b) < P * R // 2:
                fq_mul(a, b)
    # all-ones limbs (limbs < 2^30 as after a carry-free add) at a safe value
    a = [MASK] * 8 + [0]
    fq_mul(a, worst(8))
    for a in (worst(8), worst(5), worst(1), limbs(0), limbs(P - 1), [MASK] * 8 + [0]):
        fq_sqr(a)


def test_group_law_and_bounds():
    rnd = random.Random(7)
    pts = [B.mul(B.G, rnd.randrange(1, B.R_ORDER)) for _ in range(40)]
    acc = (limbs(0), limbs(0), limbs(0), limbs(0))
    ref = None
    for i, pt in enumerate(pts):
        q = pt if i % 5 else B.neg(pt)
        acc = madd(acc, to_int(q[0]), to_int(q[1]), False)
        ref = B.add(ref, q)
        assert affine(acc) == ref
    # doubling and cancellation through madd
    a1 = madd((limbs(0),) * 4, to_int(pts[0][0]), to_int(pts[0][1]), False)
    a2 = madd(a1, to_int(pts[0][0]), to_int(pts[0][1]), False)
    assert affine(a2) == B.add(pts[0], pts[0])
    ny = to_int(B.neg(pts[0])[1])
    a3 = madd(a1, to_int(pts[0][0]), ny, False)
    assert affine(a3) is None
    a4 = madd(a3, to_int(pts[1][0]), to_int(pts[1][1]), False)
    assert affine(a4) == pts[1]
    # XYZZ adds: random projective factors, doubling, cancellation, identity
    def proj(pt, z):
        zz, zzz = z * z % P, z * z * z % P
        return (to_int(pt[0] * zz % P), to_int(pt[1] * zzz % P), to_int(zz), to_int(zzz))

    for i in range(len(pts) - 1):
        s = add(proj(pts[i], rnd.randrange(1, P)), proj(pts[i + 1], rnd.randrange(1, P)))
        assert affine(s) == B.add(pts[i], pts[i + 1])
        s2 = add(s, s)
        assert affine(s2) == B.add(affine(s), affine(s))
    x = proj(pts[2], 5)
    assert affine(add(x, proj(pts[2], 9))) == B.add(pts[2], pts[2])
    assert affine(add(x, proj(B.neg(pts[2]), 9))) is None
    assert affine(add((limbs(0),) * 4, x)) == pts[2]


def test_fused_subtraction_extremes():
    """fq_mul_sub / fq_sqr_sub at the largest subtrahends the kernels pass"""
    for k, cb in ((5, 4), (5, 4.5), (2, 1.5)):
        for a in (worst(8), worst(6.5), worst(1), limbs(0)):
            for b in (worst(2), worst(6.5), limbs(1)):
                for c in (worst(cb), limbs(0), limbs(P - 1)):
                    if val(a) * val(b) < P * R // 2:
                        fq_mul_sub(k, a, b, c)
            fq_mul_sub(k, a, a, worst(cb), sqr=True)
    # the lazy forms: T = 2Q + PPP with all-ones limbs at its largest value,
    # and Q - X3 + 4p as the wide multiplicand against a normalised R
    qq, ppp = worst(1.5), worst(1.5)
    T = [2 * qq[i] + ppp[i] for i in range(9)]
    T_ones = [3 * MASK] * 8 + [T[8]]
    for r_ in (worst(6.5), limbs(P - 1)):
        fq_mul_sub(5, r_, r_, T, sqr=True, borrow=3)
        fq_mul_sub(5, r_, r_, T_ones, sqr=True, borrow=3)
        for x3 in (worst(6.5), limbs(0)):
            t = sub_lazy(7, [MASK] * 8 + [qq[8]], x3)
            fq_mul_sub(2, r_, t, worst(1.5))
    # a subtrahend past the stated bound is caught (planted on the CPU)
    try:
        fq_mul_sub(2, worst(1), worst(1), worst(2))
    except AssertionError:
        pass
    else:
        raise AssertionError("kpr check did not fire")


def test_run_addition_law():
    """The accumulation loop's form: runs from RUN_START, doubling and
    cancellation inside a run, bounds at every step."""
    rnd = random.Random(5)
    pts = [B.mul(B.G, rnd.randrange(1, B.R_ORDER)) for _ in range(30)]
    acc, aid, ref = RUN_START, True, None
    for i, pt in enumerate(pts):
        q = pt if i % 4 else B.neg(pt)
        acc, aid = madd_run(acc, aid, to_int(q[0]), to_int(q[1]))
        ref = B.add(ref, q)
        assert not aid and affine(acc) == ref
    # doubling: acc == q
    a1, f1 = madd_run(RUN_START, True, to_int(pts[0][0]), to_int(pts[0][1]))
    a2, f2 = madd_run(a1, f1, to_int(pts[0][0]), to_int(pts[0][1]))
    assert not f2 and affine(a2) == B.add(pts[0], pts[0])
    # cancellation: acc == -q, then the run goes on from the identity
    ny = to_int(B.neg(pts[0])[1])
    a3, f3 = madd_run(a1, f1, to_int(pts[0][0]), ny)
    assert f3 and affine(a3) is None
    a4, f4 = madd_run(a3, f3, to_int(pts[1][0]), to_int(pts[1][1]))
    assert not f4 and affine(a4) == pts[1]
    # the generator has x = 1: P = x2 - 0 is not 0 from RUN_START either way
    g1, _ = madd_run(RUN_START, True, to_int(B.G[0]), to_int(B.G[1]))
    assert affine(g1) == B.G


def test_bounds_at_extremes():
    """Accumulators at the top of their stated bounds still give the right
    group element (the formulas only see residues)."""
    rnd = random.Random(11)
    for _ in range(30):
        pt = B.mul(B.G, rnd.randrange(1, B.R_ORDER))
        q = B.mul(B.G, rnd.randrange(1, B.R_ORDER))
        z = rnd.randrange(1, P)
        zz, zzz = z * z % P, z * z * z % P
        X, Y = pt[0] * zz % P, pt[1] * zzz % P
        # lift each coordinate to the largest representative under its bound
        def lift(v, bound):
            r = v * R % P
            while r + P < bound * P:
                r += P
            return limbs(r)

        acc = (lift(X, 4), lift(Y, 4), lift(zz, 1.5), lift(zzz, 1.5))
        out = madd(acc, to_int(q[0]), to_int(q[1]), False)
        assert affine(out) == B.add(pt, q)
        out2 = add(acc, acc)
        assert affine(out2) == B.add(pt, pt)
        out3, f = madd_run(acc, False, to_int(q[0]), to_int(q[1]))
        assert not f and affine(out3) == B.add(pt, q)
        # the loop's own X bound (< 6.5p), then a doubling and a general add
        acc65 = (lift(X, 6.5), lift(Y, 4), lift(zz, 2), lift(zzz, 2))
        out4, f = madd_run(acc65, False, to_int(q[0]), to_int(q[1]))
        assert not f and affine(out4) == B.add(pt, q)
        out5, _ = madd_run(out4, False, to_int(q[0]), to_int(q[1]))
        assert affine(out5) == B.add(B.add(pt, q), q)
        assert affine(add(out4, out4)) == B.add(affine(out4), affine(out4))


def jac_dbl(X, Y, Z):
    """Arith::qjac_dbl (dbl-2009-l, a = 0): X, Y, Z < 2p in and out."""
    lt(X, 2), lt(Y, 2), lt(Z, 2)
    A, B_ = fq_sqr(X), fq_sqr(Y)
    XB = fq_add(X, B_)
    C, T = fq_sqr(B_), fq_sqr(XB)
    D0 = fq_sub(4, T, fq_add(A, C))
    D = fq_add(D0, D0)
    lt(D, 11)
    D = fq_csub(2, fq_csub(4, fq_csub(8, D)))
    lt(D, 2)
    E = fq_add(fq_add(A, A), A)
    lt(E, 4.5)
    Y2 = fq_add(Y, Y)
    F, Z3 = fq_mul(E, E), fq_mul(Y2, Z)
    X3 = fq_csub(2, fq_csub(4, fq_sub(4, F, fq_add(D, D))))
    lt(X3, 2)
    C8 = fq_add(C, C)
    C8 = fq_add(C8, C8)
    C8 = fq_add(C8, C8)
    lt(C8, 12)
    Y3 = fq_mul_sub(13, E, fq_sub(2, D, X3), C8)
    lt(Y3, 14.5)
    Y3 = fq_csub(2, fq_csub(4, fq_csub(8, Y3)))
    lt(Y3, 2), lt(Z3, 1.5)
    return X3, Y3, Z3


def dbl_k(pt, k):
    """Arith::qxyzz_dbl_k: XYZZ -> Jacobian (X ZZ^2, Y ZZZ^2, ZZZ), k
    doublings, back to XYZZ (Z^2, Z^3)."""
    X, Y, ZZ, ZZZ = pt
    lt(X, 6.5), lt(Y, 4), lt(ZZ, 2), lt(ZZZ, 2)
    ZZ2, ZZZ2 = fq_sqr(ZZ), fq_sqr(ZZZ)
    X, Y, Z = fq_mul(X, ZZ2), fq_mul(Y, ZZZ2), ZZZ
    for _ in range(k):
        X, Y, Z = jac_dbl(X, Y, Z)
    ZZ = fq_sqr(Z)
    return X, Y, ZZ, fq_mul(ZZ, Z)


def test_repeated_doubling_chain():
    """The device tail's 2^k P (Jacobian chain between XYZZ forms) equals
    2^k P of the group law, from accumulators at the top of their bounds and
    from the identity, with every bound checked on the way."""
    rnd = random.Random(13)

    def lift(v, bound):
        r = v * R % P
        while r + P < bound * P:
            r += P
        return limbs(r)

    for k in (1, 2, 13, 16):
        for _ in range(4):
            pt = B.mul(B.G, rnd.randrange(1, B.R_ORDER))
            z = rnd.randrange(1, P)
            zz, zzz = z * z % P, z * z * z % P
            acc = (lift(pt[0] * zz % P, 6.5), lift(pt[1] * zzz % P, 4), lift(zz, 2), lift(zzz, 2))
            assert affine(dbl_k(acc, k)) == B.mul(pt, 1 << k)
        idp = (lift(0, 1), limbs(5), lift(0, 2), lift(0, 2))  # ZZ = ZZZ = p: the identity
        assert affine(dbl_k(idp, k)) is None
    # a wrong constant in the chain is caught: 8C replaced by 4C
    pt = B.mul(B.G, 12345)
    X, Y, Z = to_int(pt[0]), to_int(pt[1]), to_int(1)
    A, B_ = fq_sqr(X), fq_sqr(Y)
    C = fq_sqr(B_)
    D0 = fq_sub(4, fq_sqr(fq_add(X, B_)), fq_add(A, C))
    D = fq_csub(2, fq_csub(4, fq_csub(8, fq_add(D0, D0))))
    E = fq_add(fq_add(A, A), A)
    X3 = fq_csub(2, fq_csub(4, fq_sub(4, fq_mul(E, E), fq_add(D, D))))
    C4 = fq_add(fq_add(C, C), fq_add(C, C))
    Y3 = fq_mul_sub(13, E, fq_sub(2, D, X3), C4)
    Z3 = fq_mul(fq_add(Y, Y), Z)
    z3 = from_int(Z3)
    bad = (from_int(X3) * pow(z3 * z3, -1, P) % P, from_int(Y3) * pow(z3 * z3 * z3, -1, P) % P)
    assert bad != B.add(pt, pt)

