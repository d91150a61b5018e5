"""oracle/bn254_g2.py -- TEST INFRASTRUCTURE ONLY (pure-Python big integers).

Restatement of BN254 G2, the group the reference's BN254G2MSM runs on
(Sources/zkMetal/MSM/BN254G2MSMEngine.swift:289-540;
Sources/zkMetal/Curve/BN254Pairing.swift:11-80, 300-396): the sextic twist
y^2 = x^3 + 3 / (9 + u) over Fp2 = Fp[u] / (u^2 + 1), restricted to its
subgroup of prime order r.  Coordinates are Fp2 elements c0 + c1 u, each
component in the 2^256 Montgomery form of BN254 Fp; scalars are reduced mod r
(BN254G2MSMEngine.swift:214-232).  Only tests/ may import it.

What pins what:
  * Fp2: u^2 = -1, the field axioms on random elements, inverses;
  * the generator: the reference's literals (BN254Pairing.swift:389-396, the
    EIP-197 / go-ethereum generator) lie on the twist and have order r
    (r G2 = O, 2 G2 != O);
  * the group law: associativity / commutativity, k G2 by double-and-add
    against sums of G2, doubling against the chord rule's tangent case;
  * MSM: closed form (sum s_i k_i mod r) * G2 for bases k_i * G2, against
    brute force on small cases.
The C G2 Pippenger in oracle/bn254_ref.c is pinned against this module
(tests/test_oracle_g2.py): its point operations at random Z against the
affine group law here, including the doubling, inverse and identity cases,
and its MSM against the closed form and brute force.
"""
from . import bn254

P = bn254.P
R_ORDER = bn254.R_ORDER
MONT_R = pow(2, 256, P)


# ---- Fp2 = Fp[u] / (u^2 + 1), elements as (c0, c1)

def f2(c0, c1=0):
    return (c0 % P, c1 % P)


def add2(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def sub2(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def neg2(a):
    return ((-a[0]) % P, (-a[1]) % P)


def mul2(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def inv2(a):
    n = (a[0] * a[0] + a[1] * a[1]) % P
    ni = pow(n, -1, P)
    return (a[0] * ni % P, (-a[1]) * ni % P)


ZERO2 = (0, 0)
ONE2 = (1, 0)
XI = (9, 1)                   # the non-residue 9 + u
B2 = mul2((3, 0), inv2(XI))   # twist coefficient 3 / (9 + u)

# the reference's generator (BN254Pairing.swift:389-396)
G2 = (
    (0x1800deef121f1e76426a00665e5c4479674322d4f75edadd46debd5cd992f6ed,
     0x198e9393920d483a7260bfb731fb5d25f1aa493335a9e71297e485b7aef312c2),
    (0x12c85ea5db8c6deb4aab71808dcb408fe3d1e7690c43d37b4ce6cc0166fa7daa,
     0x090689d0585ff075ec9e99ad690c3395bc4b313370b38ef355acdadcd122975b),
)


def on_curve(pt):
    if pt is None:
        return True
    x, y = pt
    return sub2(mul2(y, y), add2(mul2(mul2(x, x), x), B2)) == ZERO2


def neg(pt):
    return None if pt is None else (pt[0], neg2(pt[1]))


def add(a, b):
    """affine chord-and-tangent law; None is the identity"""
    if a is None:
        return b
    if b is None:
        return a
    (x1, y1), (x2, y2) = a, b
    if x1 == x2:
        if add2(y1, y2) == ZERO2:
            return None
        lam = mul2(mul2((3, 0), mul2(x1, x1)), inv2(add2(y1, y1)))
    else:
        lam = mul2(sub2(y2, y1), inv2(sub2(x2, x1)))
    x3 = sub2(sub2(mul2(lam, lam), x1), x2)
    return (x3, sub2(mul2(lam, sub2(x1, x3)), y1))


def mul(pt, k):
    k %= R_ORDER
    acc, q = None, pt
    while k:
        if k & 1:
            acc = add(acc, q)
        q = add(q, q)
        k >>= 1
    return acc


def msm_closed_form(ks, scalars):
    """MSM over bases k_i * G2: (sum s_i k_i mod r) * G2."""
    return mul(G2, sum(s * k for s, k in zip(scalars, ks)) % R_ORDER)


def to_mont(x):
    return x * MONT_R % P


def from_mont(x):
    return x * pow(MONT_R, -1, P) % P


def fe(x):
    return int(x).to_bytes(32, "little")


def point_to_bytes(pt):
    """128-byte base as the library takes it: x.c0, x.c1, y.c0, y.c1, each in
    Fp Montgomery form (8 x u32 LE); 128 zero bytes = the identity."""
    if pt is None:
        return bytes(128)
    (x0, x1), (y0, y1) = pt
    return b"".join(fe(to_mont(v)) for v in (x0, x1, y0, y1))


def bytes_to_point(b):
    v = [from_mont(int.from_bytes(bytes(b[32 * i:32 * i + 32]), "little")) for i in range(4)]
    if bytes(b[:128]) == bytes(128):
        return None
    return ((v[0], v[1]), (v[2], v[3]))


def jacobian_bytes_to_affine(r192):
    """192-byte Jacobian result X | Y | Z (each c0 | c1, Fp Montgomery) ->
    affine (None = identity)."""
    v = [from_mont(int.from_bytes(bytes(r192[32 * i:32 * i + 32]), "little")) for i in range(6)]
    X, Y, Z = (v[0], v[1]), (v[2], v[3]), (v[4], v[5])
    if Z == ZERO2:
        return None
    zi = inv2(Z)
    zi2 = mul2(zi, zi)
    return (mul2(X, zi2), mul2(Y, mul2(zi2, zi)))
