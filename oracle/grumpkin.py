"""oracle/grumpkin.py -- TEST INFRASTRUCTURE ONLY (pure-Python big integers).

Restatement of the Grumpkin curve the reference's GrumpkinMSM runs on
(Sources/zkMetal/Curve/GrumpkinCurve.swift:1-8, 206-240;
Sources/NeonFieldOps/grumpkin_ops.c:14-25): y^2 = x^3 - 17 over BN254's
scalar field r, group order BN254's base-field prime p (the BN254 cycle).
Coordinates are stored in Fr Montgomery form (x * 2^256 mod r), scalars are
reduced mod p (GrumpkinMSMEngine.swift:178-196).  Only tests/ may import it.

What pins what:
  * the field constants: recomputed from r and compared with the
    reference's literals (grumpkin_ops.c:17-25: r, -r^-1 mod 2^64, R mod r);
  * the generator: (1, sqrt(-16)) with the y the reference quotes from the
    Aztec/Noir specification (GrumpkinCurve.swift:215-222), on the curve;
  * the group order: p * G = O and 2 G != O (so G has order p, p prime);
  * the group law: affine formulas checked by associativity / commutativity
    and k*G by double-and-add against sums of G;
  * MSM: closed form (sum s_i k_i mod p) * G for bases k_i * G.
"""
from . import bn254

R = bn254.R_ORDER           # base field of Grumpkin
ORDER = bn254.P             # group order of Grumpkin
B = (-17) % R
GEN_Y = 17631683881184975370165255887551781615748388533673675138860
G = (1, GEN_Y)
MONT_R = pow(2, 256, R)

# the reference's literals (grumpkin_ops.c:17-25), as 64-bit limbs
REF_P_LIMBS64 = [0x43E1F593F0000001, 0x2833E84879B97091, 0xB85045B68181585D, 0x30644E72E131A029]
REF_INV64 = 0xC2E1F593EFFFFFFF
REF_ONE_LIMBS64 = [0xAC96341C4FFFFFFB, 0x36FC76959F60CD29, 0x666EA36F7879462E, 0x0E0A77C19A07DF2F]


def to_mont(x):
    return (x * MONT_R) % R


def from_mont(x):
    return (x * pow(MONT_R, -1, R)) % R


def on_curve(pt):
    if pt is None:
        return True
    x, y = pt
    return (y * y - x * x * x - B) % R == 0


def neg(pt):
    return None if pt is None else (pt[0], (-pt[1]) % R)


def add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    (x1, y1), (x2, y2) = a, b
    if x1 == x2:
        if (y1 + y2) % R == 0:
            return None
        lam = 3 * x1 * x1 * pow(2 * y1, -1, R) % R
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, R) % R
    x3 = (lam * lam - x1 - x2) % R
    return (x3, (lam * (x1 - x3) - y1) % R)


def mul(pt, k):
    k %= ORDER
    acc, q = None, pt
    while k:
        if k & 1:
            acc = add(acc, q)
        q = add(q, q)
        k >>= 1
    return acc


def msm_closed_form(ks, scalars):
    """MSM over bases k_i * G: (sum s_i k_i mod p) * G."""
    return mul(G, sum(s * k for s, k in zip(scalars, ks)) % ORDER)


def point_to_bytes(pt):
    """64-byte base as the library takes it: x, y in Fr Montgomery form; (0,0) = O."""
    if pt is None:
        return bytes(64)
    return to_mont(pt[0]).to_bytes(32, "little") + to_mont(pt[1]).to_bytes(32, "little")


def jacobian_bytes_to_affine(r96):
    """96-byte Jacobian result (Fr Montgomery) -> affine (None = identity)."""
    x, y, z = (from_mont(int.from_bytes(bytes(r96[32 * i:32 * i + 32]), "little")) for i in range(3))
    if z == 0:
        return None
    zi = pow(z, -1, R)
    return (x * zi * zi % R, y * zi * zi * zi % R)
