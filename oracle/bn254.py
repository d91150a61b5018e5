"""oracle/bn254.py -- TEST INFRASTRUCTURE ONLY (pure-Python big integers).

Independent restatement of the BN254 G1 arithmetic the reference MSM path
relies on, used to pin the C oracle (oracle/bn254_ref.c) and to produce the
golden fixtures under tests/golden/.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import anything under oracle/.

What pins what (no function here is checked by retyping its own formula):
  * field/Montgomery constants: recomputed from p and compared with the
    reference's literals (Sources/NeonFieldOps/bn254_msm.c:24-32,
    Sources/zkMetal/Fields/BN254Fp.swift:11-29);
  * curve arithmetic: EIP-196 known answers held by the reference
    (Tests/zkMetalTests/EVMPrecompileTests.swift:35-46,129-180): G=(1,2),
    2G, -G=(1,p-2), 0*G=O; plus r*G = O and on-curve checks;
  * MSM: closed form sum(s_i*k_i mod r)*G for bases k_i*G, and brute force
    sum of double-and-add products on small n;
  * GLV (Sources/zkMetal/MSM/GLV.swift:37-92): k == +-k1 + lambda*(+-k2) mod r
    and phi(P) = (beta*x, y) == lambda*P on G (beta^3 == 1);
  * signed digits (Sources/Shaders/msm/msm_kernels.metal:399-439):
    sum d_w 2^(c w) == k, |d_w| <= 2^(c-1), d_w != -2^(c-1).
"""

P = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
R_ORDER = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
B = 3
G = (1, 2)
MONT_R = pow(2, 256, P)
MONT_R2 = pow(2, 512, P)
MONT_INV64 = (-pow(P, -1, 2**64)) % 2**64

# beta in Montgomery form, as stored by the reference (GLV.swift:14-17)
BETA_MONT_LIMBS64 = [0x71930C11D782E155, 0xA6BB947CFFBE3323, 0xAA303344D4741444, 0x2C3B3F0D26594943]
# GLV lattice constants (GLV.swift:20-25)
GLV_A1 = 0x89D3256894D213E3
GLV_MINUS_B1 = 0x6F4D8248EEB859FC8211BBEB7D4F1128
GLV_A2 = 0x6F4D8248EEB859FD0BE4E1541221250B
GLV_B2 = 0x89D3256894D213E3
GLV_G1 = 0x24CCEF014A773D2CF7A7BD9D4391EB18D
GLV_G2 = 0x2D91D232EC7E0B3D7

# EIP-196 vectors held by the reference (EVMPrecompileTests.swift:35-46)
EIP196_2G = (0x030644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD3,
             0x15ED738C0E0A7C92E7845F96B2AE9C0A68A6A449E3538FC7FF3EBF7A5A18A2C4)
EIP196_NEG_G_Y = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD45


def limbs_from_int(x, n=8, bits=32):
    m = (1 << bits) - 1
    return [(x >> (bits * i)) & m for i in range(n)]


def int_from_limbs(limbs, bits=32):
    v = 0
    for i, l in enumerate(limbs):
        v |= int(l) << (bits * i)
    return v


def to_mont(x):
    return (x * MONT_R) % P


def from_mont(x):
    return (x * pow(MONT_R, -1, P)) % P


BETA = from_mont(int_from_limbs(BETA_MONT_LIMBS64, bits=64))


def _lambda():
    # root of l^2 + l + 1 == 0 mod r paired with beta: phi(G) == l*G
    s = pow(R_ORDER - 3, (R_ORDER + 1) // 4, R_ORDER) if R_ORDER % 4 == 3 else _tonelli(R_ORDER - 3, R_ORDER)
    inv2 = pow(2, -1, R_ORDER)
    cands = [((-1 + s) * inv2) % R_ORDER, ((-1 - s) * inv2) % R_ORDER]
    for lam in cands:
        if mul(G, lam) == (BETA * G[0] % P, G[1]):
            return lam
    raise AssertionError("no lambda pairs with beta")


def _tonelli(n, p):
    q, s = p - 1, 0
    while q % 2 == 0:
        q //= 2
        s += 1
    z = 2
    while pow(z, (p - 1) // 2, p) != p - 1:
        z += 1
    m, c, t, r = s, pow(z, q, p), pow(n, q, p), pow(n, (q + 1) // 2, p)
    while t != 1:
        i, t2 = 0, t
        while t2 != 1:
            t2 = t2 * t2 % p
            i += 1
        b = pow(c, 1 << (m - i - 1), p)
        m, c, t, r = i, b * b % p, t * b * b % p, r * b % p
    return r


# ---- affine arithmetic, None = point at infinity ----

def on_curve(pt):
    if pt is None:
        return True
    x, y = pt
    return (y * y - x * x * x - B) % P == 0


def neg(pt):
    return None if pt is None else (pt[0], (-pt[1]) % P)


def add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    x1, y1 = a
    x2, y2 = b
    if x1 == x2:
        if (y1 + y2) % P == 0:
            return None
        lam = 3 * x1 * x1 * pow(2 * y1, -1, P) % P
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, P) % P
    x3 = (lam * lam - x1 - x2) % P
    return (x3, (lam * (x1 - x3) - y1) % P)


def mul(pt, k):
    acc, base = None, pt
    while k > 0:
        if k & 1:
            acc = add(acc, base)
        base = add(base, base)
        k >>= 1
    return acc


def msm_bruteforce(points, scalars):
    acc = None
    for pt, s in zip(points, scalars):
        acc = add(acc, mul(pt, s))
    return acc


def msm_closed_form(ks, scalars):
    """MSM over bases k_i*G: (sum s_i k_i mod r) * G."""
    return mul(G, sum(s * k for s, k in zip(scalars, ks)) % R_ORDER)


# ---- byte / limb layouts used by the C ABI ----

def affine_to_bytes(pt):
    """64 B: x, y as 8 x u32 LE Montgomery; infinity -> all zero (0,0)."""
    if pt is None:
        return bytes(64)
    x, y = pt
    return to_mont(x).to_bytes(32, "little") + to_mont(y).to_bytes(32, "little")


def scalar_to_bytes(s):
    return int(s).to_bytes(32, "little")


def jacobian_bytes_to_affine(xb, yb, zb):
    """Jacobian (X, Y, Z) Montgomery bytes -> canonical affine or None."""
    X = from_mont(int.from_bytes(bytes(xb), "little"))
    Y = from_mont(int.from_bytes(bytes(yb), "little"))
    Z = from_mont(int.from_bytes(bytes(zb), "little"))
    if Z == 0:
        return None
    zi = pow(Z, -1, P)
    return (X * zi * zi % P, Y * zi * zi * zi % P)


# ---- GLV (GLV.swift:37-92) and signed digits (msm_kernels.metal:399-439) ----

def glv_decompose(k):
    k %= R_ORDER
    c1 = (k * GLV_G1) >> 256
    c2 = (k * GLV_G2) >> 256
    k1 = (k - c2 * GLV_A1 - c1 * GLV_A2) % R_ORDER
    k2 = c2 * GLV_MINUS_B1 - c1 * GLV_B2
    neg2 = k2 < 0
    k2 = -k2 if neg2 else k2
    neg1 = k1 >= (R_ORDER >> 1)
    if neg1:
        k1 = R_ORDER - k1
    return k1, k2, neg1, neg2


def signed_digits(k, c, nwin):
    """Returns (digits, final_carry)."""
    half, full, mask = 1 << (c - 1), 1 << c, (1 << c) - 1
    out, carry = [], 0
    for w in range(nwin):
        d = ((k >> (w * c)) & mask) + carry
        carry = 0
        if d > half:
            out.append(d - full)
            carry = 1
        else:
            out.append(d)
    return out, carry


# ---- seeded synthetic inputs (same definition as the library generator) ----

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def mix64(z):
    z &= M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def word(seed, i, j):
    """Word j (0..63) of element i: splitmix64 output at counter i*64+j+1."""
    return mix64((seed + GOLDEN * (i * 64 + j + 1)) & M64)


def uniform_fr(seed, i):
    for a in range(16):
        w = [word(seed, i, 4 * a + t) for t in range(4)]
        w[3] &= (1 << 62) - 1
        v = w[0] | (w[1] << 64) | (w[2] << 128) | (w[3] << 192)
        if v < R_ORDER:
            return v
    return v % R_ORDER


DIST_UNIFORM_FR, DIST_SKEWED, DIST_FULL256 = 0, 1, 2


def gen_scalar(seed, i, dist):
    if dist == DIST_UNIFORM_FR:
        return uniform_fr(seed, i)
    if dist == DIST_FULL256:
        w = [word(seed, i, t) for t in range(4)]
        return w[0] | (w[1] << 64) | (w[2] << 128) | (w[3] << 192)
    cls = word(seed, i, 60) % 100
    if cls < 40:
        return 0
    if cls < 70:
        return word(seed, i, 61) & 0xFFFF
    if cls < 80:
        return R_ORDER - 1 if word(seed, i, 62) & 1 else 1
    return uniform_fr(seed, i)


LAMBDA = _lambda()
