"""Write the golden MSM fixtures under tests/golden/ (pure-Python big integers).

    python tests/golden/make_fixtures.py
    python tests/golden/make_fixtures.py --configs   (configs.json, ~1 min)

Each fixture holds the raw ABI inputs (64 B Montgomery affine bases, 32 B
standard-form scalars) and the expected result in canonical affine
coordinates.  Expected values come from oracle/bn254.py and are computed two
independent ways where the inputs allow it: the closed form
(sum s_i k_i mod r) * G for bases k_i * G, and brute-force double-and-add.
Edge cases follow the reference's own C-Pippenger tests
(Sources/zkbench/correctness_tests.swift:1701-1840) and EIP-196 vectors
(Tests/zkMetalTests/EVMPrecompileTests.swift:35-46).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from oracle import bn254 as B  # noqa: E402
from oracle import bn254_g2 as G2  # noqa: E402
from oracle import grumpkin as GK  # noqa: E402


def lcg_scalars(seed, n):
    """The reference harness LCG (MSMTests.swift:18-23): 8 limbs, high 32 bits."""
    out, rng = [], seed
    for _ in range(n):
        limbs = []
        for _ in range(8):
            rng = (rng * 6364136223846793005 + 1442695040888963407) & B.M64
            limbs.append(rng >> 32)
        out.append(B.int_from_limbs(limbs))
    return out


def fixture(name, points, scalars, note, ks=None):
    assert len(points) == len(scalars)
    brute = B.msm_bruteforce(points, scalars)
    if ks is not None:
        closed = B.msm_closed_form(ks, scalars)
        assert closed == brute, name
    exp = {"inf": brute is None}
    if brute is not None:
        assert B.on_curve(brute)
        exp["x"] = "%064x" % brute[0]
        exp["y"] = "%064x" % brute[1]
    return {
        "name": name,
        "note": note,
        "n": len(points),
        "points": b"".join(B.affine_to_bytes(p) for p in points).hex(),
        "scalars": b"".join(B.scalar_to_bytes(s) for s in scalars).hex(),
        "expected": exp,
    }


def config_fixtures(seed=0x5EED):
    """Expected results of the benchmarked configurations, from their seeds:
    bases k_i * G (G2, Grumpkin G) with k_i = gen_scalar(seed, i, 0), scalars
    gen_scalar(seed + 1, i, 0), i < n.  Element i does not depend on n, so
    one pass gives every size as a prefix; the integer sum sum(s_i k_i) is
    reduced mod r (BN254 G1, G2) or mod p (Grumpkin's group order)."""
    sizes = {1 << 16: None, 1 << 20: None, 1 << 22: None}
    acc = 0
    for i in range(1 << 22):
        acc += B.uniform_fr(seed, i) * B.uniform_fr(seed + 1, i)
        if i + 1 in sizes:
            sizes[i + 1] = acc
    cases = []

    def g1(pt):
        return {"inf": pt is None} if pt is None else {"inf": False, "x": "%064x" % pt[0], "y": "%064x" % pt[1]}

    for n, dot in sizes.items():
        cases.append({"curve": "bn254", "n": n, "expected": g1(B.mul(B.G, dot % B.R_ORDER))})
    dot = sizes[1 << 22]
    cases.append({"curve": "grumpkin", "n": 1 << 22, "expected": g1(GK.mul(GK.G, dot % GK.ORDER))})
    q = G2.mul(G2.G2, dot % G2.R_ORDER)
    cases.append({"curve": "g2", "n": 1 << 22,
                  "expected": {"inf": False, "x": ["%064x" % v for v in q[0]], "y": ["%064x" % v for v in q[1]]}})
    return {
        "note": "plain MSMs as bench.py runs them: bases k_i * G with k_i = gen_scalars(seed_bases), "
                "scalars gen_scalars(seed_scalars), uniform in Fr; expected canonical affine (not Montgomery)",
        "seed_bases": seed,
        "seed_scalars": seed + 1,
        "cases": cases,
    }


def main():
    if "--configs" in sys.argv:
        with open(os.path.join(HERE, "configs.json"), "w") as fh:
            json.dump(config_fixtures(), fh, indent=1)
        print("wrote configs.json")
        return
    fx = []
    seed_k, seed_s = 0x5EED0001, 0x5EED0002
    ks = [B.uniform_fr(seed_k, i) for i in range(256)]
    bases = [B.mul(B.G, k) for k in ks]
    for n in (1, 2, 3, 4, 7, 16, 64, 200):
        sc = [B.uniform_fr(seed_s, i) for i in range(n)]
        fx.append(fixture(f"uniform_n{n}", bases[:n], sc, "bases k_i*G, scalars uniform in Fr", ks[:n]))
    # the reference harness: bases i*G, LCG scalars (may exceed r)
    n = 256
    lin = [B.mul(B.G, i + 1) for i in range(n)]
    sc = lcg_scalars(0xDEADBEEFCAFEBABE, n)
    fx.append(fixture("lcg_n256_bases_iG", lin, sc, "MSMTests.swift:34-62 generator", [i + 1 for i in range(n)]))
    sc = [B.gen_scalar(0x5EED0003, i, B.DIST_SKEWED) for i in range(128)]
    fx.append(fixture("skewed_n128", bases[:128], sc, "40% 0, 30% <2^16, 10% {1,r-1}, 20% uniform", ks[:128]))
    sc = [B.gen_scalar(0x5EED0004, i, B.DIST_FULL256) for i in range(32)]
    fx.append(fixture("full256_n32", bases[:32], sc, "uniform 256-bit scalars, unreduced", ks[:32]))
    P0, P1 = bases[0], bases[1]
    fx.append(fixture("same_base_1234", [P0] * 4, [1, 2, 3, 4], "4 equal bases -> 10 P (doubling in a bucket)", [ks[0]] * 4))
    fx.append(fixture("identity_bases", [P0, None, P1, None], [3, 5, 7, 11], "(0,0) bases are skipped"))
    fx.append(fixture("zero_scalars", bases[:5], [0] * 5, "all-zero scalars -> identity", ks[:5]))
    fx.append(fixture("scalar_one", [P0], [1], "1*P = P", ks[:1]))
    fx.append(fixture("scalar_zero", [P0], [0], "0*P = O", ks[:1]))
    fx.append(fixture("r_minus_1", [P0, P1], [B.R_ORDER - 1, B.R_ORDER - 1], "(r-1) P = -P", ks[:2]))
    fx.append(fixture("unreduced", [P0, P1], [(1 << 256) - 1, 0x3FFFFFFF << 224 | ((1 << 224) - 1)], "scalars >= r", ks[:2]))
    fx.append(fixture("bit252", [P0, P1], [1 << 252, 1], "high-bit scalar", ks[:2]))
    fx.append(fixture("cancel", [P0, B.neg(P0), P1], [5, 5, 0], "5P + 5(-P) = O"))
    fx.append(fixture("equal_scalars", bases[:8], [42] * 8, "all scalars equal", ks[:8]))
    fx.append(fixture("generator_small", [B.G, B.G, B.G], [1, 2, 2], "5 G from G (EIP-196 base)", [1, 1, 1]))
    for f in fx:
        path = os.path.join(HERE, f"msm_{f['name']}.json")
        with open(path, "w") as fh:
            json.dump(f, fh, indent=1)
    eip = {
        "source": "Tests/zkMetalTests/EVMPrecompileTests.swift:35-46",
        "G": ["%064x" % B.G[0], "%064x" % B.G[1]],
        "twoG": ["%064x" % B.EIP196_2G[0], "%064x" % B.EIP196_2G[1]],
        "negG_y": "%064x" % B.EIP196_NEG_G_Y,
    }
    with open(os.path.join(HERE, "eip196.json"), "w") as fh:
        json.dump(eip, fh, indent=1)
    print(f"wrote {len(fx)} MSM fixtures")


if __name__ == "__main__":
    main()
