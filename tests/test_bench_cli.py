"""CPU tests of bench.py's contract pieces that need no GPU: the one-GPU
refusal, the algorithmic multiply counts its roofline divides by, and the
G2 CPU-baseline leg."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one MI355X" in r.stderr and r.stdout == ""


def test_algorithmic_counts_match_the_published_table():
    # F(n, c*) = 10 (2n) W + 28 W 2^(c-1) + n minimised over c: 1.43e7, 1.76e8
    # and 6.43e8 field multiplies at 2^16, 2^20, 2^22 (c* = 13, 16, 19)
    assert round(bench.algorithmic_field_muls(1 << 16) / 1e7, 2) == 1.43
    assert round(bench.algorithmic_field_muls(1 << 20) / 1e8, 2) == 1.76
    assert round(bench.algorithmic_field_muls(1 << 22) / 1e8, 2) == 6.43


def test_g2_cpu_baseline_leg_compares_in_affine():
    import random
    from oracle import bn254_g2 as G
    rnd = random.Random(5)
    ks = [rnd.randrange(1, G.R_ORDER) for _ in range(6)]
    ss = [rnd.randrange(2**256) for _ in range(6)]
    pts = b"".join(G.point_to_bytes(G.mul(G.G2, k)) for k in ks)
    sc = b"".join(s.to_bytes(32, "little") for s in ss)
    # a Jacobian result as the library returns it, at Z = 2
    x, y = G.msm_closed_form(ks, ss)
    z = (2, 0)
    X, Y = G.mul2(x, G.mul2(z, z)), G.mul2(y, G.mul2(G.mul2(z, z), z))
    good = b"".join(G.fe(G.to_mont(v)) for v in (*X, *Y, *z))
    line = bench.cpu_baseline(pts, sc, 6, good, 2, curve="g2")
    assert line["matches_gpu"] is True and line["kind"] == "port" and line["cores"] == 2
    bad = b"".join(G.fe(G.to_mont(v)) for v in (*X, *G.neg2(Y), *z))
    assert bench.cpu_baseline(pts, sc, 6, bad, 2, curve="g2")["matches_gpu"] is False
