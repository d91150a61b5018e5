"""CPU tests of bench.py's contract pieces that need no GPU: the one-GPU
refusal and the algorithmic multiply counts its roofline divides by."""
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
