"""The benchmarked configurations' expected results (tests/golden/configs.json,
written by tests/golden/make_fixtures.py --configs in pure Python from the
seeds) recomputed a second way on the CPU: the library's host scalar
generator (zkmsm_gen_scalars) and the C oracle's sum(s_i k_i) mod the group
order, then one scalar multiple of the generator.  The GPU side of the same
fixture is in tests/test_gpu_configs.py."""
import pytest

import zkmetal_amd as zk
from oracle import bn254 as B
from oracle import bn254_g2 as G2
from oracle import grumpkin as GK
from oracle import ref
from helpers import config_expected_point as expected_point, config_fixture as load


def closed_form(case, seed_bases, seed_scalars):
    n = case["n"]
    ks = zk.gen_scalars(seed_bases, n, 0)
    ss = zk.gen_scalars(seed_scalars, n, 0)
    if case["curve"] == "grumpkin":
        return GK.mul(GK.G, ref.scalar_dot(ss.tobytes(), ks.tobytes(), n, curve="grumpkin"))
    dot = ref.scalar_dot(ss.tobytes(), ks.tobytes(), n)
    return G2.mul(G2.G2, dot) if case["curve"] == "g2" else B.mul(B.G, dot)


@pytest.mark.parametrize("idx", range(5))
def test_config_fixture_recomputed_from_the_seeds(idx):
    fx = load()
    case = fx["cases"][idx]
    assert closed_form(case, fx["seed_bases"], fx["seed_scalars"]) == expected_point(case)


def test_config_fixture_covers_the_benchmarked_sizes_and_differs_per_curve():
    fx = load()
    got = {(c["curve"], c["n"]) for c in fx["cases"]}
    assert got == {("bn254", 1 << 16), ("bn254", 1 << 20), ("bn254", 1 << 22),
                   ("grumpkin", 1 << 22), ("g2", 1 << 22)}
    # a wrong seed gives a different point (the fixture is not seed-blind)
    case = fx["cases"][0]
    assert closed_form(case, fx["seed_bases"] + 1, fx["seed_scalars"]) != expected_point(case)
