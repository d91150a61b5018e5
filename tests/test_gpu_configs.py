"""GPU runs of the benchmarked configurations (plain BN254 G1 MSM at 2^16,
2^20, 2^22; Grumpkin G1 and BN254 G2 at 2^22; bases and scalars from the
seeds bench.py uses, resident in device memory) against the committed
expected results in tests/golden/configs.json."""
import pytest

import zkmetal_amd as zk
from oracle import bn254_g2 as G2
from oracle import grumpkin as GK
from helpers import config_expected_point as expected_point, config_fixture as load, result_affine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    e = zk.MetalMSM()
    yield e
    e.close()


@pytest.mark.parametrize("idx", range(5))
def test_config_matches_the_golden_fixture(engine, idx):
    fx = load()
    case = fx["cases"][idx]
    n, curve = case["n"], case["curve"]
    print(f"config {curve} n={n}", flush=True)
    sc = zk.gen_scalars(fx["seed_scalars"], n, 0)
    ds = zk.DeviceBuffer(sc.nbytes)
    ds.upload(sc)
    try:
        if curve == "bn254":
            engine.srs_generate(fx["seed_bases"], n)
            bases = engine.srs_download(n)
            dp = zk.DeviceBuffer(bases.nbytes)
            dp.upload(bases)
            try:
                got = result_affine(engine.msm_batch_device(dp, ds, [n])[0])
            finally:
                dp.free()
        elif curve == "grumpkin":
            dp = zk.DeviceBuffer(64 * n)
            try:
                engine.grumpkin_points_generate(fx["seed_bases"], n, dp)
                got = GK.jacobian_bytes_to_affine(engine.grumpkin_msm_batch_device(dp, ds, [n])[0])
            finally:
                dp.free()
        else:
            dp = zk.DeviceBuffer(128 * n)
            try:
                engine.g2_points_generate(fx["seed_bases"], n, dp)
                got = G2.jacobian_bytes_to_affine(engine.g2_msm_batch_device(dp, ds, [n])[0])
            finally:
                dp.free()
    finally:
        ds.free()
    assert got == expected_point(case)
