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


@pytest.mark.parametrize("detect", [True, False])
def test_drop_in_host_call_at_2_22(engine, detect):
    """The zkMetal entry point a caller binds (zkmetal_bn254_msm, host
    buffers, PCIe included) at the headline size: with the bases recognised
    as the resident SRS (no base upload), and with the recognition off (bases
    uploaded and the table built for the call), against the golden result."""
    fx = load()
    case = next(c for c in fx["cases"] if c["curve"] == "bn254" and c["n"] == 1 << 22)
    n = case["n"]
    print(f"zkmetal_bn254_msm n={n} srs_detect={detect}", flush=True)
    engine.srs_generate(fx["seed_bases"], n)
    bases = engine.srs_download(n)
    sc = zk.gen_scalars(fx["seed_scalars"], n, 0)
    engine.set_srs_detect(detect)
    try:
        hits0 = engine.srs_detect_hits()
        got = result_affine(engine.msm(bases, sc))
        assert (engine.srs_detect_hits() > hits0) == detect
    finally:
        engine.set_srs_detect(True)
    assert got == expected_point(case)
