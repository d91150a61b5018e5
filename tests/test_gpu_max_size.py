"""The largest size the configurations use: a 2^23-point resident SRS (the
serving workload's table) and MSMs over all of it, alone and in a ragged
batch with a one-point MSM and an odd size, against the closed form
(sum s_i k_i mod r) * G with the C oracle's exact dot product."""
import pytest

import zkmetal_amd as zk
from oracle import bn254 as B
from oracle import ref
from helpers import result_affine

pytestmark = pytest.mark.gpu

N = 1 << 23


def test_msm_over_a_2_23_resident_srs_alone_and_in_a_ragged_batch():
    eng = zk.MetalMSM()
    counts = [N, 1, (1 << 22) + 3]
    total = sum(counts)
    ds = zk.DeviceBuffer(32 * total)
    try:
        print(f"srs 2^23, counts {counts}", flush=True)
        eng.srs_generate(0x23A, N)
        sc = zk.gen_scalars(0x23B, total, 0)
        ds.upload(sc)
        ks = zk.gen_scalars(0x23A, N, 0).tobytes()
        sb = sc.tobytes()
        want, off = [], 0
        for n in counts:
            want.append(B.mul(B.G, ref.scalar_dot(sb[32 * off:32 * (off + n)], ks[:32 * n], n)))
            off += n
        got = [result_affine(r) for r in eng.msm_batch_resident(ds, counts)]
        assert got == want
        alone = eng.msm_batch_resident(ds, [N])[0]
        assert result_affine(alone) == want[0]
    finally:
        ds.free()
        eng.close()
