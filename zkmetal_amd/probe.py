"""Profiling driver: a calibrated copy plus BN254 G1 MSMs, for rocprofv3.

    rocprofv3 --kernel-trace --stats -d OUT -- python3 -m zkmetal_amd.probe --logn 22
    rocprofv3 --pmc FETCH_SIZE -d OUT -- python3 -m zkmetal_amd.probe --logn 22 --reps 1

It launches k_copy over --copy-mib MiB (known traffic: that many bytes read
and written), then --reps plain MSMs of 2^logn points with bases and scalars
resident in device memory, and prints the per-stage HIP-event timings.
With --batched: --reps calls of the batched workload instead (64 MSMs over a
resident 2^20 table, sizes log-uniform in [2^12, 2^20], one call each).
With --curve grumpkin or g2: the same plain MSMs on Grumpkin G1 or BN254 G2,
bases generated on the device (k_i * G, k_i * G2).
"""
import argparse
import json

import numpy as np

import zkmetal_amd as zk


def batched(eng, a):
    rnd = np.random.default_rng(0x5EED)
    counts = [int(2 ** u) for u in rnd.uniform(12, 20, size=64)]
    eng.srs_generate(0x5EED, 1 << 20)
    sc = zk.gen_scalars(0x5EED + 2, sum(counts), a.dist)
    ds = zk.DeviceBuffer(sc.nbytes)
    ds.upload(sc)
    for r in range(a.reps):
        eng.msm_batch_resident(ds, counts)
        print(json.dumps({"rep": r, **{k: round(v, 4) for k, v in eng.last_timing().items()}}), flush=True)
    ds.free()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--logn", type=int, default=22)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--window-bits", type=int, default=0)
    ap.add_argument("--dist", type=int, default=0)
    ap.add_argument("--batched", action="store_true")
    ap.add_argument("--curve", default="bn254", choices=["bn254", "grumpkin", "g2"])
    a = ap.parse_args()
    eng = zk.MetalMSM()
    copy_ms = eng.probe_copy(a.copy_mib << 20)
    print(json.dumps({"copy_bytes": a.copy_mib << 20, "copy_ms": copy_ms,
                      "copy_GBps_each_way": (a.copy_mib << 20) / (copy_ms * 1e-3) / 1e9}), flush=True)
    if a.batched:
        batched(eng, a)
        eng.close()
        return
    n = 1 << a.logn
    if a.curve == "bn254":
        eng.srs_generate(0x5EED, n)
        bases = eng.srs_download(n)
        dp = zk.DeviceBuffer(bases.nbytes)
        dp.upload(bases)
        run = eng.msm_batch_device
    elif a.curve == "grumpkin":
        dp = zk.DeviceBuffer(64 * n)
        eng.grumpkin_points_generate(0x5EED, n, dp)
        run = eng.grumpkin_msm_batch_device
    else:
        dp = zk.DeviceBuffer(128 * n)
        eng.g2_points_generate(0x5EED, n, dp)
        run = eng.g2_msm_batch_device
    sc = zk.gen_scalars(0x5EEE, n, a.dist)
    ds = zk.DeviceBuffer(sc.nbytes)
    ds.upload(sc)
    if a.window_bits:
        eng.set_window_bits(a.window_bits)
    for r in range(a.reps):
        run(dp, ds, [n])
        print(json.dumps({"rep": r, **{k: round(v, 4) for k, v in eng.last_timing().items()}}), flush=True)
    dp.free()
    ds.free()
    eng.close()


if __name__ == "__main__":
    main()
