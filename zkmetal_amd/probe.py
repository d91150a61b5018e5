"""Profiling driver: a calibrated copy plus BN254 G1 MSMs, for rocprofv3.

    rocprofv3 --kernel-trace --stats -d OUT -- python3 -m zkmetal_amd.probe --logn 22
    rocprofv3 --pmc FETCH_SIZE -d OUT -- python3 -m zkmetal_amd.probe --logn 22 --reps 1

It launches k_copy over --copy-mib MiB (known traffic: that many bytes read
and written), then --reps plain MSMs of 2^logn points with bases and scalars
resident in device memory, and prints the per-stage HIP-event timings.
"""
import argparse
import json

import zkmetal_amd as zk


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--logn", type=int, default=22)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--window-bits", type=int, default=0)
    ap.add_argument("--dist", type=int, default=0)
    a = ap.parse_args()
    eng = zk.MetalMSM()
    copy_ms = eng.probe_copy(a.copy_mib << 20)
    print(json.dumps({"copy_bytes": a.copy_mib << 20, "copy_ms": copy_ms,
                      "copy_GBps_each_way": (a.copy_mib << 20) / (copy_ms * 1e-3) / 1e9}), flush=True)
    n = 1 << a.logn
    eng.srs_generate(0x5EED, n)
    bases = eng.srs_download(n)
    sc = zk.gen_scalars(0x5EEE, n, a.dist)
    dp = zk.DeviceBuffer(bases.nbytes)
    dp.upload(bases)
    ds = zk.DeviceBuffer(sc.nbytes)
    ds.upload(sc)
    if a.window_bits:
        eng.set_window_bits(a.window_bits)
    for r in range(a.reps):
        eng.msm_batch_device(dp, ds, [n])
        print(json.dumps({"rep": r, **{k: round(v, 4) for k, v in eng.last_timing().items()}}), flush=True)
    dp.free()
    ds.free()
    eng.close()


if __name__ == "__main__":
    main()
