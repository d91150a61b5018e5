#!/usr/bin/env python3
"""bench.py -- BN254 G1 multi-scalar multiplication on one MI355X.

Default workload (the configuration the headline metric is quoted on): a
plain BN254 G1 MSM of n = 2^22 points, scalars uniform in Fr, bases
k_i * G for seeded uniform k_i (affine, Montgomery), bases and scalars
resident in HBM when the timed region starts.  One step = one complete MSM
call (GLV/endomorphism table, decomposition, sort, accumulation, reduction,
window combine, normalisation) returning the result to the host.

Prints one JSON line.  `value` is ms per MSM at n = 2^22 (lower is better);
`roofline` is the dominant kernel (bucket accumulation) against the MI355X
32x32->64 integer multiply peak (256 CUs x 64 lanes x 2.4 GHz); `cpu_baseline` is the CPU
restatement of the reference's Pippenger on this host's cores, on the same
inputs, timed in the same run.

    python bench.py [--steps K] [--warmup W] [--logn 22] [--no-cpu]
    python bench.py --config batched|skewed|serve   (the batched, skewed and
                                                     serving workloads)
    python bench.py --config grumpkin              (the same MSM on Grumpkin G1)
    python bench.py --config g2                    (the same MSM on BN254 G2)
    python bench.py --profile [--config ...] --steps K  (profiling run for
                                                     rocprofv3: copy-kernel
                                                     calibration, then K calls)
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import zkmetal_amd as zk  # noqa: E402

PRODUCTS_PER_FIELD_MUL = 136  # 8x8 a*b + 8x8 m*p + 8 m_i (32-bit limbs)

# Peak 32x32->64 integer multiply rate of an MI355X: v_mad_u64_u32 issues one
# wave64 instruction per 4 clocks per SIMD (64 lanes per clock per CU; the
# probe's independent-chain microbenchmark sustains that rate at the clock it
# runs at), 256 CUs at the published 2400 MHz peak engine clock.
MI355X_CUS = 256
MI355X_PEAK_CLOCK_HZ = 2.4e9
PEAK_INT32_MUL = MI355X_CUS * 64 * MI355X_PEAK_CLOCK_HZ


def algorithmic_field_muls(n):
    """F(n, c*) for GLV + signed windows with XYZZ arithmetic: 10 muls per
    mixed addition over 2n half-scalar entries per window, 2 additions of 14
    muls per bucket in the running-sum reduction, one phi multiply per base;
    c* minimises F (so the count does not depend on this library's choice)."""
    best = None
    for c in range(4, 25):
        w = math.ceil(128 / c)
        f = 10 * (2 * n) * w + 28 * w * 2 ** (c - 1) + n
        best = f if best is None else min(best, f)
    return best


def algorithmic_field_muls_grumpkin(n):
    """The same count without GLV: n entries per window over ceil(254/c)
    windows of the scalar (< p < 2^254)."""
    best = None
    for c in range(4, 25):
        w = math.ceil(254 / c)
        f = 10 * n * w + 28 * w * 2 ** (c - 1)
        best = f if best is None else min(best, f)
    return best


def algorithmic_field_muls_g2(n):
    """BN254 G2 (no GLV): Fp multiplies with Karatsuba Fp2 arithmetic, 28 per
    mixed addition (8 Fp2 multiplies x 3 + 2 squarings x 2) over n entries
    per window, 2 additions of 40 per bucket, ceil(254/c) windows."""
    best = None
    for c in range(4, 25):
        w = math.ceil(254 / c)
        f = 28 * n * w + 80 * w * 2 ** (c - 1)
        best = f if best is None else min(best, f)
    return best


def timed_steps(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    samples = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        fn()
        samples.append((time.perf_counter() - s) * 1e3)
    total = (time.perf_counter() - t0) * 1e3
    return total / steps, samples


def cpu_baseline(bases_host, scalars_host, n, gpu_result, threads, curve="bn254"):
    from oracle import ref  # the oracle is only the checker / CPU baseline leg

    t = []
    res = None
    for _ in range(4):  # one warmup, then the median of three
        s = time.perf_counter()
        if curve == "g2":
            res = ref.g2_pippenger(bases_host, scalars_host, n, 0, threads)
        else:
            res = ref.pippenger(bases_host, scalars_host, n, 0, threads, curve=curve)
        t.append((time.perf_counter() - s) * 1e3)
    if curve == "g2":
        # the library's G2 result is Jacobian too: compare both in affine
        from oracle import bn254_g2
        match = bn254_g2.bytes_to_point(ref.g2_jacobian_to_affine_bytes(res)) == \
            bn254_g2.jacobian_bytes_to_affine(gpu_result)
        fn = "the BN254 G2 Pippenger (the reference's G2 point formulas, the G1 window scheme)"
    else:
        match = ref.jacobian_to_affine_bytes(res, curve=curve) == bytes(gpu_result[:64])
        fn = "bn254_pippenger_msm" if curve == "bn254" else "grumpkin_pippenger_msm"
    return {
        "value": round(statistics.median(t[1:]), 1),
        "unit": "ms per MSM",
        "cores": threads,
        "kind": "port",
        "sample": f"the full n=2^{int(math.log2(n))} MSM on the same inputs, 1 warmup + median of 3 runs; "
                  f"C restatement of {fn}, one thread per 16-bit window",
        "matches_gpu": bool(match),
        "host_cpu": _cpu_model(),
        "host_nproc": os.cpu_count(),
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def plain(args, eng, peak):
    n = 1 << args.logn
    eng.srs_generate(args.seed, n)
    bases = eng.srs_download(n)
    scal = zk.gen_scalars(args.seed + 1, n, 0)
    dpts = zk.DeviceBuffer(bases.nbytes)
    dpts.upload(bases)
    dsc = zk.DeviceBuffer(scal.nbytes)
    dsc.upload(scal)
    result = {}

    def step():
        result["r"] = eng.msm_batch_device(dpts, dsc, [n])[0]
        result.setdefault("t", []).append(eng.last_timing())

    ms, samples = timed_steps(step, args.steps, args.warmup)
    timings = result["t"][args.warmup:]
    acc_ms = statistics.mean(t["accumulate_ms"] for t in timings)
    valid = statistics.mean(t["valid_entries"] for t in timings)
    stage = {k: round(statistics.mean(t[k] for t in timings), 4)
             for k in ("decompose_ms", "sort_ms", "accumulate_ms", "merge_ms", "reduce_ms",
                       "copyout_ms", "device_ms", "host_tail_ms")}
    acc_products = valid * 10 * PRODUCTS_PER_FIELD_MUL
    achieved = acc_products / (acc_ms * 1e-3)
    msm_products = algorithmic_field_muls(n) * PRODUCTS_PER_FIELD_MUL
    extra = {
        "ms_p50": round(statistics.median(samples), 4),
        "msms_per_s": round(1e3 / ms, 2),
        "stage_ms": stage,
        "window_bits": int(timings[-1]["window_bits"]),
        "msm_roofline_frac": round(msm_products / (ms * 1e-3) / peak, 4),
        "msm_algorithmic_field_muls": algorithmic_field_muls(n),
        "measured_rates": {k: float("%.4g" % v) for k, v in args.rates.items()},
    }
    # the same call repeated under the library's default graph mode (auto:
    # the second call of a shape captures, later ones replay one hipGraph);
    # the headline above is eager, whose per-stage HIP events time the
    # accumulation for the roofline
    eng.set_graph_mode(eng.GRAPH_AUTO)
    ms_rep, _ = timed_steps(lambda: eng.msm_batch_device(dpts, dsc, [n]), args.steps, max(2, args.warmup))
    eng.set_graph_mode(eng.GRAPH_OFF)
    extra["ms_graph_replay"] = round(ms_rep, 4)
    # PCIe-inclusive host-buffer calls (zkmetal_bn254_msm): bases that are
    # the resident SRS (detected, no base upload) and the same call with the
    # detection off (bases uploaded, table rebuilt); not the headline value
    if not args.skip_host_calls:
        host_pts, host_sc = bases.tobytes(), scal.tobytes()
        hc = {}
        for label, on in (("srs_detected", True), ("bases_uploaded", False)):
            eng.set_srs_detect(on)
            hms, _ = timed_steps(lambda: eng.msm(host_pts, host_sc), 3, 1)
            hc[label] = round(hms, 3)
        eng.set_srs_detect(True)
        extra["host_buffer_call_ms"] = hc
    # the metric also names n = 2^20: same buffers, first 2^20 points
    if args.logn >= 20 and not args.skip_2_20:
        n20 = 1 << 20
        ms20, s20 = timed_steps(lambda: eng.msm_batch_device(dpts, dsc, [n20]), max(5, args.steps), 2)
        extra["ms_2_20"] = round(ms20, 4)
        extra["ms_2_20_p50"] = round(statistics.median(s20), 4)
        extra["msm_roofline_frac_2_20"] = round(
            algorithmic_field_muls(n20) * PRODUCTS_PER_FIELD_MUL / (ms20 * 1e-3) / peak, 4)
    line = {
        "metric": f"BN254 G1 MSM ms at n=2^{args.logn}",
        "value": round(ms, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"plain BN254 G1 MSM, n=2^{args.logn}, scalars uniform in Fr (seeded), "
                        f"bases k_i*G (seeded), affine Montgomery; bases+scalars resident in HBM",
            "n": n,
            "scalars": "uniform Fr",
            "window_bits": extra["window_bits"],
        },
        "roofline": {
            "bound": "valu_int32_mul",
            "kernel": "k_accumulate",
            "achieved": float("%.4g" % (achieved / 1e12)),
            "peak": float("%.4g" % (peak / 1e12)),
            "unit": "T int32 products/s",
            "frac": round(achieved / peak, 4),
            "traffic": _traffic_from_profiles("accumulate_n22_pmc.md") if args.logn == 22 else None,
            # traffic is not measured in this run: it is the figure recorded in
            # the named profile (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes)
            "traffic_source": "profiles/accumulate_n22_pmc.md" if args.logn == 22 else None,
        },
        "extra": extra,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(bases.tobytes(), scal.tobytes(), n, result["r"], args.cpu_threads)
    dpts.free()
    dsc.free()
    return line


def grumpkin(args, eng, peak):
    """Grumpkin G1 MSM (no GLV: 255-bit signed windows over n entries), bases
    k_i * G generated on the device, scalars uniform below r, both resident."""
    n = 1 << args.logn
    dpts = zk.DeviceBuffer(64 * n)
    eng.grumpkin_points_generate(args.seed, n, dpts)
    scal = zk.gen_scalars(args.seed + 1, n, 0)
    dsc = zk.DeviceBuffer(scal.nbytes)
    dsc.upload(scal)
    result = {}

    def step():
        result["r"] = eng.grumpkin_msm_batch_device(dpts, dsc, [n])[0]
        result.setdefault("t", []).append(eng.last_timing())

    ms, samples = timed_steps(step, args.steps, args.warmup)
    timings = result["t"][args.warmup:]
    acc_ms = statistics.mean(t["accumulate_ms"] for t in timings)
    valid = statistics.mean(t["valid_entries"] for t in timings)
    stage = {k: round(statistics.mean(t[k] for t in timings), 4)
             for k in ("decompose_ms", "sort_ms", "accumulate_ms", "merge_ms", "reduce_ms",
                       "copyout_ms", "device_ms", "host_tail_ms")}
    achieved = valid * 10 * PRODUCTS_PER_FIELD_MUL / (acc_ms * 1e-3)
    msm_products = algorithmic_field_muls_grumpkin(n) * PRODUCTS_PER_FIELD_MUL
    line = {
        "metric": f"Grumpkin G1 MSM ms at n=2^{args.logn}",
        "value": round(ms, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"plain Grumpkin G1 MSM, n=2^{args.logn}, scalars uniform below BN254 r (seeded), "
                        f"bases k_i*G (seeded, generated on the device), affine Fr Montgomery; "
                        f"bases+scalars resident in HBM",
            "n": n,
            "window_bits": int(timings[-1]["window_bits"]),
        },
        "roofline": {
            "bound": "valu_int32_mul",
            "kernel": "k_accumulate",
            "achieved": float("%.4g" % (achieved / 1e12)),
            "peak": float("%.4g" % (peak / 1e12)),
            "unit": "T int32 products/s",
            "frac": round(achieved / peak, 4),
            "traffic": _traffic_from_profiles("grumpkin_accumulate_n22_pmc.md") if args.logn == 22 else None,
            # traffic is not measured in this run: it is the figure recorded in
            # the named profile (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes)
            "traffic_source": "profiles/grumpkin_accumulate_n22_pmc.md" if args.logn == 22 else None,
        },
        "extra": {
            "ms_p50": round(statistics.median(samples), 4),
            "msms_per_s": round(1e3 / ms, 2),
            "stage_ms": stage,
            "msm_roofline_frac": round(msm_products / (ms * 1e-3) / peak, 4),
            "msm_algorithmic_field_muls": algorithmic_field_muls_grumpkin(n),
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(dpts.download(), scal.tobytes(), n, result["r"],
                                            args.cpu_threads, curve="grumpkin")
    dpts.free()
    dsc.free()
    return line


def g2(args, eng, peak):
    """BN254 G2 MSM (Fp2 coordinates, 255-bit signed windows over n entries),
    bases k_i * G2 generated on the device, scalars uniform in Fr, both
    resident.  Accumulation roofline: 28 Fp multiplies per entry."""
    n = 1 << args.logn
    dpts = zk.DeviceBuffer(128 * n)
    eng.g2_points_generate(args.seed, n, dpts)
    scal = zk.gen_scalars(args.seed + 1, n, 0)
    dsc = zk.DeviceBuffer(scal.nbytes)
    dsc.upload(scal)
    result = {}

    def step():
        result["r"] = eng.g2_msm_batch_device(dpts, dsc, [n])[0]
        result.setdefault("t", []).append(eng.last_timing())

    ms, samples = timed_steps(step, args.steps, args.warmup)
    timings = result["t"][args.warmup:]
    acc_ms = statistics.mean(t["accumulate_ms"] for t in timings)
    valid = statistics.mean(t["valid_entries"] for t in timings)
    stage = {k: round(statistics.mean(t[k] for t in timings), 4)
             for k in ("decompose_ms", "sort_ms", "accumulate_ms", "merge_ms", "reduce_ms",
                       "copyout_ms", "device_ms", "host_tail_ms")}
    achieved = valid * 28 * PRODUCTS_PER_FIELD_MUL / (acc_ms * 1e-3)
    msm_products = algorithmic_field_muls_g2(n) * PRODUCTS_PER_FIELD_MUL
    line = {
        "metric": f"BN254 G2 MSM ms at n=2^{args.logn}",
        "value": round(ms, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"plain BN254 G2 MSM, n=2^{args.logn}, scalars uniform in Fr (seeded), "
                        f"bases k_i*G2 (seeded, generated on the device), affine Fp2 Montgomery; "
                        f"bases+scalars resident in HBM",
            "n": n,
            "window_bits": int(timings[-1]["window_bits"]),
        },
        "roofline": {
            "bound": "valu_int32_mul",
            "kernel": "k_accumulate",
            "achieved": float("%.4g" % (achieved / 1e12)),
            "peak": float("%.4g" % (peak / 1e12)),
            "unit": "T int32 products/s",
            "frac": round(achieved / peak, 4),
            "traffic": _traffic_from_profiles("g2_accumulate_n22_pmc.md") if args.logn == 22 else None,
            # traffic is not measured in this run: it is the figure recorded in
            # the named profile (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes)
            "traffic_source": "profiles/g2_accumulate_n22_pmc.md" if args.logn == 22 else None,
        },
        "extra": {
            "ms_p50": round(statistics.median(samples), 4),
            "msms_per_s": round(1e3 / ms, 2),
            "stage_ms": stage,
            "msm_roofline_frac": round(msm_products / (ms * 1e-3) / peak, 4),
            "msm_algorithmic_field_muls": algorithmic_field_muls_g2(n),
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(dpts.download(128 * n), scal.tobytes(), n, result["r"],
                                            args.cpu_threads, curve="g2")
    dpts.free()
    dsc.free()
    return line


def batched(args, eng, peak, kind=None):
    """kind "batched": 64 MSMs sharing a 2^20 resident table, sizes log-uniform
    in [2^12, 2^20], uniform scalars; "skewed": 32 MSMs of 2^18 witness-like
    scalars.  Each step is one call holding every MSM, replayed from a
    captured hipGraph (graph mode; --no-graph: eager launches), against the
    same MSMs as back-to-back plain calls (eager, one MSM per call).  Returns
    the result line and what the closed-form check needs."""
    kind = kind or args.config
    rnd = np.random.default_rng(args.seed)
    n_tab = 1 << 20
    counts = [int(2 ** u) for u in rnd.uniform(12, 20, size=64)] if kind == "batched" \
        else [1 << 18] * 32
    dist = 0 if kind == "batched" else 1
    eng.srs_generate(args.seed, n_tab)
    scal = zk.gen_scalars(args.seed + 2, sum(counts), dist)
    dsc = zk.DeviceBuffer(scal.nbytes)
    dsc.upload(scal)
    # stage split from one eager call (a graph replay has one device span)
    eng.set_graph_mode(eng.GRAPH_OFF)
    eager_res = eng.msm_batch_resident(dsc, counts)
    tb = eng.last_timing()
    stage = {k: round(tb[k], 4) for k in ("decompose_ms", "sort_ms", "accumulate_ms", "merge_ms",
                                           "reduce_ms", "device_tail_ms", "copyout_ms", "device_ms",
                                           "host_tail_ms")}
    graph = not args.no_graph
    eng.set_graph_mode(eng.GRAPH_ON if graph else eng.GRAPH_OFF)
    caps0 = eng.graph_captures()
    res = {}

    def step():
        res["r"] = eng.msm_batch_resident(dsc, counts)

    ms_b, samples = timed_steps(step, args.steps, args.warmup)
    batched_res = res["r"]
    captures = eng.graph_captures() - caps0
    device_ms = eng.last_timing()["device_ms"]
    eng.set_graph_mode(eng.GRAPH_OFF)
    # back-to-back plain launches of the same MSMs on the same device
    offs = np.concatenate([[0], np.cumsum(counts)])
    views = []
    for j, c in enumerate(counts):
        b = zk.DeviceBuffer(c * 32)
        b.upload(scal[offs[j]:offs[j] + c])
        views.append(b)

    def seq():
        return [eng.msm_batch_resident(views[j], [c])[0] for j, c in enumerate(counts)]

    ms_s, _ = timed_steps(seq, max(2, args.steps // 2), 1)
    seq_res = seq()
    for v in views:
        v.free()
    dsc.free()
    same = seq_res == batched_res and eager_res == batched_res
    line = {
        "metric": f"{'batched' if kind == 'batched' else 'skewed batched'} BN254 G1 MSMs/s",
        "value": round(len(counts) / (ms_b * 1e-3), 2),
        "unit": "MSMs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_b, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"{len(counts)} MSMs over a resident 2^20 table, sizes {min(counts)}..{max(counts)}, "
                               f"scalars {'uniform Fr' if dist == 0 else 'witness-like mix'}, "
                               f"{'one hipGraph replay' if graph else 'one eager launch sequence'} per call",
                   "msms": len(counts), "points": int(sum(counts)), "graph": graph},
        "extra": {"ms_p50": round(statistics.median(samples), 4),
                  "device_ms": round(device_ms, 4), "graph_captures": captures,
                  "stage_ms_eager": stage, "entries_per_lane": tb["entries_per_lane"],
                  "back_to_back_ms": round(ms_s, 3),
                  "back_to_back_msms_per_s": round(len(counts) / (ms_s * 1e-3), 2),
                  "speedup_vs_back_to_back": round(ms_s / ms_b, 3),
                  "results_identical": bool(same)},
    }
    check = {"srs_seed": args.seed, "srs_n": n_tab, "counts": counts, "scalars": scal,
             "results": batched_res}
    return line, check


def serve(args, eng):
    """Open-loop serving: a seeded Poisson stream of commitment requests
    (mean `--rate` per second, sizes log-uniform in [2^10, 2^22], witness-like
    scalar mix) against a resident 2^23-point SRS.  Scalars sit in page-locked
    host memory (a pool the requests slice), as a prover's witness would.
    Latency = submit to result on the host; p50/p99 over all requests."""
    import threading

    rnd = np.random.default_rng(args.seed)
    n_srs = 1 << 23
    nreq = args.requests
    sizes = [int(2 ** u) for u in rnd.uniform(10, 22, size=nreq)]
    gaps = rnd.exponential(1.0 / args.rate, size=nreq)
    arrivals = np.cumsum(gaps) - gaps[0]
    offs = [int(rnd.integers(0, n_srs - n + 1)) for n in sizes]
    eng.srs_generate(args.seed, n_srs)
    pool = zk.HostBuffer(n_srs * 32)
    pool.array[:] = zk.gen_scalars(args.seed + 3, n_srs, 1).reshape(-1)
    pool2d = pool.array.reshape(n_srs, 32)
    views = [pool2d[o:o + n] for o, n in zip(offs, sizes)]
    srv = zk.Server(eng, slots=args.slots, max_batch=args.max_batch,
                    hbm_budget_bytes=args.hbm_budget_gib << 30)
    # warm the slots' workspaces (not part of the stream)
    for t in [srv.submit(views[i], copy=False) for i in range(min(8, nreq))]:
        srv.wait(t)
    srv.close()
    srv = zk.Server(eng, slots=args.slots, max_batch=args.max_batch,
                    hbm_budget_bytes=args.hbm_budget_gib << 30)
    tickets = [None] * nreq

    def client():
        t0 = time.perf_counter()
        for i in range(nreq):
            d = t0 + arrivals[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(views[i], copy=False)

    th = threading.Thread(target=client)
    wall0 = time.perf_counter()
    th.start()
    results = []
    for i in range(nreq):
        while tickets[i] is None:
            time.sleep(0.0002)
        results.append(srv.wait(tickets[i]))
    th.join()
    wall = time.perf_counter() - wall0
    st = srv.stats()
    srv.close()
    # saturation: the first `burst` requests of the same stream submitted at
    # once (the server never idles), against the same MSMs as back-to-back
    # plain calls on one engine
    burst = min(args.burst, nreq)
    srv = zk.Server(eng, slots=args.slots, max_batch=args.max_batch,
                    hbm_budget_bytes=args.hbm_budget_gib << 30)
    t0 = time.perf_counter()
    tk = [srv.submit(views[i], copy=False) for i in range(burst)]
    burst_res = [srv.wait(t)[0] for t in tk]
    burst_s = time.perf_counter() - t0
    bst = srv.stats()
    srv.close()
    t0 = time.perf_counter()
    seq_res = [eng.msm_batch_resident(views[i], [sizes[i]])[0] for i in range(burst)]
    seq_s = time.perf_counter() - t0
    burst_same = burst_res == seq_res
    # bit-exact: a sample of requests recomputed alone equals the served result
    check = sorted(rnd.choice(nreq, size=min(12, nreq), replace=False).tolist())
    same = all(eng.msm_batch_resident(views[i], [sizes[i]])[0] == results[i][0] for i in check)
    closed = {"srs_seed": args.seed, "srs_n": n_srs, "pool_seed": args.seed + 3,
              "requests": [(offs[i], sizes[i], results[i][0]) for i in check]}
    lat = sorted(r[1] for r in results)
    big = [r[1] for r, n in zip(results, sizes) if n >= (1 << 20)]
    del views, pool2d
    pool.free()
    return {
        "metric": f"served BN254 G1 MSMs/s at {args.rate:g} req/s offered (p50/p99 latency)",
        "value": round(st["completed"] / wall, 2),
        "unit": "MSMs/s",
        "n_gpus": 1,
        "steps": nreq,
        "warmup": 8,
        "ms_per_step": round(wall * 1e3 / nreq, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"Poisson stream, mean {args.rate:g} req/s, {nreq} requests, sizes log-uniform "
                               f"in [2^10, 2^22], witness-like scalars (40% 0, 30% <2^16, 10% 1 or r-1, "
                               f"20% Fr), resident 2^23-point SRS, scalars in pinned host memory",
                   "slots": args.slots, "max_batch": args.max_batch,
                   "hbm_budget_gib": args.hbm_budget_gib, "points": int(sum(sizes))},
        "extra": {"p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                  "mean_ms": round(st["mean_ms"], 3), "max_ms": round(st["max_ms"], 3),
                  "p50_ms_host": round(lat[len(lat) // 2], 3),
                  "p99_ms_n_ge_2_20": round(sorted(big)[max(0, math.ceil(0.99 * len(big)) - 1)], 3) if big else None,
                  "offered_req_per_s": round(nreq / max(arrivals[-1], 1e-9), 2),
                  "batches": st["batches"], "mean_batch": round(st["mean_batch"], 3),
                  "hbm_peak_gib": round(st["hbm_peak_bytes"] / 2**30, 3),
                  "points_per_s": round(sum(sizes) / wall, 1),
                  "results_identical_sample": bool(same), "checked": len(check),
                  "saturation": {"requests": burst, "served_msms_per_s": round(burst / burst_s, 2),
                                 "back_to_back_msms_per_s": round(burst / seq_s, 2),
                                 "speedup_vs_back_to_back": round(seq_s / burst_s, 3),
                                 "mean_batch": round(bst["mean_batch"], 3),
                                 "p50_ms": round(bst["p50_ms"], 3), "p99_ms": round(bst["p99_ms"], 3),
                                 "results_identical": bool(burst_same)}},
    }, closed


def profile(args, eng):
    """Profiling run for rocprofv3 (kernel traces, PMC counters): the
    workload's setup and --warmup untimed calls, then one plain copy kernel
    over --copy-mib MiB (known traffic: that many bytes read and written, the
    HBM counters' calibration; it also marks where the measured calls start
    in a trace), then --steps calls of the workload with their stage timings,
    one JSON line each.  --config plain, grumpkin or g2: plain MSMs of 2^logn
    points, bases and scalars resident; batched / skewed: the 64-MSM and the
    32 x 2^18 witness-like calls over a resident 2^20 table."""
    n = 1 << args.logn
    bufs = []
    if args.config in ("batched", "skewed"):
        rnd = np.random.default_rng(args.seed)
        counts = [int(2 ** u) for u in rnd.uniform(12, 20, size=64)] if args.config == "batched" \
            else [1 << 18] * 32
        eng.srs_generate(args.seed, 1 << 20)
        sc = zk.gen_scalars(args.seed + 2, sum(counts), 0 if args.config == "batched" else 1)
        ds = zk.DeviceBuffer(sc.nbytes)
        ds.upload(sc)
        bufs.append(ds)

        def run():
            eng.msm_batch_resident(ds, counts)
    else:
        if args.config == "grumpkin":
            dp = zk.DeviceBuffer(64 * n)
            eng.grumpkin_points_generate(args.seed, n, dp)
            fn = eng.grumpkin_msm_batch_device
        elif args.config == "g2":
            dp = zk.DeviceBuffer(128 * n)
            eng.g2_points_generate(args.seed, n, dp)
            fn = eng.g2_msm_batch_device
        else:
            eng.srs_generate(args.seed, n)
            bases = eng.srs_download(n)
            dp = zk.DeviceBuffer(bases.nbytes)
            dp.upload(bases)
            fn = eng.msm_batch_device
        sc = zk.gen_scalars(args.seed + 1, n, 0)
        ds = zk.DeviceBuffer(sc.nbytes)
        ds.upload(sc)
        bufs += [dp, ds]

        def run():
            fn(dp, ds, [n])
    for _ in range(args.warmup):
        run()
    copy_ms = eng.probe_copy(args.copy_mib << 20)
    print(json.dumps({"copy_bytes": args.copy_mib << 20, "copy_ms": round(copy_ms, 4)}), flush=True)
    for r in range(args.steps):
        run()
        print(json.dumps({"step": r, **{k: round(v, 4) for k, v in eng.last_timing().items()}}), flush=True)
    for b in bufs:
        b.free()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--logn", type=int, default=22)
    ap.add_argument("--seed", type=int, default=0x5EED)
    ap.add_argument("--config", default="plain", choices=["plain", "batched", "skewed", "serve", "grumpkin", "g2"])
    ap.add_argument("--requests", type=int, default=1000, help="serve: requests in the stream")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: mean arrivals per second")
    ap.add_argument("--slots", type=int, default=2, help="serve: engine slots")
    ap.add_argument("--max-batch", type=int, default=64, help="serve: requests per batch")
    ap.add_argument("--hbm-budget-gib", type=int, default=16, help="serve: admission budget")
    ap.add_argument("--burst", type=int, default=300, help="serve: requests in the saturation phase")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-2-20", action="store_true")
    ap.add_argument("--skip-host-calls", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--profile", action="store_true",
                    help="profiling run (rocprofv3): a calibrated copy kernel, then --steps calls of "
                         "the workload with their stage timings; no JSON result line")
    ap.add_argument("--copy-mib", type=int, default=256, help="profile: copy kernel size")
    ap.add_argument("--no-legs", action="store_true",
                    help="plain: skip the batched / skewed / served legs")
    ap.add_argument("--no-graph", action="store_true", help="batched calls as eager launches")
    ap.add_argument("--leg-steps", type=int, default=5, help="timed calls of the batched legs")
    ap.add_argument("--leg-requests", type=int, default=300, help="requests of the served leg")
    ap.add_argument("--leg-burst", type=int, default=200, help="saturation requests of the served leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported in this build "
              f"(a multi-GPU MSM would shard points across GPUs and gather one point per GPU)",
              file=sys.stderr)
        sys.exit(2)
    eng = zk.MetalMSM()
    # every call eagerly, so that HIP events time each kernel stage; the
    # batched legs turn graph replays on for their timed steps themselves
    eng.set_graph_mode(eng.GRAPH_OFF)
    if args.profile:
        profile(args, eng)
        eng.close()
        return
    args.rates = eng.probe_rates()
    peak = PEAK_INT32_MUL
    checks = []
    if args.config == "plain":
        line = plain(args, eng, peak)
        if not args.no_legs:
            # the three system workloads, measured in the same run: batched and
            # skewed calls replayed from captured hipGraphs, and a served stream
            legs = {}
            la = argparse.Namespace(**vars(args))
            la.steps, la.warmup = args.leg_steps, 2
            for kind in ("batched", "skewed"):
                leg, chk = batched(la, eng, peak, kind)
                legs[kind] = _leg_summary(leg)
                checks.append((kind, chk))
            la.requests, la.burst = args.leg_requests, args.leg_burst
            leg, chk = serve(la, eng)
            legs["serve"] = _leg_summary(leg)
            checks.append(("serve", chk))
            line["extra"]["system_legs"] = legs
    elif args.config == "serve":
        line, chk = serve(args, eng)
        checks.append(("serve", chk))
    elif args.config == "grumpkin":
        line = grumpkin(args, eng, peak)
    elif args.config == "g2":
        line = g2(args, eng, peak)
    else:
        line, chk = batched(args, eng, peak)
        checks.append((args.config, chk))
    eng.close()
    if checks and not args.no_cpu:
        # part of the CPU leg: the system workloads' results against the closed
        # form, after every timed region
        line.setdefault("cpu_baseline", {})["closed_form_checks"] = closed_form_checks(checks)
    print(json.dumps(line), flush=True)


def _leg_summary(leg):
    keep = ("metric", "value", "unit", "steps", "warmup", "ms_per_step", "config", "extra")
    return {k: leg[k] for k in keep if k in leg}


def closed_form_checks(checks):
    """Checker (CPU leg, after the timed regions): each checked result of the
    system workloads against (sum s_i k_i mod r) * G of the resident bases
    k_i * G, with the oracle's exact dot product and scalar multiplication."""
    from oracle import bn254 as B
    from oracle import ref

    out = {}
    t0 = time.perf_counter()
    for kind, chk in checks:
        ks = zk.gen_scalars(chk["srs_seed"], chk["srs_n"], 0).tobytes()
        if kind == "serve":
            pool = zk.gen_scalars(chk["pool_seed"], chk["srs_n"], 1).tobytes()
            items = [(pool[32 * o:32 * (o + n)], n, r) for o, n, r in chk["requests"]]
        else:
            sb = chk["scalars"].tobytes()
            offs = np.concatenate([[0], np.cumsum(chk["counts"])])
            items = [(sb[32 * int(offs[j]):32 * int(offs[j + 1])], n, chk["results"][j])
                     for j, n in enumerate(chk["counts"])]
        ok = 0
        for sc, n, r in items:
            want = B.mul(B.G, ref.scalar_dot(sc, ks[:32 * n], n))
            ok += B.jacobian_bytes_to_affine(r[:32], r[32:64], r[64:96]) == want
        out[kind] = {"checked": len(items), "equal_closed_form": int(ok)}
    out["seconds"] = round(time.perf_counter() - t0, 2)
    return out


def _traffic_from_profiles(name):
    """HBM bytes per k_accumulate launch at n = 2^22 measured with rocprofv3
    --pmc (FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel calibration),
    recorded in profiles/<name> as a line 'traffic_bytes_per_launch: N'."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
