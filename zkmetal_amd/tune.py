"""Autotuner: window width c per size class, measured on the device.

Two tables, both indexed by ceil(log2 n):
  * plain -- an MSM called alone: the c with the lowest wall time per call;
  * batch -- an MSM inside a multi-MSM call: the c with the lowest wall time
    per MSM when `batch_points` worth of MSMs of that class run as one call
    (latency terms amortise there, so the throughput-optimal c can differ).
Candidates are the cost-model default and its neighbours; bases are a
resident seeded table, scalars seeded and device resident, so only the MSM
pipeline is timed.  The result is a JSON-able dict; `apply` loads it into an
engine (a request server created on that engine copies it).

    python -m zkmetal_amd.tune --out tuned.json [--min-logn 10 --max-logn 22]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import zkmetal_amd as zk  # noqa: E402

C_MIN, C_MAX = 4, 16


def _candidates(n, spread):
    c0 = zk.default_window_bits(n)
    return [c for c in range(c0 - spread, c0 + spread + 1) if C_MIN <= c <= C_MAX]


def _time(fn, reps):
    fn()  # warm (workspace growth, caches)
    best = None
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        dt = (time.perf_counter() - t) * 1e3
        best = dt if best is None else min(best, dt)
    return best


def _pick(res, which, default, margin):
    """fastest candidate, but the cost-model default unless beaten by > margin"""
    best = min(res, key=lambda c: res[c][which])
    if default in res and res[best][which] > res[default][which] * (1.0 - margin):
        return default
    return best


def tune(engine=None, min_logn=10, max_logn=22, reps=3, spread=2, dist=0, seed=0x7E57,
         batch_points=1 << 21, margin=0.03, log=None):
    """Sweep c per size class; returns {"plain": {logn: c}, "batch": {...},
    "ms": {...}} and leaves the engine with the tuned tables applied.  A
    candidate replaces the cost-model default only when it is faster by more
    than `margin` (run-to-run noise is a few percent)."""
    own = engine is None
    eng = engine or zk.MetalMSM()
    nmax = 1 << max_logn
    eng.srs_generate(seed, nmax)
    scal = zk.gen_scalars(seed + 1, max(nmax, batch_points), dist)
    dsc = zk.DeviceBuffer(scal.nbytes)
    dsc.upload(scal)
    out = {"arch": "gfx950", "dist": dist, "plain": {}, "batch": {}, "ms": {}}
    try:
        for logn in range(min_logn, max_logn + 1):
            n = 1 << logn
            k = max(1, batch_points // n) if n < batch_points else 1
            res = {}
            for c in _candidates(n, spread):
                eng.set_window_bits(c)
                alone = _time(lambda: eng.msm_batch_resident(dsc, [n]), reps)
                many = _time(lambda: eng.msm_batch_resident(dsc, [n] * k), reps) / k if k > 1 else alone
                res[c] = (alone, many)
                if log:
                    log(f"logn {logn} c {c}: alone {alone:.3f} ms, in batch of {k} {many:.3f} ms/MSM")
            c0 = zk.default_window_bits(n)
            out["plain"][logn] = _pick(res, 0, c0, margin)
            out["batch"][logn] = _pick(res, 1, c0, margin)
            out["ms"][logn] = {str(c): [round(a, 4), round(b, 4)] for c, (a, b) in res.items()}
    finally:
        eng.set_window_bits(0)
        dsc.free()
    apply(eng, out)
    if own:
        eng.close()
    return out


def tables(result):
    """(plain, batch) lists of 33 entries, 0 where a class was not tuned."""
    t = [[0] * 33, [0] * 33]
    for i, key in enumerate(("plain", "batch")):
        for logn, c in result.get(key, {}).items():
            t[i][int(logn)] = int(c)
    return t


def apply(engine, result):
    plain, batch = tables(result)
    engine.set_window_table(plain)
    engine.set_window_table(batch, batch=True)


def save(result, path):
    with open(path, "w") as fh:
        json.dump(result, fh, indent=1, sort_keys=True)


def load(path):
    with open(path) as fh:
        return json.load(fh)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--min-logn", type=int, default=10)
    ap.add_argument("--max-logn", type=int, default=22)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--spread", type=int, default=2)
    ap.add_argument("--dist", type=int, default=0)
    ap.add_argument("--margin", type=float, default=0.03)
    args = ap.parse_args()
    res = tune(min_logn=args.min_logn, max_logn=args.max_logn, reps=args.reps, spread=args.spread,
               dist=args.dist, margin=args.margin, log=lambda s: print(s, flush=True))
    print(json.dumps({"plain": res["plain"], "batch": res["batch"]}), flush=True)
    if args.out:
        save(res, args.out)
    np.set_printoptions(linewidth=200)
    p, b = tables(res)
    print("plain table:", p)
    print("batch table:", b)


if __name__ == "__main__":
    main()
