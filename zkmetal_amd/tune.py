"""Autotuner: window width c per MSM size class and the launch schedule per
call size class, measured on the device.

Window tables, indexed by ceil(log2 n):
  * plain -- an MSM called alone: the c with the lowest wall time per call;
  * batch -- an MSM inside a multi-MSM call: the c with the lowest wall time
    per MSM when `batch_points` worth of MSMs of that class run as one call
    (latency terms amortise there, so the throughput-optimal c can differ).
Candidates are the cost-model default and its neighbours.

Schedule tables, indexed by ceil(log2 e) of a call's sorted entries e (the
engine's own size measure for a call; one table for calls of one MSM, one
for calls of several): the pipeline parts (how the call's (MSM, window)
segments are split so that the sort of one part runs on a second stream
beside the accumulation of the previous one) and the bucket reduction's
buckets per lane (the tile of buckets a lane sums before the lanes of a row
or column meet in a shuffle tree).  Each is swept on single MSMs of every
size and on batches of log-uniform sizes from 2^19 to 2^24 points, parts
first (at the built-in reduction tile), then the tile.

Bases are a resident seeded table, scalars seeded and device resident, so
only the MSM pipeline is timed.  A candidate replaces the built-in choice
only when faster by more than `margin`.  The result is a JSON-able dict;
`apply` loads it into an engine (a request server created on that engine
copies it).

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
         batch_points=1 << 21, margin=0.03, log=None, schedule=True, sched_batch_logn=(19, 24)):
    """Sweep c per size class, then (schedule=True) the launch schedule per
    call size class with the tuned widths in place; returns {"plain": {logn:
    c}, "batch": {...}, "ms": {...}, "schedule": {"plain": {cls: [parts,
    per]}, "batch": {...}}, "schedule_ms": {...}} and leaves the engine with
    the tuned tables applied.  A candidate replaces the default only when it
    is faster by more than `margin` (run-to-run noise is a few percent)."""
    own = engine is None
    eng = engine or zk.MetalMSM()
    nmax = 1 << max_logn
    eng.srs_generate(seed, nmax)
    sched_points = (1 << sched_batch_logn[1]) + nmax if schedule else 0
    scal = zk.gen_scalars(seed + 1, max(nmax, batch_points, sched_points), dist)
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
        eng.set_window_bits(0)
        if schedule:
            plain, batch = tables(out)
            eng.set_window_table(plain)
            eng.set_window_table(batch, batch=True)
            out["schedule"], out["schedule_ms"] = tune_schedule(eng, dsc, min_logn, max_logn, reps, margin,
                                                                seed + 2, log, sched_batch_logn)
    finally:
        eng.set_window_bits(0)
        dsc.free()
    apply(eng, out)
    if own:
        eng.close()
    return out


PARTS = [1, 2, 3, 4, 6, 8]
PER = [4, 8, 16, 32, 64]


def _sched_class(eng):
    """size class ceil(log2 entries) of the engine's last call"""
    e = int(eng.last_timing()["entries"])
    return max(0, (e - 1).bit_length())


def tune_schedule(eng, dsc, min_logn, max_logn, reps, margin, seed, log=None, batch_logn=(19, 24)):
    """Sweep pipeline parts, then the reduction tile, per call size class on
    single MSMs of 2^min_logn .. 2^max_logn points and on batches of
    log-uniform sizes (2^batch_logn[0] .. 2^batch_logn[1] points in all, MSMs
    of 2^min_logn points up to a quarter of the batch, at most 2^max_logn and
    2^20); dsc holds at least 2^batch_logn[1] + 2^max_logn scalars.  Returns
    ({"plain": {cls: [parts, per]}, "batch": {...}}, {"plain": {cls: ms
    table}, ...})."""
    rnd = np.random.default_rng(seed)
    work = [("plain", [1 << k]) for k in range(min_logn, max_logn + 1)]
    for lt in range(batch_logn[0], batch_logn[1] + 1):
        hi = max(min_logn, min(lt - 2, 20, max_logn))
        counts, tot = [], 0
        while tot < (1 << lt):
            c = int(2 ** rnd.uniform(min_logn, hi))
            counts.append(c)
            tot += c
        work.append(("batch", counts))
    sched = {"plain": {}, "batch": {}}
    times = {"plain": {}, "batch": {}}
    for kind, counts in work:
        batch = kind == "batch"
        empty = [0] * 33
        eng.set_schedule_table(empty, empty, batch=batch)
        fn = lambda: eng.msm_batch_resident(dsc, counts)  # noqa: E731
        fn()
        cls = _sched_class(eng)
        if cls in times[kind]:  # a size class is tuned on its first workload
            continue
        base_parts = int(eng.last_timing()["parts"])
        base_per = int(eng.last_timing()["buckets_per_group"])
        base = _time(fn, reps)
        res = {"builtin": round(base, 4)}

        def run(parts, per):
            pt, pr = [0] * 33, [0] * 33
            pt[cls], pr[cls] = parts, per
            eng.set_schedule_table(pt, pr, batch=batch)
            ms = _time(fn, reps)
            res[f"p{parts}/g{per}"] = round(ms, 4)
            return ms

        best_p, best = 0, base
        for p in PARTS:
            if p == base_parts:
                continue
            ms = run(p, 0)
            if ms < best * (1.0 - margin):
                best_p, best = p, ms
        best_g = 0
        for g in PER:
            if g == base_per:
                continue
            ms = run(best_p, g)
            if ms < best * (1.0 - margin):
                best_g, best = g, ms
        eng.set_schedule_table(empty, empty, batch=batch)
        if best_p or best_g:
            sched[kind][cls] = [best_p, best_g]
        times[kind][cls] = res
        if log:
            log(f"{kind} class 2^{cls} ({len(counts)} MSMs, {sum(counts)} points; built-in parts {base_parts}, "
                f"per {base_per}): {base:.3f} ms -> parts {best_p or base_parts}, per {best_g or base_per}: "
                f"{best:.3f} ms")
    return sched, times


def tables(result):
    """(plain, batch) lists of 33 entries, 0 where a class was not tuned."""
    t = [[0] * 33, [0] * 33]
    for i, key in enumerate(("plain", "batch")):
        for logn, c in result.get(key, {}).items():
            t[i][int(logn)] = int(c)
    return t


def schedule_tables(result):
    """{"plain"|"batch": (parts[33], per[33])}, 0 where a class keeps the
    built-in rule."""
    out = {}
    for key in ("plain", "batch"):
        pt, pr = [0] * 33, [0] * 33
        for cls, (p, g) in result.get("schedule", {}).get(key, {}).items():
            pt[int(cls)], pr[int(cls)] = int(p), int(g)
        out[key] = (pt, pr)
    return out


def apply(engine, result):
    plain, batch = tables(result)
    engine.set_window_table(plain)
    engine.set_window_table(batch, batch=True)
    if "schedule" in result:
        for key, (pt, pr) in schedule_tables(result).items():
            engine.set_schedule_table(pt, pr, batch=key == "batch")


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
    ap.add_argument("--no-schedule", action="store_true", help="window widths only")
    args = ap.parse_args()
    res = tune(min_logn=args.min_logn, max_logn=args.max_logn, reps=args.reps, spread=args.spread,
               dist=args.dist, margin=args.margin, log=lambda s: print(s, flush=True),
               schedule=not args.no_schedule)
    print(json.dumps({"plain": res["plain"], "batch": res["batch"], "schedule": res.get("schedule")}),
          flush=True)
    if args.out:
        save(res, args.out)
    np.set_printoptions(linewidth=200)
    p, b = tables(res)
    print("plain table:", p)
    print("batch table:", b)
    if "schedule" in res:
        for key, (pt, pr) in schedule_tables(res).items():
            print(f"{key} schedule parts:", pt)
            print(f"{key} schedule per:", pr)


if __name__ == "__main__":
    main()
