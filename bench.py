#!/usr/bin/env python3
"""Throughput benchmark: large single transposes on one MI355X.

Workload (named large cases): seven ~256 MiB transposes, float32
and float64, covering the Tiled, TiledCopy and Packed/PackedSplit paths:

    f32 (8192,8192)/(1,0)           f32 (1024,256,256)/(0,2,1)
    f32 (2,8,16,32,64,128)/(5,3,1,0,2,4)
    f64 (4096,8192)/(1,0)           f64 (512,256,256)/(0,2,1)
    f64 (4,4,8,32,32,256)/(2,5,0,4,1,3)
    f64 (5,97,211,331)/(3,1,0,2)

A step is one pass over the seven cases with inputs resident in HBM. The
metric is 2*bytes/t summed over the step (the reference's librettTimer count,
src/Timer.cpp:151). The line also carries a same-run hipMemcpyAsync D2D
ceiling, the roofline of the dominant kernel (HIP events on the plan's
stream) and a CPU baseline (the OpenMP host permutation on a bounded sample).

Unless --quick is given, the same line then carries the other workloads,
each with its outputs checked on the device: "batched" (4096 uneven
problems in one persistent graph launch vs a per-call librettExecute loop),
"serving" (Poisson arrivals through the continuous-batching engine: warm and
cold plan caches, p50/p99 over a load sweep, per-call baselines),
"large_random" (48 seeded ~256 MiB cases, librettPlan vs librettPlanMeasure)
and "selector" (held-out slowdown of the plan choice vs exhaustive timing).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CASES = [
    ((8192, 8192), (1, 0), 4),
    ((1024, 256, 256), (0, 2, 1), 4),
    ((2, 8, 16, 32, 64, 128), (5, 3, 1, 0, 2, 4), 4),
    ((4096, 8192), (1, 0), 8),
    ((512, 256, 256), (0, 2, 1), 8),
    ((4, 4, 8, 32, 32, 256), (2, 5, 0, 4, 1, 3), 8),
    ((5, 97, 211, 331), (3, 1, 0, 2), 8),
]
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--measure", action="store_true", help="choose plans with librettPlanMeasure")
    ap.add_argument("--analytic", action="store_true",
                    help="librettPlan with the analytic model only (no learned selector)")
    ap.add_argument("--exact-counts", action="store_true",
                    help="LIBRETT_EXACT_COUNTS=1: librettPlan refines the model's sampled line counts with the "
                         "device counters")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--workload", default="large",
                    choices=["large", "large-random", "batched", "serving", "selector"],
                    help="large: the seven ~256 MiB cases (default); large-random: 48 seeded ~256 MiB cases "
                         "(keep/move/small leading dim x f32/f64 x 8 seeds), librettPlan vs librettPlanMeasure, "
                         "checked on the device; batched: 4096 uneven problems; serving: Poisson request stream "
                         "through the continuous-batching engine; selector: learned plan selector vs analytic "
                         "model on held-out problems")
    ap.add_argument("--heldout", type=int, default=2000, help="selector workload: held-out problems")
    ap.add_argument("--selector-problems", type=int, default=200,
                    help="held-out problems of the selector part of the default run")
    ap.add_argument("--quick", action="store_true",
                    help="default run: only the seven named cases (no batched / serving / large-random / selector "
                         "parts)")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "RUN_LOG"),
                    help="summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of "
                         "'bench.py --steps 1 --warmup 0 --no-cpu' (RUN_LOG: that run's stderr, for the plans) "
                         "into HBM traffic per named case; no GPU needed")
    ap.add_argument("--pmc-json", default=None, help="with --pmc-summary: also write the summary as JSON here")
    ap.add_argument("--pmc-requests", nargs=2, metavar=("READ_DIR", "WRITE_DIR"),
                    help="summarise two rocprofv3 --pmc runs of the raw TCC/EA request counters (reads: "
                         "TCC_EA0_RDREQ_sum TCC_BUBBLE_sum TCC_EA0_RDREQ_32B_sum TCC_EA0_RDREQ_DRAM_sum; writes: "
                         "TCC_EA0_WRREQ_sum TCC_EA0_WRREQ_64B_sum TCC_EA0_WRREQ_DRAM_sum) of "
                         "'bench.py --steps 1 --warmup 0 --no-cpu' and exit")
    ap.add_argument("--requests", type=int, default=20000, help="serving workload: arrivals per rate")
    ap.add_argument("--arena-gb", type=float, default=64.0, help="serving workload: device arena budget (GiB)")
    ap.add_argument("--slots", type=int, default=8, help="serving workload: micro-batches in flight")
    ap.add_argument("--max-batch", type=int, default=64, help="serving workload: requests per micro-batch")
    ap.add_argument("--clients", type=int, default=1, help="serving workload: client threads submitting requests")
    ap.add_argument("--problems", type=int, default=4096, help="batched workload: number of problems")
    ap.add_argument("--tune", type=int, default=3, help="batched workload: autotuner candidates per problem")
    args = ap.parse_args()
    if args.pmc_summary:
        return pmc_summary(*args.pmc_summary, json_path=args.pmc_json)
    if args.pmc_requests:
        return pmc_requests(*args.pmc_requests)
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs on a single MI355X and has no "
              "multi-GPU path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.exact_counts:
        os.environ["LIBRETT_EXACT_COUNTS"] = "1"
    import librett_amd as L
    from librett_amd import data, hip

    hip.check(hip.runtime().hipSetDevice(0), "hipSetDevice")
    stream = hip.stream_create()
    L.librettInitialize()  # one-time set-up, as a LibreTT caller would do first
    if args.analytic:
        L.selector_set(None, None, None)
    sub = {"batched": bench_batched, "serving": bench_serving, "selector": bench_selector,
           "large-random": bench_large_random}
    if args.workload in sub:
        print(json.dumps(sub[args.workload](args, stream)), flush=True)
        return
    line = bench_named(args, stream)
    if not args.quick:
        # The other workloads, each checked on the device, in the same line.
        t0 = time.perf_counter()
        line["batched"] = summarize_batched(bench_batched(args, stream))
        log("batched workload done (%.0f s)" % (time.perf_counter() - t0))
        t0 = time.perf_counter()
        line["serving"] = summarize_serving(bench_serving(args, stream))
        log("serving workload done (%.0f s)" % (time.perf_counter() - t0))
        t0 = time.perf_counter()
        line["large_random"] = summarize_large_random(bench_large_random(args, stream))
        log("large-random workload done (%.0f s)" % (time.perf_counter() - t0))
        t0 = time.perf_counter()
        args.heldout = args.selector_problems
        line["selector"] = summarize_selector(bench_selector(args, stream))
        log("selector workload done (%.0f s)" % (time.perf_counter() - t0))
    hip.stream_destroy(stream)
    print(json.dumps(line), flush=True)


def bench_named(args, stream):
    """The seven named ~256 MiB cases: the headline line."""
    import librett_amd as L
    from librett_amd import data, hip
    selection = ("librettPlanMeasure" if args.measure else
                 "learned selector (librettPlan)" if L.selector_generation() else "gfx950 analytic model (librettPlan)")
    if args.exact_counts and not args.measure:
        selection += " with exact line counts"

    # One random source buffer, uploaded as the input of every case.
    max_bytes = max(int(np.prod(d)) * e for d, _, e in CASES)
    src = data.uniform(max_bytes // 8 + 1, 8, 20261202).view(np.uint8)[:max_bytes]
    cases = []
    for dim, perm, elem in CASES:
        vol = int(np.prod(dim))
        nbytes = vol * elem
        a = hip.DeviceBuffer(nbytes)
        b = hip.DeviceBuffer(nbytes)
        hip.htod(a.ptr, src[:nbytes])
        hip.memset(b.ptr, 0, nbytes)
        t0 = time.perf_counter()
        if args.measure:
            r, h = L.librettPlanMeasure(len(dim), list(dim), list(perm), elem, stream, a.ptr, b.ptr)
        else:
            r, h = L.librettPlan(len(dim), list(dim), list(perm), elem, stream)
        plan_s = time.perf_counter() - t0
        if r != L.LIBRETT_SUCCESS:
            raise SystemExit("plan failed for %s %s: %d" % (dim, perm, r))
        info = L.plan_info(h)
        cases.append(dict(dim=dim, perm=perm, elem=elem, vol=vol, bytes=2 * nbytes, a=a, b=b, h=h, info=info,
                          plan_s=plan_s))
        log("plan %-28s %-20s %d B -> %s (%.1f ms)" % (dim, perm, elem, info["desc"], plan_s * 1e3))

    def run_step(events=None, k=0):
        for i, c in enumerate(cases):
            if events is not None:
                events[i][k][0].record(stream)
            rc = L.librettExecute(c["h"], c["a"].ptr, c["b"].ptr)
            if rc != L.LIBRETT_SUCCESS:
                raise SystemExit("execute failed: %d" % rc)
            if events is not None:
                events[i][k][1].record(stream)

    for _ in range(args.warmup):
        run_step()
    hip.stream_sync(stream)
    log("warmup done (%d steps)" % args.warmup)

    events = [[(hip.Event(), hip.Event()) for _ in range(args.steps)] for _ in cases]
    hip.device_sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        run_step(events, k)
    hip.stream_sync(stream)
    hip.device_sync()
    wall = time.perf_counter() - t0
    log("timed %d steps in %.3f s" % (args.steps, wall))

    step_bytes = sum(c["bytes"] for c in cases)
    value = step_bytes * args.steps / wall / 1e9
    per_case = []
    by_kernel = {}
    for i, c in enumerate(cases):
        ts = sorted(events[i][k][1].elapsed_ms(events[i][k][0]) for k in range(args.steps))
        avg = sum(ts) / len(ts)
        med = ts[len(ts) // 2]
        gbs = c["bytes"] / (avg * 1e-3) / 1e9
        per_case.append(dict(dim=list(c["dim"]), perm=list(c["perm"]), elem=c["elem"], plan=c["info"]["desc"],
                             method=c["info"]["method_name"], avg_ms=round(avg, 4), median_ms=round(med, 4),
                             best_ms=round(ts[0], 4), gbs=round(gbs, 1),
                             predicted_ms=round(c["info"]["predicted_s"] * 1e3, 4)))
        kname = c["info"]["method_name"] + ":" + c["info"]["desc"].split(" grid")[0]
        kb = by_kernel.setdefault(kname, [0.0, 0, 0])
        kb[0] += avg * args.steps
        kb[1] += c["bytes"] * args.steps
        kb[2] += args.steps
    dom_name, (dom_ms, dom_bytes, dom_launches) = max(by_kernel.items(), key=lambda kv: kv[1][0])
    dom_gbs = dom_bytes / (dom_ms * 1e-3) / 1e9
    # HBM bytes per launch of the dominant kernel from the committed counter
    # summary, when it was taken with the same plan.
    dom_traffic = None
    matched = []
    for c in cases:
        if c["info"]["method_name"] + ":" + c["info"]["desc"].split(" grid")[0] == dom_name:
            t = measured_traffic(c["dim"], c["perm"], c["elem"])
            matched.append(t["traffic_bytes"] if t and t.get("plan") == c["info"]["desc"] else None)
    if matched and all(m is not None for m in matched):
        dom_traffic = int(round(sum(matched) / len(matched)))  # every matched case runs once per step

    # Ceilings measured the same way: each transpose of a step replaced by a
    # copy of the same bytes between the same buffers, for the same steps.
    def ceiling(copy_fn):
        for c in cases:
            copy_fn(c)
        hip.stream_sync(stream)
        per = [[] for _ in cases]
        ev = [[(hip.Event(), hip.Event()) for _ in range(args.steps)] for _ in cases]
        t0 = time.perf_counter()
        for k in range(args.steps):
            for i, c in enumerate(cases):
                ev[i][k][0].record(stream)
                copy_fn(c)
                ev[i][k][1].record(stream)
        hip.stream_sync(stream)
        wall_c = time.perf_counter() - t0
        for i in range(len(cases)):
            per[i] = sum(ev[i][k][1].elapsed_ms(ev[i][k][0]) for k in range(args.steps)) / args.steps
        return step_bytes * args.steps / wall_c / 1e9, per

    d2d_gbs, d2d_per = ceiling(lambda c: hip.memcpy_d2d_async(c["b"].ptr, c["a"].ptr, c["bytes"] // 2, stream))
    devcopy_gbs, dc_per = ceiling(
        lambda c: L.device_copy(c["b"].ptr, c["a"].ptr, c["bytes"] // 2 // 16 * 16, stream, 1, 1, 0))
    for i, c in enumerate(per_case):
        c["d2d_gbs"] = round(cases[i]["bytes"] / (d2d_per[i] * 1e-3) / 1e9, 1)
        c["copy_kernel_gbs"] = round(cases[i]["bytes"] / (dc_per[i] * 1e-3) / 1e9, 1)

    for c in per_case:
        c["frac_d2d"] = round(c["gbs"] / c["d2d_gbs"], 3)
        c["frac_copy_kernel"] = round(c["gbs"] / c["copy_kernel_gbs"], 3)
        log("case %-28s %-20s %2d B %-11s %8.1f GB/s (%5.1f%% of D2D, %5.1f%% of copy kernel, %5.1f%% of 8 TB/s) "
            "model %.3f ms, run %.3f ms" % (tuple(c["dim"]), tuple(c["perm"]), c["elem"], c["method"], c["gbs"],
                                            100 * c["frac_d2d"], 100 * c["frac_copy_kernel"],
                                            100 * c["gbs"] / HBM_PEAK_GBS, c["predicted_ms"], c["avg_ms"]))
    log("D2D hipMemcpyAsync %.1f GB/s, streaming copy kernel %.1f GB/s; aggregate %.1f GB/s (%.1f%% of D2D, "
        "%.1f%% of the copy kernel)" % (d2d_gbs, devcopy_gbs, value, 100 * value / d2d_gbs, 100 * value / devcopy_gbs))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_reps)

    for c in cases:
        L.librettDestroy(c["h"])
        c["a"].free()
        c["b"].free()

    line = {
        "metric": "transpose GB/s (2*bytes/t) vs HBM peak",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (opaque 4/8-byte moves)",
        "data": "synthetic (seeded uniform [-1,1))",
        "config": {
            "workload": "large single transposes ~256 MiB each (7 named cases, f32 and f64)",
            "cases": len(cases),
            "bytes_per_step": step_bytes,
            "plan_selection": selection,
        },
        "d2d_copy_gbs": round(d2d_gbs, 1),
        "frac_of_d2d": round(value / d2d_gbs, 3),
        "copy_kernel_gbs": round(devcopy_gbs, 1),
        "frac_of_copy_kernel": round(value / devcopy_gbs, 3),
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": round(dom_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
            "traffic": dom_traffic,
            "traffic_over_algorithmic": round(dom_traffic / (dom_bytes / dom_launches), 4) if dom_traffic else None,
            "launches": dom_launches,
            "avg_launch_ms": round(dom_ms / dom_launches, 4),
        },
        "cpu_baseline": cpu,
        "reference_hip_build": "not measured: the reference sources are not available on this machine (no LIBRETT_REF_DIR)",
        "plan_ms": {"%s %s %dB" % (tuple(c["dim"]), tuple(c["perm"]), c["elem"]): round(c["plan_s"] * 1e3, 2)
                    for c in cases},
        "cases": per_case,
    }
    return line


def bench_large_random(args, stream):
    """Seeded ~256 MiB transposes of three classes (leading dim kept, moved,
    small) in f32 and f64: GB/s of librettPlan's and librettPlanMeasure's
    plans (HIP events, median of the timed steps), their ratio, the same-
    buffer D2D copy, and an on-device check of every output."""
    import librett_amd as L
    from librett_amd import hip, workloads as W
    cap = 300 << 20
    A, B = hip.DeviceBuffer(cap), hip.DeviceBuffer(cap)
    e0, e1 = hip.Event(), hip.Event()

    def timed(fn, n):
        fn()
        hip.stream_sync(stream)
        ts = []
        for _ in range(n):
            e0.record(stream)
            fn()
            e1.record(stream)
            e1.synchronize()
            ts.append(e1.elapsed_ms(e0) * 1e-3)
        return sorted(ts)[len(ts) // 2]

    def timed_pair(f1, f2, n):
        # Both warmed up, then their runs alternate (f1 first on even rounds,
        # f2 first on odd ones), so neither is timed in a systematically
        # different state of the device; medians of each.
        f1()
        f2()
        hip.stream_sync(stream)
        t1, t2 = [], []
        for i in range(n):
            order = ((f1, t1), (f2, t2)) if i % 2 == 0 else ((f2, t2), (f1, t1))
            for fn, ts in order:
                e0.record(stream)
                fn()
                e1.record(stream)
                e1.synchronize()
                ts.append(e1.elapsed_ms(e0) * 1e-3)
        return sorted(t1)[len(t1) // 2], sorted(t2)[len(t2) // 2]

    rows = []
    steps = max(3, args.steps // 2)
    for elem in (4, 8):
        for cls in ("keep", "move", "small"):
            for seed in range(8):
                d, p = W.large_case(20261202 + seed, elem, cls)
                n = int(np.prod(d)) * elem
                L.check_fill(A.ptr, n, stream)
                r1, h1 = L.librettPlan(len(d), d, p, elem, stream)
                r2, h2 = L.librettPlanMeasure(len(d), d, p, elem, stream, A.ptr, B.ptr)
                if r1 != L.LIBRETT_SUCCESS or r2 != L.LIBRETT_SUCCESS:
                    raise SystemExit("plan failed for %s %s" % (d, p))
                t_plan, t_meas = timed_pair(lambda: L.librettExecute(h1, A.ptr, B.ptr),
                                            lambda: L.librettExecute(h2, A.ptr, B.ptr), steps)
                L.librettExecute(h1, A.ptr, B.ptr)
                bad1 = L.check_transpose(B.ptr, d, p, elem, stream)[0]
                L.librettExecute(h2, A.ptr, B.ptr)
                bad2 = L.check_transpose(B.ptr, d, p, elem, stream)[0]
                t_d2d = timed(lambda: hip.memcpy_d2d_async(B.ptr, A.ptr, n, stream), steps)
                if bad1 or bad2:
                    raise SystemExit("wrong output for %s %s: %d / %d elements" % (d, p, bad1, bad2))
                row = dict(cls=cls, elem=elem, seed=seed, dim=d, perm=p, bytes=2 * n,
                           plan=L.plan_info(h1)["desc"], measured=L.plan_info(h2)["desc"],
                           plan_gbs=round(2 * n / t_plan / 1e9, 1), measure_gbs=round(2 * n / t_meas / 1e9, 1),
                           d2d_gbs=round(2 * n / t_d2d / 1e9, 1), plan_over_measure=round(t_plan / t_meas, 3))
                rows.append(row)
                log("%-5s f%d seed %d %-32s %-22s plan %7.1f GB/s measure %7.1f GB/s (x%.3f) D2D %7.1f" % (
                    cls, 8 * elem, seed, d, p, row["plan_gbs"], row["measure_gbs"], row["plan_over_measure"],
                    row["d2d_gbs"]))
                L.librettDestroy(h1)
                L.librettDestroy(h2)
    tot = sum(r["bytes"] for r in rows)
    t_plan = sum(r["bytes"] / (r["plan_gbs"] * 1e9) for r in rows)
    t_meas = sum(r["bytes"] / (r["measure_gbs"] * 1e9) for r in rows)
    t_d2d = sum(r["bytes"] / (r["d2d_gbs"] * 1e9) for r in rows)
    by_cls = {}
    for cls in ("keep", "move", "small"):
        rs = [r for r in rows if r["cls"] == cls]
        by_cls[cls] = {"plan_gbs": round(sum(r["bytes"] for r in rs) / sum(r["bytes"] / (r["plan_gbs"] * 1e9)
                                                                          for r in rs) / 1e9, 1),
                       "frac_d2d_min": round(min(r["plan_gbs"] / r["d2d_gbs"] for r in rs), 3),
                       "frac_d2d_median": round(float(np.median([r["plan_gbs"] / r["d2d_gbs"] for r in rs])), 3)}
    line = {
        "metric": "transpose GB/s (2*bytes/t), seeded ~256 MiB cases, librettPlan vs librettPlanMeasure",
        "value": round(tot / t_plan / 1e9, 1), "unit": "GB/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": round(t_plan * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64 (opaque 4/8-byte moves)", "data": "synthetic (index pattern, checked on the device)",
        "config": {"workload": "48 seeded ~256 MiB transposes: leading dim kept / moved / small, f32 and f64, "
                               "8 seeds each, rank 2-6"},
        "measure_gbs": round(tot / t_meas / 1e9, 1), "d2d_gbs": round(tot / t_d2d / 1e9, 1),
        "frac_of_d2d": round(t_d2d / t_plan, 3), "plan_over_measure_time_median":
            float(np.median([r["plan_over_measure"] for r in rows])),
        "plan_over_measure_time_max": max(r["plan_over_measure"] for r in rows),
        "by_class": by_cls, "cases": rows, "roofline": None, "cpu_baseline": None,
    }
    A.free()
    B.free()
    return line


def summarize_large_random(r):
    worst = sorted(r["cases"], key=lambda c: -c["plan_over_measure"])[:5]
    return {"gbs": r["value"], "measure_gbs": r["measure_gbs"], "d2d_gbs": r["d2d_gbs"],
            "frac_of_d2d": r["frac_of_d2d"], "plan_over_measure_time_median": r["plan_over_measure_time_median"],
            "plan_over_measure_time_max": r["plan_over_measure_time_max"], "by_class": r["by_class"],
            "checked": "every output checked on the device (index pattern)", "cases": len(r["cases"]),
            "worst": [{k: c[k] for k in ("cls", "elem", "seed", "dim", "perm", "plan_gbs", "measure_gbs",
                                         "plan_over_measure")} for c in worst]}


def bench_selector(args, stream):
    """Learned plan selector: every candidate of each held-out problem timed
    with librettPlanMeasure's loop; the slowdown of the selector's choice and
    of the analytic model's choice versus the fastest candidate."""
    import librett_amd as L
    from librett_amd import selector as S
    if not L.selector_generation():
        raise SystemExit("no selector weights loaded (librett_amd/selector_weights.bin)")
    A, B = S._buffers(512 << 20)
    t0 = time.perf_counter()
    recs = S.collect(S.problems(args.heldout, 1001), stream, A.ptr, B.ptr, log)
    coll = time.perf_counter() - t0
    p, mean, std = S.read_weights(S.DEFAULT_WEIGHTS)
    t1 = time.perf_counter()
    sc = S.scores(p, mean, std, recs, stream)
    infer = time.perf_counter() - t1
    margin = L.selector_margin()
    sel, ana = S.evaluate(recs, sc, margin)
    pure, _ = S.evaluate(recs, sc)
    ss, sa, sp = S.summary(sel), S.summary(ana), S.summary(pure)
    ncand = sum(len(r["times"]) for r in recs)
    # One librettPlan-style choice per problem (candidates, features, the
    # ensemble's inference kernel and the margin rule), timed on the host.
    nch = min(100, len(recs))
    t3 = time.perf_counter()
    for r in recs[:nch]:
        L.selector_choose(r["dims"], r["perm"], r["elem"])
    choose_us = (time.perf_counter() - t3) / max(1, nch) * 1e6
    log("selector choice per problem (candidates, features, inference): %.0f us (mean of %d)" % (choose_us, nch))
    # The pairs that decide each slowdown, timed again head to head.
    t2 = time.perf_counter()
    pk = S.picks(recs, sc, margin)
    rs = S.summary(S.retime(recs, [(k, i, b) for k, i, m, b in pk], stream, A.ptr, B.ptr))
    ra = S.summary(S.retime(recs, [(k, m, b) for k, i, m, b in pk if m >= 0], stream, A.ptr, B.ptr))
    log("re-timed head to head (15 runs each): librettPlan mean %.4f (p50 %.3f p90 %.3f p95 %.3f p99 %.3f max %.3f, "
        "%.1f%% within 5%%); analytic %.4f (p90 %.3f p95 %.3f p99 %.3f); %.0f s"
        % (rs["mean"], rs["p50"], rs["p90"], rs["p95"], rs["p99"], rs["max"], 100 * rs["within_5pct"], ra["mean"],
           ra["p90"], ra["p95"], ra["p99"], time.perf_counter() - t2))
    log("librettPlan (selector, margin %.2f): mean slowdown %.4f (p50 %.3f p90 %.3f p95 %.3f p99 %.3f, %.1f%% within 5%%); "
        "selector alone %.4f (p90 %.3f, %.1f%% within 5%%); analytic %.4f (p50 %.3f p90 %.3f p99 %.3f, %.1f%% within "
        "5%%); %d problems, %d candidates timed in %.0f s; inference %.1f ms"
        % (margin, ss["mean"], ss["p50"], ss["p90"], ss["p95"], ss["p99"], 100 * ss["within_5pct"], sp["mean"], sp["p90"],
           100 * sp["within_5pct"], sa["mean"], sa["p50"], sa["p90"], sa["p99"], 100 * sa["within_5pct"], len(recs),
           ncand, coll, infer * 1e3))
    line = {
        "metric": "plan selection: mean slowdown of the chosen plan vs the exhaustive best (held-out problems)",
        "value": round(ss["mean"], 4), "unit": "x", "n_gpus": 1, "steps": len(recs), "warmup": 0,
        "ms_per_step": round(coll / max(1, len(recs)) * 1e3, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (MLP); f32/f64/c128 opaque moves timed", "data": "synthetic",
        "config": {"workload": "learned selector: %d held-out problems, rank 2-7, 1 MB-512 MB, f32/f64/c128"
                               % len(recs), "seed": 1001, "candidates_timed": ncand},
        "selector": ss, "selector_margin": margin, "selector_alone": sp, "analytic": sa,
        "retimed": {"librettPlan": rs, "analytic": ra, "runs_each": 15},
        "inference_ms_all_candidates": round(infer * 1e3, 2), "choice_us_per_problem": round(choose_us, 1),
        "roofline": None, "cpu_baseline": None,
    }
    A.free()
    B.free()
    return line


def summarize_selector(r):
    return {"problems": r["config"]["workload"], "candidates_timed": r["config"]["candidates_timed"],
            "librettPlan_slowdown": r["selector"], "selector_alone_slowdown": r["selector_alone"],
            "analytic_slowdown": r["analytic"], "margin": r["selector_margin"],
            "retimed_head_to_head": r["retimed"], "choice_us_per_problem": r["choice_us_per_problem"]}


def batched_problems(n, seed=20261202):
    """Uneven batch: rank 3-5, volume log-uniform 1e3..1e6 elements, float64
    or complex128, random non-identity permutation per problem."""
    from librett_amd import workloads as W
    return W.batched_problems(n, seed)


def bench_batched(args, stream):
    """Batched workload: the whole batch in one persistent graph launch vs a
    loop of per-problem librettExecute calls on the same buffers. Inputs hold
    the index pattern, and every output of the timed batch is checked on the
    device afterwards."""
    import librett_amd as L
    from librett_amd import hip
    probs = batched_problems(args.problems)
    sizes = [int(np.prod(d)) * e for d, p, e in probs]
    offs, o = [], 0
    for s_ in sizes:
        offs.append(o)
        o += (s_ + 255) // 256 * 256
    total = o
    alg = sum(2 * s_ for s_ in sizes)
    log("batched: %d problems, %.2f GB algorithmic, %.2f GB per side" % (len(probs), alg / 1e9, total / 1e9))
    A = hip.DeviceBuffer(total)
    B = hip.DeviceBuffer(total)
    for off, s_ in zip(offs, sizes):
        L.check_fill(A.ptr + off, s_, stream)
    hip.memset(B.ptr, 0, total)
    hip.stream_sync(stream)
    ins = [A.ptr + o for o in offs]
    outs = [B.ptr + o for o in offs]
    e0, e1 = hip.Event(), hip.Event()

    def timed(fn, steps, warm):
        for _ in range(warm):
            fn()
        hip.stream_sync(stream)
        e0.record(stream)
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return (time.perf_counter() - t0) / steps, e1.elapsed_ms(e0) * 1e-3 / steps

    def check_all(what):
        hip.stream_sync(stream)
        bad = [i for i, ((d, p, e), off) in enumerate(zip(probs, offs))
               if L.check_transpose(B.ptr + off, d, p, e, stream)[0]]
        if bad:
            raise SystemExit("batched (%s): %d wrong outputs, first %s" % (what, len(bad), probs[bad[0]]))
        hip.memset(B.ptr, 0, total)

    # Per-call loop of librettExecute with plans built beforehand.
    t0 = time.perf_counter()
    handles = []
    for d, p, e in probs:
        r, h = L.librettPlan(len(d), list(d), list(p), e, stream)
        if r != L.LIBRETT_SUCCESS:
            raise SystemExit("plan failed")
        handles.append(h)
    plan_loop_s = time.perf_counter() - t0

    def loop():
        for h, i, o_ in zip(handles, ins, outs):
            if L.librettExecute(h, i, o_) != L.LIBRETT_SUCCESS:
                raise SystemExit("librettExecute failed in the per-call loop")
    wall_loop, ev_loop = timed(loop, max(1, args.steps // 4), 1)
    check_all("per-call loop")
    log("per-call loop: %.2f ms/step (%.1f GB/s), planning %.1f s" % (wall_loop * 1e3, alg / wall_loop / 1e9,
                                                                       plan_loop_s))
    res = {}
    for tune in (0, args.tune):
        t0 = time.perf_counter()
        bt = L.Batch(probs, stream, tune=tune)
        bt.set_buffers(ins, outs)
        info = bt.info()
        build_s = time.perf_counter() - t0

        def run(graph):
            if bt.execute(use_graph=graph) != L.LIBRETT_SUCCESS:
                raise SystemExit("batch execute failed")
        wall_b, ev_b = timed(lambda: run(True), args.steps, args.warmup)
        check_all("graph, tune %d" % tune)
        wall_n, ev_n = timed(lambda: run(False), args.steps, args.warmup)
        check_all("plain launch, tune %d" % tune)
        res[tune] = dict(wall=wall_b, ev=ev_b, wall_plain=wall_n, info=info, build_s=build_s)
        log("batch tune=%d: graph %.3f ms (%.1f GB/s), plain launch %.3f ms; units %d grid %d; methods %s; "
            "plan %.2f s tune %.2f s" % (tune, wall_b * 1e3, alg / wall_b / 1e9, wall_n * 1e3, info["total_units"],
                                        info["grid"], info["method_count"], info["plan_seconds"],
                                        info["tune_seconds"]))
        bt.destroy()
    # Same-buffer copy ceiling: every problem's bytes copied in one pass.
    d2d = timed(lambda: hip.memcpy_d2d_async(B.ptr, A.ptr, total, stream), args.steps, 1)[0]
    best_tune = min(res, key=lambda k: res[k]["wall"])
    wall = res[best_tune]["wall"]
    value = alg / wall / 1e9
    for h in handles:
        L.librettDestroy(h)
    line = {
        "metric": "batched transpose GB/s (2*bytes/t), one persistent graph launch per batch",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+c128 (opaque 8/16-byte moves)",
        "data": "synthetic (index pattern; every output checked on the device)",
        "config": {"workload": "batched: %d uneven transposes, rank 3-5, vol 1e3-1e6, f64/c128" % len(probs),
                   "bytes_per_step": alg, "autotune_candidates": best_tune},
        "per_call_loop_gbs": round(alg / wall_loop / 1e9, 1),
        "speedup_vs_per_call": round(wall_loop / wall, 2),
        "d2d_copy_gbs": round(2 * total / d2d / 1e9, 1),
        "frac_of_d2d": round(value / (2 * total / d2d / 1e9), 3),
        "model_only_gbs": round(alg / res[0]["wall"] / 1e9, 1),
        "autotuned_gbs": round(alg / res[args.tune]["wall"] / 1e9, 1),
        "plain_launch_gbs": round(alg / res[best_tune]["wall_plain"] / 1e9, 1),
        "autotune_seconds": round(res[args.tune]["info"]["tune_seconds"], 2),
        "batch_info": res[best_tune]["info"],
    }
    A.free()
    B.free()
    return line


def summarize_batched(r):
    keys = ("value", "ms_per_step", "per_call_loop_gbs", "speedup_vs_per_call", "d2d_copy_gbs", "frac_of_d2d",
            "model_only_gbs", "autotuned_gbs", "plain_launch_gbs", "autotune_seconds")
    out = {("gbs" if k == "value" else k): r[k] for k in keys}
    out["workload"] = r["config"]["workload"]
    out["checked"] = r["data"]
    return out


def serving_checked_pass(srv, stream, n, seed):
    """Requests of the serving distribution submitted from the host with
    index-pattern inputs in the arena, all in flight together; every output
    checked on the device. Returns the number of requests checked."""
    import librett_amd as L
    from librett_amd import hip, workloads as W
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n):
        e = int(rng.choice([4, 8, 16]))
        nb = float(np.exp(rng.uniform(np.log(4096), np.log(64 << 20))))
        r = int(rng.integers(2, 7))
        reqs.append(W.random_problem(rng, r, r, max(2.0, nb / e), max(2.0, nb / e), [e]))
    for wave in range(0, n, 100):
        part = reqs[wave:wave + 100]
        bufs = []
        for d, p, e in part:
            nb = int(np.prod(d)) * e
            pin, pout = srv.alloc(nb), srv.alloc(nb)
            if pin is None or pout is None:
                raise SystemExit("serving check: arena full")
            L.check_fill(pin, nb, stream)
            bufs.append((pin, pout))
        hip.stream_sync(stream)
        tickets = [srv.submit(d, p, e, pi, po) for (d, p, e), (pi, po) in zip(part, bufs)]
        for t in tickets:
            srv.wait(t)
        for (d, p, e), (pi, po) in zip(part, bufs):
            bad = L.check_transpose(po, d, p, e, stream)[0]
            if bad:
                raise SystemExit("serving: wrong output for %s %s %d B (%d elements)" % (d, p, e, bad))
            srv.free(pi)
            srv.free(po)
    return n


def bench_serving(args, stream):
    """Serving workload: Poisson arrivals of requests 4 KiB-64 MiB (log-uniform),
    rank 2-6, f32/f64/c128, random permutations, through the continuous-
    batching engine; admission reserves input and output bytes in the device
    arena. Measured with a warm plan cache (shapes from a catalog planned up
    front) and with a new shape per request (planned on arrival by the client
    threads); compared with per-call librettExecute loops over the same mix."""
    import librett_amd as L
    from librett_amd import hip
    pool_bytes = 1 << 30
    pool = hip.DeviceBuffer(pool_bytes)
    hip.memset(pool.ptr, 5, pool_bytes)
    srv = L.Server(int(args.arena_gb * (1 << 30)), slots=args.slots, max_batch=args.max_batch)
    # One client thread submits by default: its admission and submission take
    # a few microseconds a request, so on a slow host the request rate, not
    # the device, can bound saturation (more client threads measured slower:
    # they contend in the submission path).
    kw = dict(catalog=1024, min_bytes=4096, max_bytes=64 << 20, rank_lo=2, rank_hi=6, elems=(4, 8, 16),
              inputs_in_arena=True, clients=args.clients)
    # Capacity: a rate far above what the device sustains (closed-loop saturation).
    srv.run_poisson(pool.ptr, pool_bytes, qps=1e7, requests=2000, seed=1, **kw)  # warm-up
    srv.reset_stats()
    sat = srv.run_poisson(pool.ptr, pool_bytes, qps=1e7, requests=args.requests, seed=2, **kw)
    mean_req_bytes = sat["bytes"] / max(1, sat["completed"])
    cap_qps = sat["gbs"] * 1e9 / mean_req_bytes
    log("saturation (warm plans): %.1f GB/s, %.0f req/s, mean request %.2f MB moved, mean batch %.1f" % (
        sat["gbs"], cap_qps, mean_req_bytes / 1e6, sat["mean_batch"]))
    rates = []
    for frac in (0.25, 0.5, 0.75, 0.9):
        srv.reset_stats()
        st = srv.run_poisson(pool.ptr, pool_bytes, qps=frac * cap_qps, requests=args.requests, seed=3, **kw)
        rates.append(dict(load=frac, qps=round(frac * cap_qps), gbs=round(st["gbs"], 1), p50_us=round(st["p50_us"], 1),
                          p90_us=round(st["p90_us"], 1), p99_us=round(st["p99_us"], 1),
                          mean_batch=round(st["mean_batch"], 2), stalls=st["admission_stalls"],
                          arena_peak_gb=round(st["arena_peak_bytes"] / 1e9, 2)))
        log("load %.2f: %8.0f req/s  %7.1f GB/s  p50 %8.1f us  p99 %8.1f us  batch %.1f  stalls %d" % (
            frac, frac * cap_qps, st["gbs"], st["p50_us"], st["p99_us"], st["mean_batch"], st["admission_stalls"]))
    # New shape per request: plans are made on arrival by 8 client threads.
    fresh = dict(kw, fresh=True, clients=max(8, args.clients))
    srv.reset_stats()
    fsat = srv.run_poisson(pool.ptr, pool_bytes, qps=1e7, requests=args.requests, seed=5, **fresh)
    fcap = fsat["gbs"] * 1e9 / max(1.0, fsat["bytes"] / max(1, fsat["completed"]))
    srv.reset_stats()
    fhalf = srv.run_poisson(pool.ptr, pool_bytes, qps=0.5 * fcap, requests=args.requests, seed=6, **fresh)
    plan_us = 1e6 * fsat["plan_seconds"] / max(1, fsat["plan_misses"])
    log("fresh shapes: saturation %.1f GB/s (%.0f req/s), cold plan %.0f us per shape; at half that rate p50 %.1f us "
        "p99 %.1f us" % (fsat["gbs"], fcap, plan_us, fhalf["p50_us"], fhalf["p99_us"]))
    checked = serving_checked_pass(srv, stream, 300, 8)
    srv.destroy()
    # Per-call baseline: the same catalog through librettExecute, one stream,
    # plans built beforehand, back to back.
    cat = serving_catalog(kw, 1024)
    handles = []
    for d, p, e in cat:
        r, h = L.librettPlan(len(d), list(d), list(p), e, stream)
        if r != L.LIBRETT_SUCCESS:
            raise SystemExit("plan failed for %s %s" % (d, p))
        handles.append((h, int(np.prod(d)) * e))
    out = hip.DeviceBuffer(64 << 20)
    rng = np.random.default_rng(4)
    picks = rng.integers(0, len(handles), size=args.requests)
    offs = rng.integers(0, (pool_bytes - (64 << 20)) // 256, size=args.requests) * 256

    def call(i):
        h, nb = handles[picks[i]]
        if L.librettExecute(h, pool.ptr + int(offs[i]), out.ptr) != L.LIBRETT_SUCCESS:
            raise SystemExit("librettExecute failed in the per-call loop")
        return 2 * nb
    for i in range(200):
        call(i)
    hip.stream_sync(stream)
    t0 = time.perf_counter()
    moved = 0
    for i in range(args.requests):
        moved += call(i)
    hip.stream_sync(stream)
    loop_gbs = moved / (time.perf_counter() - t0) / 1e9
    log("per-call loop: %.1f GB/s" % loop_gbs)
    # The same calls with a stream synchronisation after each one: a request
    # handler without batching that returns every result before taking the
    # next request.
    t0 = time.perf_counter()
    moved = 0
    for i in range(args.requests):
        moved += call(i)
        hip.stream_sync(stream)
    sync_gbs = moved / (time.perf_counter() - t0) / 1e9
    log("per-call loop, synchronised per request: %.1f GB/s" % sync_gbs)
    for h, _ in handles:
        L.librettDestroy(h)
    out.free()
    pool.free()
    line = {
        "metric": "serving transpose GB/s and p50/p99 latency (Poisson arrivals, continuous batching)",
        "value": round(sat["gbs"], 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.requests,
        "warmup": 2000,
        "ms_per_step": round(sat["wall_seconds"] / max(1, sat["completed"]) * 1e3, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64/c128 (opaque moves)",
        "data": "synthetic",
        "config": {"workload": "serving: Poisson arrivals, 4 KiB-64 MiB log-uniform, rank 2-6, f32/f64/c128, "
                               "%g GiB arena gating input + output bytes, LRU plan cache, %d client threads" % (
                                   args.arena_gb, args.clients),
                   "slots": args.slots, "max_batch": args.max_batch, "clients": args.clients},
        "saturation": {"gbs": round(sat["gbs"], 1), "req_per_s": round(cap_qps), "p50_us": round(sat["p50_us"], 1),
                       "p99_us": round(sat["p99_us"], 1), "mean_batch": round(sat["mean_batch"], 2),
                       "shapes": "catalog of 1024, planned up front"},
        "load_sweep": rates,
        "fresh_shapes": {"saturation_gbs": round(fsat["gbs"], 1), "req_per_s": round(fcap),
                         "cold_plan_us_mean": round(plan_us, 1), "plan_misses": fsat["plan_misses"],
                         "client_threads": max(8, args.clients), "half_load_p50_us": round(fhalf["p50_us"], 1),
                         "half_load_p99_us": round(fhalf["p99_us"], 1), "half_load_gbs": round(fhalf["gbs"], 1)},
        "per_call_loop_gbs": round(loop_gbs, 1),
        "speedup_vs_per_call": round(sat["gbs"] / loop_gbs, 2),
        "per_call_synchronised_gbs": round(sync_gbs, 1),
        "speedup_vs_per_call_synchronised": round(sat["gbs"] / sync_gbs, 2),
        "checked_requests": checked,
    }
    return line


def summarize_serving(r):
    out = {k: r[k] for k in ("saturation", "load_sweep", "fresh_shapes", "per_call_loop_gbs", "speedup_vs_per_call",
                             "per_call_synchronised_gbs", "speedup_vs_per_call_synchronised")}
    out["workload"] = r["config"]["workload"]
    out["checked"] = "%d requests of the same distribution checked on the device" % r["checked_requests"]
    return out


def serving_catalog(kw, n):
    """Host copy of the request shapes the serving engine draws (for the
    per-call baseline): same distribution, numpy generator."""
    from librett_amd import workloads as W
    rng = np.random.default_rng(kw.get("seed", 77))
    out = []
    for _ in range(n):
        e = int(rng.choice(kw["elems"]))
        nbytes = float(np.exp(rng.uniform(np.log(kw["min_bytes"]), np.log(kw["max_bytes"]))))
        rank = int(rng.integers(kw["rank_lo"], kw["rank_hi"] + 1))
        d, p, _ = W.random_problem(rng, rank, rank, max(2.0, nbytes / e), max(2.0, nbytes / e), [e])
        out.append((d, p, e))
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_named_large_cases.json")


def _pmc_rows(d):
    import csv
    import glob
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection.csv under %s" % d)
    rows = list(csv.DictReader(open(files[0])))
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    return rows


def pmc_summary(fetch_dir, write_dir, run_log, json_path=None):
    """HBM traffic of the named cases from two counter runs of
    'bench.py --steps 1 --warmup 0 --no-cpu'. The counters are scaled by
    what they read for the streaming copy kernel of the same run, whose
    traffic is known (Grid_Size 16-byte words read and written)."""
    fr, wr = _pmc_rows(fetch_dir), _pmc_rows(write_dir)
    if len(fr) != len(wr):
        raise SystemExit("the two runs dispatched different kernel counts")
    fetch = [float(r["Counter_Value"]) * 1024 for r in fr]   # KiB -> bytes
    write = [float(r["Counter_Value"]) * 1024 for r in wr]
    names = [r["Kernel_Name"] for r in fr]
    grids = [int(r["Grid_Size"]) for r in fr]
    cal_r, cal_w = [], []
    for n, g, f, w in zip(names, grids, fetch, write):
        if "k_devcopy" in n:
            cal_r.append(g * 16 / f)
            cal_w.append(g * 16 / w)
    if not cal_r:
        raise SystemExit("no copy-kernel dispatches to calibrate against")
    fr_k, fw_k = float(np.median(cal_r)), float(np.median(cal_w))
    transposes = [i for i, n in enumerate(names) if any(k in n for k in ("k_tiled", "k_copy<", "k_packed"))]
    plans = [ln.split(" -> ", 1)[1].rsplit(" (", 1)[0].strip() for ln in open(run_log) if ln.startswith("plan ")]
    out = {"fetch_factor": round(fr_k, 4), "write_factor": round(fw_k, 4), "cases": []}
    lines = ["# HBM traffic of the named large cases (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)", "",
             "FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by the factor that makes the streaming copy",
             "kernel of the same run read and write exactly its bytes: reads x%.4f, writes x%.4f." % (fr_k, fw_k), "",
             "| case | kernel | algorithmic MB | read MB | write MB | traffic / algorithmic |", "|---|---|---|---|---|---|"]
    for (dim, perm, elem), i in zip(CASES, transposes[:len(CASES)]):
        alg = 2 * int(np.prod(dim)) * elem
        rd, wt = fetch[i] * fr_k, write[i] * fw_k
        ent = {"dim": list(dim), "perm": list(perm), "elem": elem, "kernel": names[i],
               "plan": plans[len(out["cases"])] if len(plans) > len(out["cases"]) else None, "algorithmic_bytes": alg,
               "read_bytes": round(rd), "write_bytes": round(wt), "traffic_bytes": round(rd + wt),
               "ratio": round((rd + wt) / alg, 4), "read_ratio": round(rd / (alg / 2), 4),
               "write_ratio": round(wt / (alg / 2), 4)}
        out["cases"].append(ent)
        lines.append("| %s %s %dB | `%s` | %.1f | %.1f | %.1f | %.3f (reads %.3f, writes %.3f) |" % (
            tuple(dim), tuple(perm), elem, names[i].split("(")[0], alg / 1e6, rd / 1e6, wt / 1e6, ent["ratio"],
            ent["read_ratio"], ent["write_ratio"]))
    print("\n".join(lines))
    if json_path:
        with open(json_path, "w") as fh:
            fh.write("{\"fetch_factor\": %s, \"write_factor\": %s, \"cases\": [\n" % (out["fetch_factor"],
                                                                                   out["write_factor"]))
            fh.write(",\n".join(json.dumps(c) for c in out["cases"]))
            fh.write("\n]}\n")


def _pmc_table(d):
    """{dispatch id: (kernel, grid, {counter: value})} of one counter run."""
    import csv
    import glob
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection.csv under %s" % d)
    t = {}
    for r in csv.DictReader(open(files[0])):
        k = int(r["Dispatch_Id"])
        ent = t.setdefault(k, (r["Kernel_Name"], int(r["Grid_Size"]), {}))
        ent[2][r["Counter_Name"]] = ent[2].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return t


def pmc_requests(read_dir, write_dir):
    """Checks what FETCH_SIZE and WRITE_SIZE count on this GPU against the raw
    TCC/EA request counters they are built from. FETCH_SIZE charges 128 B per
    TCC_BUBBLE, 32 B per TCC_EA0_RDREQ_32B and 64 B per other TCC_EA0_RDREQ;
    WRITE_SIZE 64 B per TCC_EA0_WRREQ_64B and 32 B per other TCC_EA0_WRREQ.
    For the streaming copy kernel the bytes are known (Grid_Size 16-byte words
    each way), so bytes / requests gives the request size the hardware
    actually moves."""
    rd, wr = _pmc_table(read_dir), _pmc_table(write_dir)
    lines = ["| kernel | dispatches | MB read (alg.) | RDREQ | BUBBLE | RDREQ_32B | RDREQ_DRAM | B per RDREQ | "
             "FETCH_SIZE formula / bytes | MB written (alg.) | WRREQ | WRREQ_64B | WRREQ_DRAM | B per WRREQ | "
             "WRITE_SIZE formula / bytes |", "|---" * 15 + "|"]
    seen = {}
    # The first transpose dispatches are the named cases in order (one step,
    # no warm-up); their algorithmic bytes per side are vol * elem.
    case_bytes = [float(np.prod(d)) * e for d, _, e in CASES]
    ntr = 0
    for k in sorted(rd):
        name, grid, c = rd[k]
        w = wr.get(k, (name, grid, {}))[2]
        short = name.split("(")[0].replace("void ", "")
        alg = None
        if "k_devcopy" in name:
            alg = grid * 16.0
        elif any(t in name for t in ("k_tiled", "k_copy<", "k_packed")):
            if ntr < len(case_bytes):
                alg = case_bytes[ntr]
                short = "%s %s" % (short, "x".join(str(x) for x in CASES[ntr][0]))
            ntr += 1
        ent = seen.setdefault(short, {"n": 0, "alg": 0.0, "rd": 0.0, "bub": 0.0, "r32": 0.0, "rdram": 0.0,
                                      "wr": 0.0, "w64": 0.0, "wdram": 0.0, "known": alg is not None})
        ent["n"] += 1
        ent["alg"] += alg or 0.0
        ent["rd"] += c.get("TCC_EA0_RDREQ_sum", 0.0)
        ent["bub"] += c.get("TCC_BUBBLE_sum", 0.0)
        ent["r32"] += c.get("TCC_EA0_RDREQ_32B_sum", 0.0)
        ent["rdram"] += c.get("TCC_EA0_RDREQ_DRAM_sum", 0.0)
        ent["wr"] += w.get("TCC_EA0_WRREQ_sum", 0.0)
        ent["w64"] += w.get("TCC_EA0_WRREQ_64B_sum", 0.0)
        ent["wdram"] += w.get("TCC_EA0_WRREQ_DRAM_sum", 0.0)
    for short, e in seen.items():
        if not e["known"] or e["rd"] <= 0 or e["wr"] <= 0:
            continue
        fetch = e["bub"] * 128 + (e["rd"] - e["bub"] - e["r32"]) * 64 + e["r32"] * 32
        write = (e["wr"] - e["w64"]) * 32 + e["w64"] * 64
        lines.append("| `%s` | %d | %.1f | %.0f | %.0f | %.0f | %.0f | %.1f | %.4f | %.1f | %.0f | %.0f | %.0f | %.1f | "
                     "%.4f |" % (short, e["n"], e["alg"] / 1e6, e["rd"], e["bub"], e["r32"], e["rdram"],
                                 e["alg"] / e["rd"], fetch / e["alg"], e["alg"] / 1e6, e["wr"], e["w64"],
                                 e["wdram"], e["alg"] / e["wr"], write / e["alg"]))
    lines += ["", "Every dispatch, all kernels (requests, not bytes):", "",
              "| dispatch | kernel | RDREQ | BUBBLE | RDREQ_32B | RDREQ_DRAM | WRREQ | WRREQ_64B | WRREQ_DRAM |",
              "|---|---|---|---|---|---|---|---|---|"]
    for k in sorted(rd):
        name, grid, c = rd[k]
        w = wr.get(k, (name, grid, {}))[2]
        lines.append("| %d | `%s` | %.0f | %.0f | %.0f | %.0f | %.0f | %.0f | %.0f |" % (
            k, name.split("(")[0].replace("void ", "")[:60], c.get("TCC_EA0_RDREQ_sum", 0), c.get("TCC_BUBBLE_sum", 0),
            c.get("TCC_EA0_RDREQ_32B_sum", 0), c.get("TCC_EA0_RDREQ_DRAM_sum", 0), w.get("TCC_EA0_WRREQ_sum", 0),
            w.get("TCC_EA0_WRREQ_64B_sum", 0), w.get("TCC_EA0_WRREQ_DRAM_sum", 0)))
    print("\n".join(lines))


def measured_traffic(dim, perm, elem):
    """HBM bytes per launch of a named case from the committed counter summary, or None."""
    try:
        data = json.load(open(TRAFFIC_FILE))
    except (OSError, ValueError):
        return None
    for c in data.get("cases", []):
        if c["dim"] == list(dim) and c["perm"] == list(perm) and c["elem"] == elem:
            return c
    return None


def cpu_baseline(reps):
    """The oracle's OpenMP permutation on the host, over the same seven cases
    at full size (the bounded sample is one median-of-`reps` run per case)."""
    import oracle
    from librett_amd import data
    nthreads = min(oracle.max_threads(), int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or oracle.max_threads())
    total_bytes = 0
    total_s = 0.0
    t_start = time.perf_counter()
    max_bytes = max(int(np.prod(d)) * e for d, _, e in CASES)
    src = data.uniform(max_bytes // 8 + 1, 8, 5).view(np.uint8)
    for dim, perm, elem in CASES:
        vol = int(np.prod(dim))
        x = src[:vol * elem].view(oracle.ELEM_DTYPES[elem])
        oracle.transpose(list(dim), list(perm), elem, x, variant="omp", nthreads=nthreads)  # warm-up
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            oracle.transpose(list(dim), list(perm), elem, x, variant="omp", nthreads=nthreads)
            ts.append(time.perf_counter() - t0)
        ts.sort()
        total_s += ts[len(ts) // 2]
        total_bytes += 2 * vol * elem
        log("cpu %-28s %.3f s" % (dim, ts[len(ts) // 2]))
    log("cpu baseline took %.1f s wall" % (time.perf_counter() - t_start))
    return {
        "value": round(total_bytes / total_s / 1e9, 2),
        "unit": "GB/s",
        "cores": nthreads,
        "kind": "port",
        "sample": "the 7 named cases at full size, median of %d OpenMP runs each after one warm-up" % reps,
    }


if __name__ == "__main__":
    main()
