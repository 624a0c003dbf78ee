#!/usr/bin/env python3
"""Benchmark: lineage-steps/s of the HIP Independent algorithm on one MI355X.

Headline workload ("config4-window"): the almost-infinite scenario with
Gaussian dispersal sigma=50, every lineage of a circle of radius 2000
(12,566,345 lineages), speciation 1e-6, Seed(42), delta_t 2, DedupCache::None
(so the event stream is bit-identical between the forms and the CPU path),
up to pause_before = T (default 2000: about 1000 events per lineage,
1.26e10 lineage-steps).  One bench step simulates the whole window from the
initial lineages with the persistent form (autotuned launch shape, each
water-level round one hipGraph); lineages are already resident in HBM when
the timed region starts.

The same JSON line also reports: the plain one-launch-per-slice form on the
same window (with a bit-identity check of the two event streams), the same
scenario with the reference's default dedup cache, config 2 (full run),
config 3 (128x128 map, the alias-gather workload, with its gather roofline),
serving (config 5: p50/p99 latency), the dominant kernel's roofline, and the
CPU oracle timed on the host's cores on a bounded sample of the headline
workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "necsim-rust_amd"))

import numpy as np  # noqa: E402

import necsim_hip as nh  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# bytes of lineage state one persistent round reads and writes per lineage
# (SoA: ref 8, t_last 8, x/y/index 12, next_t 8, next_state 8, flags 1)
LINEAGE_STATE_BYTES = 45
EVENT_BYTES = 40
CACHE_BUCKET_BYTES = 32


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_job(ctx, tables, cogs, lineages, args, seed, pause):
    job = nh.Job(ctx, tables, args, nh.WyHash.seed_from_u64(seed), cogs.speciation_probability,
                 cogs.turnover_rate, pause_before=pause)
    job.add_lineages(lineages)
    job.start()  # device allocation + upload, outside the timed region
    return job


def run_job(job):
    fin = False
    while not fin:
        fin = job.step_round()
        job.drain_events()
    o = job.outcome()
    ms, launches = job.kernel_time()
    return o, ms, launches


def timed_runs(ctx, tables, cogs, lineages, args, seed, pause, steps, warmup):
    for _ in range(warmup):
        j = make_job(ctx, tables, cogs, lineages, args, seed, pause)
        run_job(j)
        j.close()
    jobs = [make_job(ctx, tables, cogs, lineages, args, seed, pause) for _ in range(steps)]
    t0 = time.perf_counter()
    results = [run_job(j) for j in jobs]
    wall = time.perf_counter() - t0
    for j in jobs:
        j.close()
    return wall, results


def cpu_baseline(workload, pause, budget_s=20.0):
    """The oracle (C restatement of the reference CPU Independent algorithm)
    on the host's cores, IsolatedIndividuals partitions over a bounded sample
    of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
    if workload == "config4-window":
        scn = O.Scenario.almost_infinite_normal(50.0)
        circle = (2**31 - 1, 2**31 - 1, 2000)
        nu, seed, total = 1e-6, 42, 12566345
    else:
        raise ValueError(workload)
    params = O.make_params(seed, nu, pause_before=pause)
    # calibrate: a small sample first, then size the sample to the budget
    parts = 4096
    r = O.simulate_partitioned(scn, params, 1.0, threads, parts, circle=circle)
    rate = r["steps"] / max(r["seconds"], 1e-6)
    sample_lineages = total * threads / parts
    steps_per_lineage = r["steps"] / max(sample_lineages, 1)
    want_lineages = budget_s * rate / max(steps_per_lineage, 1e-9)
    parts = int(max(threads, min(parts, total * threads / max(want_lineages, 1))))
    r = O.simulate_partitioned(scn, params, 1.0, threads, parts, circle=circle)
    return {"value": r["steps"] / r["seconds"], "unit": "lineage-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{threads} of {parts} IsolatedIndividuals partitions of the same workload "
                      f"({r['steps']} steps, {r['seconds']:.1f} s, one thread per partition)"}


def config2_run(ctx):
    """config 2 (sigma 10, r 100, nu 1e-4, Seed(7)) to completion, persistent form"""
    cogs = nh.AlmostInfiniteNormalDispersalScenario(100, 10.0).new(1e-4)
    lin = cogs.sample_habitat(nh.OriginPreSampler.all())
    tables = nh.Tables(ctx, cogs)
    out = {}
    for name, dedup in (("dedup_relative_1", nh.DedupCache.Relative(1.0)), ("dedup_none", nh.DedupCache.none())):
        wall, res = timed_runs(ctx, tables, cogs, lin, nh.IndependentArguments(dedup_cache=dedup), 7, None, 1, 1)
        o, ms, k = res[0]
        out[name] = {"steps": int(o.steps), "ms": wall * 1e3, "value": o.steps / wall, "kernel_ms": ms,
                     "speciations": int(o.n_events), "rounds": int(o.rounds)}
    return {"lineages": int(len(lin)), **out}


GATHER_BYTES_PER_STEP = 48  # cell record 32 B + alias atom 16 B per dispersal step


def config3_run(ctx, window):
    """config 3 (128x128 map, nu 1e-5): the alias-gather workload"""
    h, rp, c, w = nh.generate_map(128, 128, seed=42)
    scn = nh.SpatiallyExplicitUniformTurnoverScenario(128, 128, h, rp, c, w)
    cogs = scn.new(1e-5)
    lin = cogs.sample_habitat(nh.OriginPreSampler.all())
    tables = nh.Tables(ctx, cogs)
    wall, res = timed_runs(ctx, tables, cogs, lin, nh.IndependentArguments(), 42, window, 1, 1)
    o, ms, k = res[0]
    achieved = o.steps * GATHER_BYTES_PER_STEP / (ms * 1e-3) / 1e9 if ms else 0.0
    return {"lineages": int(len(lin)), "window": window, "steps": int(o.steps), "ms": wall * 1e3,
            "value": o.steps / wall, "kernel_ms": ms, "atoms": int(len(scn.tables()[2])),
            "gather_roofline": {"bound": "hbm", "bytes_per_step": GATHER_BYTES_PER_STEP,
                                "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(achieved / HBM_PEAK_GBS, 5),
                                "roofline_steps_per_s": HBM_PEAK_GBS * 1e9 / GATHER_BYTES_PER_STEP}}


def serving_run(ctx, n_requests=150, rate=50.0, cap=1000.0, tick_steps=250, seed=1):
    """config 5: Poisson stream of simulations on a shared 256x256 map"""
    h, rp, c, w = nh.generate_map(256, 256, seed=42)
    scn = nh.SpatiallyExplicitUniformTurnoverScenario(256, 256, h, rp, c, w)
    base = scn.new(1e-3)
    tables = nh.Tables(ctx, base)
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    reqs = []
    for k in range(n_requests):
        nu = float(10 ** rng.uniform(-7, -3))
        pct = float(rng.uniform(0.01, 1.0))
        lin = base.sample_habitat(nh.OriginPreSampler.all().percentage(pct))
        reqs.append((nu, lin))
    srv = nh.Server(ctx, tables, max_lineages=1 << 24, hbm_budget_bytes=48 << 30, tick_steps=tick_steps)
    args = nh.IndependentArguments()
    t0 = time.perf_counter()
    submitted = 0
    lat, steps, done = [], 0, 0
    submit_t = {}
    while done < n_requests:
        now = time.perf_counter() - t0
        while submitted < n_requests and arrivals[submitted] <= now:
            nu, lin = reqs[submitted]
            rid = srv.submit(lin, nh.WyHash.seed_from_u64(submitted), nu, args, pause_before=cap)
            submit_t[rid] = t0 + arrivals[submitted]
            submitted += 1
        active, queued = srv.tick()
        while True:
            r = srv.poll()
            if r is None:
                break
            lat.append((time.perf_counter() - submit_t[r.request_id]) * 1e3)
            steps += int(r.steps)
            srv.release(r.request_id)
            done += 1
        if active == 0 and queued == 0 and submitted < n_requests:
            time.sleep(max(0.0, min(0.002, arrivals[submitted] - (time.perf_counter() - t0))))
    wall = time.perf_counter() - t0
    srv.close()
    lat = np.array(lat)
    return {"requests": n_requests, "arrival_rate_per_s": rate, "pause_before_cap": cap,
            "tick_steps": tick_steps, "wall_s": wall, "steps": steps, "value": steps / wall,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}


def run_collect(job):
    fin = False
    chunks = []
    while not fin:
        fin = job.step_round()
        chunks.append(job.drain_events())
    return np.concatenate(chunks) if chunks else np.zeros(0, nh.EVENT_DTYPE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--window", type=float, default=2000.0, help="pause_before of the headline workload")
    ap.add_argument("--dedup-window", type=float, default=20000.0,
                    help="pause_before of the dedup-on config 4 comparison")
    ap.add_argument("--no-extra", action="store_true", help="headline only")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="default launch shape unless given below")
    ap.add_argument("--form", choices=("persistent", "plain"), default="persistent",
                    help="form of the headline runs")
    ap.add_argument("--block-size", type=int, default=0, help="fixed launch shape (skips the autotuner)")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--inner-steps", type=int, default=0)
    ap.add_argument("--exec-fill", type=int, default=0)
    ap.add_argument("--graph", action="store_true",
                    help="capture each water-level round (kernel + event sort) in a hipGraph; the "
                         "kernel timing then covers the whole graph")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for {a.gpus})",
              file=sys.stderr)
        sys.exit(2)

    ctx = nh.Context.get(0)
    cus, arch = ctx.info()
    log(f"device {arch} with {cus} CUs")

    # ---- headline: config 4, dedup None window, persistent form ----
    t = time.perf_counter()
    cogs = nh.AlmostInfiniteNormalDispersalScenario(2000, 50.0).new(1e-6)
    lineages = cogs.sample_habitat(nh.OriginPreSampler.all())
    tables = nh.Tables(ctx, cogs)
    setup_s = time.perf_counter() - t
    log(f"config 4: {len(lineages)} lineages, setup {setup_s:.2f} s")
    form = nh.PERSISTENT if a.form == "persistent" else nh.PLAIN
    fixed = a.block_size or a.blocks_per_cu or a.inner_steps or a.exec_fill
    args = nh.IndependentArguments(dedup_cache=nh.DedupCache.none(), form=form,
                                   use_graph=a.graph and form == nh.PERSISTENT,
                                   block_size=a.block_size, blocks_per_cu=a.blocks_per_cu,
                                   inner_steps=a.inner_steps, exec_fill=a.exec_fill)
    tuned = None
    if not (a.no_tune or fixed or form == nh.PLAIN):
        t = time.perf_counter()
        targs, tms = nh.autotune(ctx, tables, args, nh.WyHash.seed_from_u64(42), cogs, lineages[::8], 50.0,
                                 tune_step_slice=False)
        args = nh.IndependentArguments(dedup_cache=args.dedup_cache, form=nh.PERSISTENT, use_graph=args.use_graph,
                                       block_size=targs.block_size, blocks_per_cu=targs.blocks_per_cu,
                                       inner_steps=targs.inner_steps, exec_fill=targs.exec_fill)
        tuned = {"block_size": targs.block_size, "blocks_per_cu": targs.blocks_per_cu,
                 "inner_steps": targs.inner_steps, "exec_fill": targs.exec_fill, "trial_ms": tms,
                 "tune_s": time.perf_counter() - t}
        log(f"autotuned: {tuned}")
    wall, results = timed_runs(ctx, tables, cogs, lineages, args, 42, a.window, a.steps, a.warmup)
    steps_total = sum(int(o.steps) for o, _, _ in results)
    value = steps_total / wall
    kern_ms = sum(ms for _, ms, _ in results)
    launches = sum(k for _, _, k in results)
    o0 = results[0][0]
    log(f"persistent: {steps_total} steps in {wall:.3f} s -> {value:.4e} steps/s; "
        f"device {kern_ms:.1f} ms over {launches} launches")

    # roofline of the dominant kernel (k_persistent, captured with the event
    # sort in one graph per round): algorithmic bytes per launch = the lineage
    # state read and written back once per round plus the speciation events
    nl = max(results[0][2], 1)
    per_launch_steps = o0.steps / nl
    per_launch_bytes = len(lineages) * 2 * LINEAGE_STATE_BYTES / nl + o0.n_events / nl * EVENT_BYTES
    mean_launch_ms = results[0][1] / nl
    achieved = per_launch_bytes / (mean_launch_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None,
                "kernel": "k_persistent<almost-infinite> (+ event sort, one hipGraph per round)",
                "mean_launch_ms": round(mean_launch_ms, 3),
                "algorithmic_bytes_per_launch": int(per_launch_bytes),
                "steps_per_launch": int(per_launch_steps),
                "note": "no table gathers on this workload: the kernel is bound by VALU issue "
                        "(64-bit integer hashing and f64 maths), not by HBM"}

    extra = {"autotuned": tuned,
             "launch": {"form": a.form, "block_size": args.block_size, "blocks_per_cu": args.blocks_per_cu,
                        "inner_steps": args.inner_steps, "exec_fill": args.exec_fill,
                        "use_graph": bool(args.use_graph)}}
    # plain copy kernel of known size: achievable HBM bandwidth and the
    # calibration of the FETCH_SIZE / WRITE_SIZE byte counters
    cms, cgbs = nh.copy_calibrate(ctx, 1 << 30, 10)
    extra["copy_kernel"] = {"kernel": "k_copy_calibrate", "bytes_read": 1 << 30, "bytes_written": 1 << 30,
                            "ms": cms, "GB_per_s": cgbs}
    log(f"copy kernel: {cgbs:.1f} GB/s")
    if not a.no_extra:
        # bit-exact comparison with the plain one-launch-per-slice form
        pj = make_job(ctx, tables, cogs, lineages, nh.IndependentArguments(dedup_cache=nh.DedupCache.none(),
                                                                          form=nh.PLAIN), 42, a.window)
        t = time.perf_counter()
        pev = run_collect(pj)
        pwall = time.perf_counter() - t
        po = pj.outcome()
        pms, pk = pj.kernel_time()
        pj.close()
        qj = make_job(ctx, tables, cogs, lineages, args, 42, a.window)
        qev = run_collect(qj)
        qo = qj.outcome()
        qj.close()
        same = pev.tobytes() == qev.tobytes() and po.steps == qo.steps and po.time == qo.time
        extra["plain_form"] = {"value": po.steps / pwall, "ms": pwall * 1e3, "steps": int(po.steps),
                               "launches": int(pk), "kernel_ms": pms,
                               "events_bit_identical_to_persistent": bool(same)}
        extra["persistent_over_plain"] = value / (po.steps / pwall)
        log(f"plain: {po.steps} steps in {pwall:.3f} s -> {po.steps / pwall:.4e} steps/s; identical={same}")

        # config 4 with the reference's default dedup cache (Relative(1.0)):
        # coalesced lineages are dropped, the window is mostly a coalescence
        # workload; plain form = the reference's queue order, bit-exact
        dd = {}
        for name, form in (("persistent", nh.PERSISTENT), ("plain", nh.PLAIN)):
            dargs = nh.IndependentArguments(form=form, use_graph=form == nh.PERSISTENT)
            dwall, dres = timed_runs(ctx, tables, cogs, lineages, dargs, 42, a.dedup_window, 1, 1)
            do = dres[0][0]
            dd[name] = {"value": do.steps / dwall, "ms": dwall * 1e3, "steps": int(do.steps),
                        "kernel_ms": dres[0][1], "launches": int(dres[0][2]),
                        "remaining_lineages": int(do.n_remaining), "speciations": int(do.n_events)}
        dd["persistent_over_plain_wall"] = dd["plain"]["ms"] / dd["persistent"]["ms"]
        extra["config4_dedup_window"] = {"pause_before": a.dedup_window, **dd}
        log(f"config 4 dedup window: {extra['config4_dedup_window']}")
        del tables
        extra["config2_full_run"] = config2_run(ctx)
        log(f"config 2: {extra['config2_full_run']}")
        extra["config3_window"] = config3_run(ctx, 2000.0)
        log(f"config 3: {extra['config3_window']}")
        extra["serving_config5"] = serving_run(ctx)
        log(f"serving: {extra['serving_config5']}")

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline("config4-window", a.window)
        log(f"cpu baseline: {cpu['value']:.4e} steps/s on {cpu['cores']} cores")

    line = {
        "metric": "lineage-steps/sec on 1 MI355X vs CPU Independent; p50/p99 sweep latency",
        "value": value, "unit": "lineage-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64+u64", "data": "synthetic",
        "config": {"workload": "config4-window: almost-infinite normal sigma=50, circle r=2000 "
                               f"(12,566,345 lineages), nu=1e-6, Seed(42), pause_before={a.window}, "
                               "DedupCache::None, persistent form (hipGraph rounds, autotuned launch), "
                               "delta_t 2, event_slice 2.0",
                   "lineages": int(len(lineages)), "steps_per_run": int(o0.steps),
                   "speciations_per_run": int(o0.n_events), "rounds": int(o0.rounds),
                   "remaining_lineages": int(o0.n_remaining)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
