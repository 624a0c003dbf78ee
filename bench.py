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
import dataclasses
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "necsim-rust_amd"))

import numpy as np  # noqa: E402

import necsim_hip as nh  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# MI355X VALU issue: 4 SIMDs per CU, each retiring one full-rate wave64
# vector instruction per 4 cycles (16 lanes per cycle) = one wave
# instruction per CU per cycle at the 2.4 GHz peak engine clock (the same
# rate as the 78.6 TFLOP/s f64 vector FMA peak: 64 lanes x 2 flops per CU-cycle)
ENGINE_CLOCK_HZ = 2.4e9
# committed rocprofv3 summaries of the headline window: the pooled kernel and
# the one-lineage-per-lane kernel (the roofline uses the one of the kernel run)
PROFILE_SUMMARIES = {"pooled": os.path.join(ROOT, "profiles", "k_pooled_config4_window.txt"),
                     "lane": os.path.join(ROOT, "profiles", "k_persistent_config4_window.txt")}
GATHER_PROFILE = os.path.join(ROOT, "profiles", "k_persistent_alias_gather_map256.txt")
# bytes of lineage state one persistent round reads and writes per lineage
# (SoA: ref 8, t_last 8, x/y/index 12, next_t 8, next_state 8, flags 1)
LINEAGE_STATE_BYTES = 45
MOVING_STATE_BYTES = LINEAGE_STATE_BYTES - 8  # all but the reference, which never changes
EVENT_BYTES = 40
# lineages per wave of the pooled kernel in the dispersal-reporting rounds
# (128 and 192 measured equal)
DISPERSAL_POOL = 128
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


def host_threads():
    """Threads of the CPU baseline: the CPUs this process may run on, capped
    at the host's CPU share per GPU where OMP_NUM_THREADS states it (16 per
    GPU on the MI355X hosts, whatever os.cpu_count() shows)."""
    n = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(n, share) if share > 0 else n), n


def cpu_baseline(workload, pause, budget_s=20.0, single_budget_s=8.0):
    """The oracle (C restatement of the reference CPU Independent algorithm,
    gcc -O3 -ffp-contract=off) on the host's cores: IsolatedIndividuals
    partitions over a bounded sample of the same workload, one thread per
    partition, on every thread of the host's share; and one thread alone on
    a smaller sample (the measured per-core rate)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads, visible = host_threads()
    if workload == "config4-window":
        scn = O.Scenario.almost_infinite_normal(50.0)
        circle = (2**31 - 1, 2**31 - 1, 2000)
        nu, seed, total = 1e-6, 42, 12566345
    else:
        raise ValueError(workload)
    params = O.make_params(seed, nu, pause_before=pause)

    def sized_run(nthreads, budget):
        # calibrate on a small sample, then size the sample to the budget
        parts = 4096
        r = O.simulate_partitioned(scn, params, 1.0, nthreads, parts, circle=circle)
        rate = r["steps"] / max(r["seconds"], 1e-6)
        steps_per_lineage = r["steps"] / max(total * nthreads / parts, 1)
        want_lineages = budget * rate / max(steps_per_lineage, 1e-9)
        parts = int(max(nthreads, min(1 << 20, total * nthreads / max(want_lineages, 1))))
        r = O.simulate_partitioned(scn, params, 1.0, nthreads, parts, circle=circle)
        return r, parts

    r, parts = sized_run(threads, budget_s)
    r1, parts1 = sized_run(1, single_budget_s)
    value, single = r["steps"] / r["seconds"], r1["steps"] / r1["seconds"]
    return {"value": value, "unit": "lineage-steps/s", "cores": threads, "kind": "port",
            "single_thread_value": single, "all_threads_over_single": value / single,
            "host_cpus": {"os_cpu_count": os.cpu_count(), "affinity": visible,
                          "omp_num_threads": os.environ.get("OMP_NUM_THREADS")},
            "sample": f"{threads} of {parts} IsolatedIndividuals partitions of the same workload on "
                      f"{threads} threads ({r['steps']} steps, {r['seconds']:.1f} s); single thread: 1 of "
                      f"{parts1} partitions ({r1['steps']} steps, {r1['seconds']:.1f} s)"}


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
# the alias-gather window's launch shape, the autotuner's choice on it
# (nh.autotune over windows of 100 and 500: execution once an eighth of a
# wave is ready, 13.4 ms against 14.3 ms at the default five eighths); fixed
# here so that its profile holds the window's launches only
GATHER_SHAPE = dict(block_size=256, blocks_per_cu=4, inner_steps=16, exec_fill=1)


def map_window_run(ctx, size, nu, window, dedup, seed=42, tune=False, shape=None):
    """A window of the spatially explicit map scenario (the generator of
    configs 3 and 5): every dispersal gathers the target cell's record and
    one alias atom.  Returns timings and the gather-bound roofline.  With
    tune, the launch shape is chosen by untimed runs of the same window; with
    shape (IndependentArguments fields), that launch shape is used."""
    h, rp, c, w = nh.generate_map(size, size, seed=42)
    scn = nh.SpatiallyExplicitUniformTurnoverScenario(size, size, h, rp, c, w)
    # resident tables: alias build on the host + upload, or on the GPU
    t = time.perf_counter()
    host_tables = nh.Tables(ctx, scn.new(nu))
    host_s = time.perf_counter() - t
    del host_tables
    t = time.perf_counter()
    cogs = scn.new(nu, host_tables=False)
    tables = nh.Tables(ctx, cogs)
    device_s = time.perf_counter() - t
    lin = cogs.sample_habitat(nh.OriginPreSampler.all())
    atoms = int(len(scn.tables()[2]))
    args = nh.IndependentArguments(dedup_cache=dedup, **(shape or {}))
    tuned = dict(shape) if shape else None
    if tune:
        # launch shape chosen on the window itself: one untimed run with the
        # default shape and one per residency of 256-thread blocks.  These
        # windows hold 0.4-1.6 lineages per lane and end in serial chains,
        # which can run faster with fewer resident waves per SIMD.
        t = time.perf_counter()
        trials = {"default": args}
        for bpc in (1, 2, 3, 4):
            trials[f"256x{bpc}"] = dataclasses.replace(args, block_size=256, blocks_per_cu=bpc)
        trial_ms = {}
        for name, targs in trials.items():
            tj = make_job(ctx, tables, cogs, lin, targs, seed, window)
            tt = time.perf_counter()
            run_collect(tj)
            trial_ms[name] = round((time.perf_counter() - tt) * 1e3, 3)
            tj.close()
        best = min(trial_ms, key=trial_ms.get)
        args = trials[best]
        tuned = {"chosen": best, "trial_ms": trial_ms, "tune_s": round(time.perf_counter() - t, 3)}
    wall, res = timed_runs(ctx, tables, cogs, lin, args, seed, window, 1, 1)
    o, ms, k = res[0]
    achieved = o.steps * GATHER_BYTES_PER_STEP / (ms * 1e-3) / 1e9 if ms else 0.0
    bound = HBM_PEAK_GBS * 1e9 / GATHER_BYTES_PER_STEP
    return {"map": f"{size}x{size}", "lineages": int(len(lin)), "window": window, "speciation": nu,
            "dedup": "none" if dedup.mode == 0 else "relative 1.0", "launch_shape": tuned or "default",
            "steps": int(o.steps), "ms": wall * 1e3,
            "value": o.steps / wall, "kernel_ms": ms, "launches": int(k), "atoms": atoms,
            "table_bytes": atoms * 16 + size * size * 32,
            "tables_build_s": {"host_alias_build_and_upload": host_s, "device_alias_build": device_s},
            "gather_roofline": {"bound": "hbm", "bytes_per_step": GATHER_BYTES_PER_STEP,
                                "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(achieved / HBM_PEAK_GBS, 5), "roofline_steps_per_s": bound,
                                "steps_per_s_frac_of_roofline": round(o.steps / (ms * 1e-3) / bound, 5)}}


def chain_latency_run(ctx, window=20000.0):
    """The latency of dependent steps: lineages that never speciate in the
    window (almost-infinite sigma 50, nu 1e-9) run window / 2 steps each; one
    lineage, and 64 (which sparse rounds spread one per wave), per form."""
    out = {"window": window}
    for n_lin, radius in ((1, 0), (64, 5)):
        cogs = nh.AlmostInfiniteNormalDispersalScenario(radius, 50.0).new(1e-9)
        tables = nh.Tables(ctx, cogs)
        lin = cogs.sample_habitat(nh.OriginPreSampler.all())[:n_lin]
        for name, kw in (("persistent", {}), ("persistent_64_thread_blocks", dict(block_size=64, blocks_per_cu=1)),
                         ("plain_batched", dict(form=nh.PLAIN)), ("plain", dict(form=nh.PLAIN, plain_batch=1))):
            args = nh.IndependentArguments(dedup_cache=nh.DedupCache.none(),
                                           event_slice=nh.EventSlice.Absolute(10**9), **kw)
            job = make_job(ctx, tables, cogs, lin, args, 5, window)
            t = time.perf_counter()
            run_collect(job)
            wall = time.perf_counter() - t
            o = job.outcome()
            job.close()
            per = o.steps / n_lin
            out[f"{n_lin}_lineages_{name}"] = {"steps": int(o.steps), "ms": wall * 1e3,
                                               "us_per_dependent_step": wall / per * 1e6}
    return out


def config3_run(ctx, window):
    """config 3 (128x128 map, nu 1e-5, the reference's default dedup)"""
    return map_window_run(ctx, 128, 1e-5, window, nh.DedupCache.Relative(1.0), tune=True)


def gather_run(ctx, window):
    """the alias-gather workload: 256x256 map (config 5's generator; its
    tables fill the 256 MB last-level cache), no dedup.  Adds the HBM
    figures of the committed rocprofv3 summary of the same window."""
    r = map_window_run(ctx, 256, 1e-5, window, nh.DedupCache.none(), shape=GATHER_SHAPE)
    prof = read_profile_summary(GATHER_PROFILE)
    if prof.get("window") == window:
        r["rocprof"] = {k: prof.get(k) for k in ("kernel_mean_ms", "hbm_bytes_per_launch", "hbm_rate_GBps",
                                                  "hbm_frac_of_peak", "fetch_over_gather_bytes",
                                                  "valu_issue_fraction", "valu_lane_utilization")}
        r["rocprof"]["source"] = os.path.relpath(GATHER_PROFILE, ROOT)
    return r


def serving_run(ctx, n_requests=150, rate=50.0, cap=1000.0, tick_steps=1000, seed=1):
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
    srv = nh.Server(ctx, tables, max_lineages=1 << 24, hbm_budget_bytes=8 << 30, tick_steps=tick_steps)
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


PACKED_EVENT_BYTES = 56


def dispersal_run(ctx, tables, cogs, lineages, args, window, drain_window, seed=42, wide_slice=20.0):
    """Dispersal reporting (a reporter with ReportDispersal) on the headline
    workload: every step is a 56-byte PackedEvent, the water level advances by
    event_slice / sum(lambda), and each round's events are sorted on the
    device.  Three ways to consume them: the speciation-only run as the
    yardstick, the device EventCounterReporter (events stay in HBM; each
    round's step kernel is one hipGraph, then the sort to a permutation and
    the count through it), and draining every sorted event to the host
    (bounded by the device-to-host copy)."""
    import dataclasses
    gargs = dataclasses.replace(args, use_graph=True)
    # the reference CUDA algorithm's default event slice, Relative(20.0)
    # (cuda/src/arguments.rs:124-138): ten times fewer, larger rounds (the
    # sort buffers for 2 x 20 x 12.6M events take about 70 GB of HBM)
    wide = dataclasses.replace(gargs, event_slice=nh.EventSlice.Relative(wide_slice or 20.0))
    out = {"window": window, "host_drain_window": drain_window, "event_slice": "relative 2.0"}
    modes = [("speciation_only", window, gargs, {}),
             ("device_counter", window, gargs, dict(report_dispersal=True, device_event_counter=True))]
    if wide_slice:
        out["device_counter_event_slice_20"] = {"event_slice": f"relative {wide_slice}"}
        modes.append(("device_counter_event_slice_20", window, wide,
                      dict(report_dispersal=True, device_event_counter=True)))
    modes.append(("host_drain", drain_window, gargs, dict(report_dispersal=True)))
    for name, win, margs, kw in modes:
        job = nh.Job(ctx, tables, margs, nh.WyHash.seed_from_u64(seed), cogs.speciation_probability,
                     cogs.turnover_rate, pause_before=win, **kw)
        job.add_lineages(lineages)
        job.start()
        t0 = time.perf_counter()
        fin, drained, species = False, 0, set()
        while not fin:
            fin = job.step_round()
            if name == "host_drain":
                drained += len(job.drain_packed_events())
            elif name == "speciation_only":
                species |= species_of(job.drain_events())
        wall = time.perf_counter() - t0
        o = job.outcome()
        ms, k = job.kernel_time()
        r = {"steps": int(o.steps), "events": int(o.n_events), "rounds": int(o.rounds), "wall_ms": wall * 1e3,
             "device_ms": ms, "steps_per_s": o.steps / wall, "device_steps_per_s": o.steps / (ms * 1e-3)}
        if name == "speciation_only":
            # coalesced lineages (dedup None) repeat their speciations; the
            # counter counts each distinct one once (event_counter.rs:57-72)
            r["distinct_speciations"] = len(species)
        if name.startswith("device_counter"):
            r["counts"] = job.event_counts()
            r["counts_raw_total_equals_steps"] = r["counts"]["raw_total"] == o.steps
        if name == "host_drain":
            r["drained_events"] = drained
            r["drained_GB_per_s"] = drained * PACKED_EVENT_BYTES / wall / 1e9
        job.close()
        out[name] = {**out.get(name, {}), **r}
        log(f"dispersal {name}: {r}")
    sp, dc = out["speciation_only"], out["device_counter"]
    if wide_slice:
        out["device_counter_event_slice_20"]["counts_equal_event_slice_2"] = \
            out["device_counter_event_slice_20"]["counts"] == dc["counts"]
    out["device_counter_species_equal_speciation_only"] = dc["counts"]["speciation"] == sp["distinct_speciations"]
    out["device_counter_over_speciation_only"] = dc["steps_per_s"] / sp["steps_per_s"]
    return out


def task_list_run(ctx, tables, cogs, lineages, pargs, window, seed=42):
    """The persistent form against the reference CUDA algorithm's launch
    structure (cuda/src/parallelisation/monolithic.rs:222-318, rebuilt on
    this library's kernels as the task-list form): a task list of grid x block
    lineages per launch, uploaded and downloaded around it, dedup cache and
    slow/fast queues on the host; its defaults are 64 x 64 tasks and a step
    slice of 150 (cuda/src/arguments.rs:124-138).  DedupCache::None, where
    every form's event stream is bit-exact, on a window short enough for the
    task-list form to finish in seconds."""
    out = {"window": window}
    streams = {}
    shapes = (("persistent", pargs),
              ("task_list_64x64", nh.IndependentArguments(dedup_cache=nh.DedupCache.none(), form=nh.TASK_LIST,
                                                          step_slice=150, block_size=64, task_list_size=4096)),
              ("task_list_gpu_filling", nh.IndependentArguments(dedup_cache=nh.DedupCache.none(), form=nh.TASK_LIST,
                                                                step_slice=150, block_size=256,
                                                                task_list_size=256 * 1024)))
    for name, args in shapes:
        job = make_job(ctx, tables, cogs, lineages, args, seed, window)
        t = time.perf_counter()
        ev = run_collect(job)
        wall = time.perf_counter() - t
        o = job.outcome()
        ms, k = job.kernel_time()
        job.close()
        streams[name] = (ev.tobytes(), int(o.steps), o.time)
        out[name] = {"value": o.steps / wall, "ms": wall * 1e3, "steps": int(o.steps), "launches": int(k),
                     "kernel_ms": ms}
        log(f"task-list comparison {name}: {out[name]}")
    for name in ("task_list_64x64", "task_list_gpu_filling"):
        out[f"persistent_over_{name}"] = out["persistent"]["value"] / out[name]["value"]
        out[f"{name}_events_bit_identical_to_persistent"] = streams[name] == streams["persistent"]
    return out


def species_of(ev):
    """a species is a distinct (time, origin) speciation event (SpeciationEvent
    equality, necsim/core/src/event.rs:301-307, as counted by
    plugins/common/src/biodiversity.rs:32-41)"""
    return set(zip(ev["event_time"].tolist(), ev["x"].tolist(), ev["y"].tolist(), ev["index"].tolist()))


def dedup_window_run(ctx, tables, cogs, lineages, dedup_window):
    """config 4 with the reference's default dedup cache (Relative(1.0)) up
    to pause_before = dedup_window: the persistent form (autotuned, then the
    residency chosen on the window) against the plain form, one host round
    trip per pass and batched.  Coalesced lineages are dropped, so the window
    is mostly a coalescence workload; the plain form keeps the reference's
    queue order, bit-exact.  Returns the results and the tuned arguments."""
    dd = {}
    t = time.perf_counter()
    dtuned, dtms = nh.autotune(ctx, tables, nh.IndependentArguments(), nh.WyHash.seed_from_u64(42), cogs,
                               lineages[::8], 200.0, tune_step_slice=True)
    # Second tuning stage on the window itself: the sample trial is
    # throughput bound and favours full residency, but this window ends
    # in long serial chains (the population collapses from 12.6M to 15k
    # lineages), which run faster with fewer resident waves per SIMD.
    # One untimed run per residency of 256-thread blocks picks it.
    residency = {}
    for bpc in (1, 2, 3):
        targs = dataclasses.replace(dtuned, block_size=256, blocks_per_cu=bpc)
        tj = make_job(ctx, tables, cogs, lineages, targs, 42, dedup_window)
        tt = time.perf_counter()
        run_collect(tj)
        residency[bpc] = round((time.perf_counter() - tt) * 1e3, 2)
        tj.close()
    best_bpc = min(residency, key=residency.get)
    dtuned = dataclasses.replace(dtuned, block_size=256, blocks_per_cu=best_bpc)
    dd["autotuned"] = {"step_slice": dtuned.step_slice, "block_size": dtuned.block_size,
                       "blocks_per_cu": dtuned.blocks_per_cu, "inner_steps": dtuned.inner_steps,
                       "exec_fill": dtuned.exec_fill, "residency_trial_ms": residency,
                       "tune_s": time.perf_counter() - t}
    species, streams = {}, {}
    # plain: one host round trip per pass (the reference's launch structure);
    # plain_batched: the same passes, 16 per host synchronisation
    for name, dargs in (("persistent", dtuned), ("plain", nh.IndependentArguments(form=nh.PLAIN, plain_batch=1)),
                        ("plain_batched", nh.IndependentArguments(form=nh.PLAIN))):
        dj = make_job(ctx, tables, cogs, lineages, dargs, 42, dedup_window)
        t = time.perf_counter()
        dev = run_collect(dj)
        dwall = time.perf_counter() - t
        do = dj.outcome()
        dms, dk = dj.kernel_time()
        dj.close()
        # duplicate lineages that a lossy cache lets through repeat a species
        species[name] = species_of(dev)
        streams[name] = (dev.tobytes(), int(do.steps), do.time)
        dd[name] = {"value": do.steps / dwall, "ms": dwall * 1e3, "steps": int(do.steps),
                    "kernel_ms": dms, "launches": int(dk), "remaining_lineages": int(do.n_remaining),
                    "speciation_events": int(do.n_events), "species": len(species[name])}
    dd["species_sets_identical"] = species["persistent"] == species["plain"] == species["plain_batched"]
    dd["persistent_over_plain_wall"] = dd["plain"]["ms"] / dd["persistent"]["ms"]
    # bit-exact with dedup on: the batched plain form keeps the reference's queue order
    dd["plain_batched_events_bit_identical_to_plain"] = streams["plain_batched"] == streams["plain"]
    dd["plain_batched_over_plain_wall"] = dd["plain"]["ms"] / dd["plain_batched"]["ms"]
    return {"pause_before": dedup_window, **dd}, dtuned


def config4_tail_run(ctx, tables, cogs, lineages, dtuned, dedup_window, window=2e5, completion=True, seed=42):
    """config 4's heavy-tailed lineage lifetimes (mean 1/(lambda nu) = 2e6
    time units) under the reference's default dedup cache: a window 10x the
    dedup window, and the whole simulation to completion.  The persistent
    form (tuned on the dedup window; each water-level round's step / sort /
    compaction cycle replayed from a hipGraph) against the plain form with one
    host round trip per pass and batched; the phase between the dedup window
    and this one is the difference of the two runs of each form."""
    out = {"window": window}
    gtuned = dataclasses.replace(dtuned, use_graph=True)
    runs = [("window", window, (("persistent", gtuned),
                                ("plain", nh.IndependentArguments(form=nh.PLAIN, plain_batch=1)),
                                ("plain_batched", nh.IndependentArguments(form=nh.PLAIN))))]
    if completion:
        runs.append(("completion", None, (("persistent", gtuned),)))
    for label, pause, forms in runs:
        r, species = {"pause_before": pause}, {}
        for name, fargs in forms:
            job = make_job(ctx, tables, cogs, lineages, fargs, seed, pause)
            t = time.perf_counter()
            ev = run_collect(job, progress=f"config 4 tail {label} {name}")
            wall = time.perf_counter() - t
            o = job.outcome()
            ms, k = job.kernel_time()
            job.close()
            species[name] = species_of(ev)
            r[name] = {"value": o.steps / wall, "ms": wall * 1e3, "steps": int(o.steps), "kernel_ms": ms,
                       "launches": int(k), "rounds": int(o.rounds), "time": o.time,
                       "remaining_lineages": int(o.n_remaining), "speciation_events": int(o.n_events),
                       "species": len(species[name])}
            log(f"config 4 tail {label} {name}: {r[name]}")
        if len(forms) > 1:
            r["species_sets_identical"] = len({frozenset(v) for v in species.values()}) == 1
            for name in ("plain", "plain_batched"):
                r[f"persistent_over_{name}_wall"] = r[name]["ms"] / r["persistent"]["ms"]
        out[label] = r
    # the tail phase (dedup window .. window) of each form
    w = out["window"]
    phase = {}
    for name in ("persistent", "plain", "plain_batched"):
        a, b = dedup_window[name], w[name]
        dsteps, dms = b["steps"] - a["steps"], b["ms"] - a["ms"]
        phase[name] = {"steps": dsteps, "ms": dms, "value": dsteps / (dms * 1e-3) if dms > 0 else None}
    for name in ("plain", "plain_batched"):
        phase[f"persistent_over_{name}_wall"] = phase[name]["ms"] / phase["persistent"]["ms"]
    out[f"phase_{dedup_window['pause_before']:g}_to_{window:g}"] = phase
    return out


def run_collect(job, progress=None):
    fin = False
    chunks = []
    while not fin:
        fin = job.step_round()
        chunks.append(job.drain_events())
        if progress:
            o = job.outcome()
            log(f"{progress}: round {o.rounds}, {o.n_remaining} lineages left, time {o.time:.4g}")
    return np.concatenate(chunks) if chunks else np.zeros(0, nh.EVENT_DTYPE)


def read_profile_summary(path):
    """key: value lines of a committed rocprofv3 summary (numbers where they parse)"""
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        if ":" not in line or line.startswith((" ", "#", "|", "-")):
            continue
        k, v = line.split(":", 1)
        v = v.strip()
        try:
            out[k.strip()] = float(v)
        except ValueError:
            out[k.strip()] = v
    return out


def summarize(dirs):
    """Summary of rocprofv3 output directories (kernel trace --stats and
    --pmc passes): per kernel, dispatches, mean duration and counters per
    dispatch.  Prints text for profiles/."""
    import csv
    import glob
    from collections import defaultdict
    for d in dirs:
        for f in sorted(glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)):
            print(f"## kernel stats ({os.path.relpath(f, d)})")
            print("| kernel | calls | mean ms | total ms |")
            print("|---|---|---|---|")
            for r in csv.DictReader(open(f)):
                print(f"| {r['Name'][:90]} | {r['Calls']} | {float(r['AverageNs']) / 1e6:.4f} | "
                      f"{float(r['TotalDurationNs']) / 1e6:.3f} |")
        for f in sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)):
            per = defaultdict(lambda: defaultdict(float))
            names = {}
            for r in csv.DictReader(open(f)):
                key = int(r["Dispatch_Id"])
                names[key] = r["Kernel_Name"]
                per[key][r["Counter_Name"]] += float(r["Counter_Value"])
            agg = defaultdict(lambda: defaultdict(float))
            calls = defaultdict(int)
            for key, cs in per.items():
                calls[names[key]] += 1
                for c, v in cs.items():
                    agg[names[key]][c] += v
            print(f"## counters per dispatch ({os.path.relpath(f, d)})")
            for kname in sorted(agg, key=lambda k: -max(agg[k].values())):
                cs = ", ".join(f"{c}={v / calls[kname]:.6g}" for c, v in sorted(agg[kname].items()))
                print(f"- {kname[:90]} (x{calls[kname]}): {cs}")


def _kt_stats(d):
    import csv
    import glob
    f = glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)[0]
    return {r["Name"]: r for r in csv.DictReader(open(f))}


def _pmc_means(d):
    """per kernel name: each counter's mean over its dispatches"""
    import csv
    import glob
    from collections import defaultdict
    per, names = defaultdict(lambda: defaultdict(float)), {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            k = int(r["Dispatch_Id"])
            per[k][r["Counter_Name"]] += float(r["Counter_Value"])
            names[k] = r["Kernel_Name"]
    agg = defaultdict(lambda: defaultdict(list))
    for k, cs in per.items():
        for c, v in cs.items():
            agg[names[k]][c].append(v)
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in agg.items()}


def _of(table, pattern, counter):
    return [v for k, v in table.items() if pattern in k][0][counter]


def derive(kind, log, dirs):
    """The derived figures of a committed profile, from the bench output
    `log` and the rocprofv3 directories `dirs` (kernel trace, FETCH_SIZE,
    WRITE_SIZE, then for headline and gather the SQ pass and for headline
    the instruction-mix pass).  Byte counters are scaled by what they read
    for k_copy_calibrate (1 GiB read + 1 GiB written per dispatch)."""
    kt, fe, wr = _kt_stats(dirs[0]), _pmc_means(dirs[1]), _pmc_means(dirs[2])
    ff = (1 << 30) / (_of(fe, "k_copy_calibrate", "FETCH_SIZE") * 1024)
    wf = (1 << 30) / (_of(wr, "k_copy_calibrate", "WRITE_SIZE") * 1024)
    lines = [l for l in open(log).read().splitlines() if l.startswith("{")]
    bj = json.loads(lines[-1])
    out = {}
    if kind in ("headline", "gather"):
        sq = _pmc_means(dirs[3])
        if kind == "headline":
            la = bj["extra"]["launch"]
            K = f"k_pooled<1, {la['pool']}, false>" if la["pool"] else "k_persistent<1, false, false, false>"
            steps = bj["roofline"]["lineage_steps_per_launch"]
        else:
            K = "k_persistent<2, false, false, false>"
            steps = bj["gather_window"]["steps"]
        mean_ms = float([r for n, r in kt.items() if K in n][0]["AverageNs"]) / 1e6
        fb, wb = _of(fe, K, "FETCH_SIZE") * 1024 * ff, _of(wr, K, "WRITE_SIZE") * 1024 * wf
        valu, grbm = _of(sq, K, "SQ_INSTS_VALU"), _of(sq, K, "GRBM_GUI_ACTIVE")
        sec = mean_ms * 1e-3
        out["window"] = 2000
        if kind == "headline":
            out.update({"form": la["form"], "block_size": la["block_size"], "blocks_per_cu": la["blocks_per_cu"],
                        "inner_steps": la["inner_steps"], "exec_fill": la["exec_fill"], "pool": la["pool"]})
        out["kernel_mean_ms"] = round(mean_ms, 4)
        out["lineage_steps_per_launch"] = steps
        if kind == "headline":
            algo = bj["roofline"]["hbm"]["algorithmic_bytes_per_launch"]
            out.update({"algorithmic_bytes_per_launch": algo, "hbm_fetch_bytes_per_launch": round(fb),
                        "hbm_write_bytes_per_launch": round(wb), "hbm_bytes_per_launch": round(fb + wb),
                        "traffic_over_algorithmic": round((fb + wb) / algo, 3)})
        else:
            gb = steps * GATHER_BYTES_PER_STEP
            out.update({"gather_bytes_per_launch": gb, "gather_rate_GBps": round(gb / sec / 1e9, 1),
                        "gather_frac_of_hbm_peak": round(gb / sec / (HBM_PEAK_GBS * 1e9), 4),
                        "hbm_fetch_bytes_per_launch": round(fb), "hbm_write_bytes_per_launch": round(wb),
                        "hbm_bytes_per_launch": round(fb + wb), "hbm_rate_GBps": round((fb + wb) / sec / 1e9, 1),
                        "hbm_frac_of_peak": round((fb + wb) / sec / (HBM_PEAK_GBS * 1e9), 4),
                        "fetch_over_gather_bytes": round(fb / gb, 3),
                        "lineage_steps_per_s": float(f"{steps / sec:.4e}"),
                        "gather_roofline_steps_per_s": float(f"{HBM_PEAK_GBS * 1e9 / GATHER_BYTES_PER_STEP:.4e}")})
        out["valu_wave_instructions_per_lineage_step"] = round(valu / steps, 3)
        out["valu_issue_fraction"] = round(valu / (grbm / 8 * 256), 4)
        out["valu_lane_utilization"] = round(_of(sq, K, "SQ_THREAD_CYCLES_VALU") /
                                             (64 * _of(sq, K, "SQ_ACTIVE_INST_VALU")), 4)
        if kind == "headline":
            mix = _pmc_means(dirs[4])
            out.update({
                "valu_int64_per_step": round(_of(mix, K, "SQ_INSTS_VALU_INT64") / steps, 3),
                "valu_int32_per_step": round(_of(mix, K, "SQ_INSTS_VALU_INT32") / steps, 3),
                "valu_f64_add_mul_fma_per_step": round(sum(_of(mix, K, c) for c in (
                    "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_ADD_F64")) / steps, 3),
                "valu_f64_trans_per_step": round(_of(mix, K, "SQ_INSTS_VALU_TRANS_F64") / steps, 4),
                "valu_cvt_per_step": round(_of(mix, K, "SQ_INSTS_VALU_CVT") / steps, 3),
                "vmem_per_step": round(_of(mix, K, "SQ_INSTS_VMEM") / steps, 4),
                "lds_per_step": round(_of(sq, K, "SQ_INSTS_LDS") / steps, 3)})
        for k, v in out.items():
            print(f"{k}: {v}")
        return
    # dispersal: one counted round of the device-counter mode, per kernel
    dc = bj["dispersal_windows"]
    counted, drained = dc["device_counter"]["rounds"], dc["host_drain"]["rounds"]
    ev_round = dc["device_counter"]["steps"] / counted
    sorted_rounds = counted + drained
    rows = [("k_pooled<1, 128, true>", "step kernel (pooled), every step a 56 B PackedEvent + its sort key and slot",
             sorted_rounds, True),
            ("k_sradix_hist", "staged radix: digit histogram per 4096-key tile (per pass)", sorted_rounds, True),
            ("k_hist_chunk_sums<256, 4096>", "histogram scan, phase 1 (per pass)", sorted_rounds, True),
            ("k_scan_chunk_sums<256, 4096>", "histogram scan, phase 2 (per pass)", sorted_rounds, True),
            ("k_hist_scan_apply<256, 4096>", "histogram scan, phase 3 (per pass)", sorted_rounds, True),
            ("k_sradix_scatter", "staged radix: tile ranked in LDS, written out in digit order (per pass)",
             sorted_rounds, True),
            ("k_tie_runs<nh::DevPacked, false>", "windows of a counted round: each record read once, runs counted, "
             "only events whose prior time differs from their run's first sorted", counted, True),
            ("k_tie_runs<nh::DevPacked, true>", "(drained rounds only, not in the total) windows ordering runs by "
             "(prior, reference)", drained, False)]
    passes = int([r for n, r in kt.items() if "k_sradix_scatter" in n][0]["Calls"]) / sorted_rounds
    head = {"events_per_round": round(ev_round), "radix_digit_passes_per_round": f"{passes:.4f}",
            "device_counter_events_per_s": f"{dc['device_counter']['steps_per_s']:.4e}",
            "device_counter_round_ms": f"{dc['device_counter']['wall_ms'] / counted:.4f}",
            "step_kernel_events_per_s": f"{dc['device_counter']['device_steps_per_s']:.4e}",
            "speciation_only_steps_per_s": f"{dc['speciation_only']['steps_per_s']:.4e}",
            "host_drain_events_per_s": f"{dc['host_drain']['steps_per_s']:.4e}",
            "host_drain_GB_per_s": f"{dc['host_drain']['drained_GB_per_s']:.4e}",
            "hbm_bound_events_per_s": f"{HBM_PEAK_GBS * 1e9 / 280:.4e}",
            "fraction_of_hbm_bound": f"{dc['device_counter']['steps_per_s'] / (HBM_PEAK_GBS * 1e9 / 280):.4f}"}
    for k, v in head.items():
        print(f"{k}: {v}")
    print("\n## per round (kernel mean x calls per round; HBM bytes per event = calibrated FETCH_SIZE + "
          "WRITE_SIZE / events per round; the scan kernels are shared with the drain and speciation-only sorts)\n")
    print("| kernel | work | mean ms | calls/round | ms/round | HBM B/event |")
    print("|---|---|---|---|---|---|")
    total = 0.0
    for pat, work, per, in_total in rows:
        r = [x for n, x in kt.items() if pat in n]
        if not r:
            continue
        r = r[0]
        mean, cpr = float(r["AverageNs"]) / 1e6, int(r["Calls"]) / per
        total += mean * cpr if in_total else 0.0
        fk = [v for k, v in fe.items() if pat in k]
        wk = [v for k, v in wr.items() if pat in k]
        b = ((fk[0]["FETCH_SIZE"] * ff + wk[0]["WRITE_SIZE"] * wf) * 1024 / ev_round * cpr) if fk and wk else \
            float("nan")
        print(f"| {r['Name'][:60]} | {work} | {mean:.3f} | {cpr:.1f} | {mean * cpr:.3f} | {b:.1f} |")
    print(f"| total | | | | {total:.3f} | |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--window", type=float, default=2000.0, help="pause_before of the headline workload")
    ap.add_argument("--dedup-window", type=float, default=20000.0,
                    help="pause_before of the dedup-on config 4 comparison")
    ap.add_argument("--tail-window", type=float, default=200000.0,
                    help="pause_before of the config 4 heavy-tail comparison (0: skip it)")
    ap.add_argument("--no-completion", action="store_true",
                    help="skip the config 4 run to completion (persistent form)")
    ap.add_argument("--part", choices=("all", "headline", "gather", "dispersal", "tail"), default="all",
                    help="headline: the headline runs only; gather: the alias-gather window only; "
                         "dispersal: the dispersal-reporting windows only")
    ap.add_argument("--dispersal-window", type=float, default=200.0,
                    help="pause_before of the dispersal-reporting runs on the device")
    ap.add_argument("--task-window", type=float, default=200.0,
                    help="pause_before of the persistent vs task-list (reference CUDA structure) comparison")
    ap.add_argument("--dispersal-wide", type=float, default=20.0,
                    help="event slice factor of a second counted dispersal run (0: skip it)")
    ap.add_argument("--drain-window", type=float, default=20.0,
                    help="pause_before of the dispersal-reporting run drained to the host")
    ap.add_argument("--no-extra", action="store_true", help="same as --part headline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="default launch shape unless given below")
    ap.add_argument("--form", choices=("persistent", "plain"), default="persistent",
                    help="form of the headline runs")
    ap.add_argument("--block-size", type=int, default=0, help="fixed launch shape (skips the autotuner)")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--inner-steps", type=int, default=0)
    ap.add_argument("--exec-fill", type=int, default=0)
    ap.add_argument("--pool", type=int, default=-1,
                    help="pooled kernel with this many lineages per wave (128, 192; skips the autotuner); "
                         "0 = the one-lineage-per-lane kernel")
    ap.add_argument("--graph", action="store_true",
                    help="capture each water-level round (kernel + event sort) in a hipGraph; the "
                         "kernel timing then covers the whole graph")
    ap.add_argument("--summarize", nargs="+", metavar="DIR",
                    help="print a summary of rocprofv3 output directories and exit (no GPU)")
    ap.add_argument("--derive", nargs="+", metavar="ARG",
                    help="KIND BENCH_LOG DIR...: the derived figures of a committed profile (KIND headline: "
                         "kernel trace, FETCH_SIZE, WRITE_SIZE, SQ and instruction-mix directories; gather: the "
                         "first four; dispersal: the first three); prints text for profiles/ (no GPU)")
    a = ap.parse_args()
    if a.summarize:
        summarize(a.summarize)
        return
    if a.derive:
        derive(a.derive[0], a.derive[1], a.derive[2:])
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for {a.gpus})",
              file=sys.stderr)
        sys.exit(2)
    if a.no_extra:
        a.part = "headline"

    ctx = nh.Context.get(0)
    cus, arch = ctx.info()
    log(f"device {arch} with {cus} CUs")

    # plain copy kernel of known size: achievable HBM bandwidth and the
    # calibration of the FETCH_SIZE / WRITE_SIZE byte counters
    cms, cgbs = nh.copy_calibrate(ctx, 1 << 30, 10)
    copy = {"kernel": "k_copy_calibrate", "bytes_read": 1 << 30, "bytes_written": 1 << 30, "ms": cms,
            "GB_per_s": cgbs}
    log(f"copy kernel: {cgbs:.1f} GB/s")

    if a.part == "gather":
        r = gather_run(ctx, 2000.0)
        log(f"gather: {r}")
        print(json.dumps({"part": "gather", "gather_window": r, "copy_kernel": copy}), flush=True)
        return
    if a.part == "dispersal":
        cogs = nh.AlmostInfiniteNormalDispersalScenario(2000, 50.0).new(1e-6)
        lineages = cogs.sample_habitat(nh.OriginPreSampler.all())
        tables = nh.Tables(ctx, cogs)
        dargs = nh.IndependentArguments(dedup_cache=nh.DedupCache.none(), block_size=a.block_size,
                                        blocks_per_cu=a.blocks_per_cu, inner_steps=a.inner_steps,
                                        exec_fill=a.exec_fill, pool=DISPERSAL_POOL if a.pool < 0 else a.pool)
        r = dispersal_run(ctx, tables, cogs, lineages, dargs, a.dispersal_window, a.drain_window,
                          wide_slice=a.dispersal_wide)
        print(json.dumps({"part": "dispersal", "dispersal_windows": r, "copy_kernel": copy}), flush=True)
        return

    # ---- headline: config 4, dedup None window ----
    t = time.perf_counter()
    cogs = nh.AlmostInfiniteNormalDispersalScenario(2000, 50.0).new(1e-6)
    lineages = cogs.sample_habitat(nh.OriginPreSampler.all())
    tables = nh.Tables(ctx, cogs)
    log(f"config 4: {len(lineages)} lineages, setup {time.perf_counter() - t:.2f} s")
    if a.part == "tail":
        dd, dtuned = dedup_window_run(ctx, tables, cogs, lineages, a.dedup_window)
        log(f"config 4 dedup window: {dd}")
        tail = config4_tail_run(ctx, tables, cogs, lineages, dtuned, dd, a.tail_window,
                                completion=not a.no_completion)
        print(json.dumps({"part": "tail", "config4_dedup_window": dd, "config4_tail": tail}), flush=True)
        return
    form = nh.PERSISTENT if a.form == "persistent" else nh.PLAIN
    pool = max(a.pool, 0)
    fixed = a.block_size or a.blocks_per_cu or a.inner_steps or a.exec_fill or pool
    args = nh.IndependentArguments(dedup_cache=nh.DedupCache.none(), form=form,
                                   use_graph=a.graph and form == nh.PERSISTENT,
                                   block_size=a.block_size, blocks_per_cu=a.blocks_per_cu,
                                   inner_steps=a.inner_steps, exec_fill=a.exec_fill, pool=pool)
    tuned = None
    if not (a.no_tune or fixed or form == nh.PLAIN):
        t = time.perf_counter()
        # trials on every lineage over a short window: the pooled kernel's
        # advantage depends on how many lineages each pool entry serves
        targs, tms = nh.autotune(ctx, tables, args, nh.WyHash.seed_from_u64(42), cogs, lineages, 100.0,
                                 tune_step_slice=False)
        args = nh.IndependentArguments(dedup_cache=args.dedup_cache, form=nh.PERSISTENT, use_graph=args.use_graph,
                                       block_size=targs.block_size, blocks_per_cu=targs.blocks_per_cu,
                                       inner_steps=targs.inner_steps, exec_fill=targs.exec_fill, pool=targs.pool)
        tuned = {"block_size": targs.block_size, "blocks_per_cu": targs.blocks_per_cu,
                 "inner_steps": targs.inner_steps, "exec_fill": targs.exec_fill, "pool": targs.pool,
                 "trial_ms": tms, "tune_s": time.perf_counter() - t}
        log(f"autotuned: {tuned}")
    t = time.perf_counter()
    make_job(ctx, tables, cogs, lineages, args, 42, a.window).close()
    upload_s = time.perf_counter() - t  # host -> HBM upload of the lineages (+ allocation)
    # origin sampling: the host walk (then the upload above) against the
    # device sampler writing straight into a job's lineage store
    t = time.perf_counter()
    host_lin = cogs.sample_habitat(nh.OriginPreSampler.all())
    host_walk_s = time.perf_counter() - t
    oj = nh.Job(ctx, tables, args, nh.WyHash.seed_from_u64(42), cogs.speciation_probability, cogs.turnover_rate,
                pause_before=a.window)
    t = time.perf_counter()
    oj.add_origins(nh.OriginPreSampler.all(), cogs.circle)
    oj.start()
    device_origins_s = time.perf_counter() - t
    origin_sampling = {"lineages": int(len(host_lin)), "host_walk_s": host_walk_s,
                       "host_walk_and_upload_s": host_walk_s + upload_s,
                       "device_sampling_into_job_s": device_origins_s,
                       "same_lineages": oj.take_lineages().tobytes() == host_lin.tobytes()}
    oj.close()
    del host_lin
    log(f"origin sampling: {origin_sampling}")
    wall, results = timed_runs(ctx, tables, cogs, lineages, args, 42, a.window, a.steps, a.warmup)
    steps_total = sum(int(o.steps) for o, _, _ in results)
    value = steps_total / wall
    kern_ms = sum(ms for _, ms, _ in results)
    launches = sum(k for _, _, k in results)
    o0 = results[0][0]
    log(f"{a.form}: {steps_total} steps in {wall:.3f} s -> {value:.4e} steps/s; "
        f"device {kern_ms:.1f} ms over {launches} launches")

    # Roofline of the dominant kernel.  The almost-infinite scenario gathers
    # no tables, so the binding limit is VALU issue (64-bit integer hashing
    # and f64 libm): achieved wave instructions per second, from the
    # committed rocprofv3 count of VALU instructions per lineage step at a
    # launch shape times the steps per second that shape runs at here (HIP
    # events on the kernel's stream), against one wave64 instruction per CU
    # per cycle.  HBM is reported beside it: algorithmic bytes of one round
    # launch are every lineage's state but its reference read once (37 B),
    # the survivors' state written back (37 B), per speciation a dead flag,
    # the reference and the 40 B event.  The plain form reads the whole state
    # (45 B) and writes the moving part (37 B) of its pass list per launch.
    nl = max(launches, 1)
    n_in, n_left, n_ev = len(lineages), int(o0.n_remaining), int(o0.n_events)
    if form == nh.PERSISTENT:
        round_bytes = (n_in * MOVING_STATE_BYTES + n_left * MOVING_STATE_BYTES
                       + n_ev * (1 + 8 + EVENT_BYTES))
        per_launch_bytes = round_bytes * a.steps / nl
    else:
        per_launch_bytes = n_in * (LINEAGE_STATE_BYTES + MOVING_STATE_BYTES)
    mean_launch_ms = kern_ms / nl
    achieved_gbs = per_launch_bytes / (mean_launch_ms * 1e-3) / 1e9
    profile_path = PROFILE_SUMMARIES["pooled" if args.pool else "lane"]
    prof = read_profile_summary(profile_path)
    shape = (args.block_size, args.blocks_per_cu, args.inner_steps, args.exec_fill, args.pool)
    prof_shape = tuple(int(prof.get(k, 0 if k == "pool" else -1))
                       for k in ("block_size", "blocks_per_cu", "inner_steps", "exec_fill", "pool"))
    same_workload = prof.get("window") == a.window and prof.get("form") == a.form
    kname = ("k_plain<almost-infinite>" if form != nh.PERSISTENT else
             f"k_pooled<almost-infinite, {args.pool}>" if args.pool else "k_persistent<almost-infinite, no dedup>")
    roofline = {"bound": "valu", "achieved": None, "peak": cus * ENGINE_CLOCK_HZ, "unit": "wave-instructions/s",
                "frac": None, "traffic": None, "kernel": kname}
    ipl = prof.get("valu_wave_instructions_per_lineage_step") if same_workload else None
    if ipl:
        if prof_shape == shape:
            v_steps, v_ms, v_launches = steps_total, kern_ms, nl
        else:
            # the instruction count was profiled at another launch shape: time
            # that shape here, so that count and time belong to the same launch
            vargs = dataclasses.replace(args, block_size=prof_shape[0], blocks_per_cu=prof_shape[1],
                                        inner_steps=prof_shape[2], exec_fill=prof_shape[3], pool=prof_shape[4])
            _, vres = timed_runs(ctx, tables, cogs, lineages, vargs, 42, a.window, 1, 0)
            v_steps, v_ms, v_launches = int(vres[0][0].steps), vres[0][1], vres[0][2]
        inst_rate = ipl * v_steps / (v_ms * 1e-3)
        roofline.update({"achieved": inst_rate, "frac": round(inst_rate / roofline["peak"], 4),
                         "lane_utilization": prof.get("valu_lane_utilization"),
                         "lane_weighted_frac": round(inst_rate / roofline["peak"]
                                                     * (prof.get("valu_lane_utilization") or 0.0), 4),
                         "wave_instructions_per_lineage_step": ipl,
                         "profiled_launch_shape": {"block_size": prof_shape[0], "blocks_per_cu": prof_shape[1],
                                                   "inner_steps": prof_shape[2], "exec_fill": prof_shape[3],
                                                   "pool": prof_shape[4]},
                         "mean_launch_ms_at_profiled_shape": round(v_ms / max(v_launches, 1), 3),
                         "source": os.path.relpath(profile_path, ROOT)})
    roofline["traffic"] = prof.get("hbm_bytes_per_launch") if same_workload and prof_shape == shape else None
    roofline.update({"mean_launch_ms": round(mean_launch_ms, 3), "launches": nl,
                     "lineage_steps_per_launch": int(steps_total / nl),
                     "hbm": {"achieved": round(achieved_gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(achieved_gbs / HBM_PEAK_GBS, 6),
                             "algorithmic_bytes_per_launch": int(per_launch_bytes),
                             "traffic": prof.get("hbm_bytes_per_launch") if same_workload else None},
                     "note": "the almost-infinite scenario gathers no tables: the kernel is bound by VALU issue "
                             "(frac); lane_weighted_frac scales it by the share of active lanes; HBM carries "
                             "only the lineage state once per round"})

    extra = {"autotuned": tuned,
             "launch": {"form": a.form, "block_size": args.block_size, "blocks_per_cu": args.blocks_per_cu,
                        "inner_steps": args.inner_steps, "exec_fill": args.exec_fill, "pool": args.pool,
                        "use_graph": bool(args.use_graph)},
             "copy_kernel": copy,
             "upload_s": upload_s,
             "origin_sampling": origin_sampling,
             "pcie_inclusive_value": steps_total / (wall + a.steps * upload_s)}
    if a.part == "all":
        # bit-exact comparison with the plain one-launch-per-slice form
        pj = make_job(ctx, tables, cogs, lineages, nh.IndependentArguments(dedup_cache=nh.DedupCache.none(),
                                                                          form=nh.PLAIN, plain_batch=1),
                      42, a.window)
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
        extra["task_list_window"] = task_list_run(ctx, tables, cogs, lineages, args, a.task_window)

        dd, dtuned = dedup_window_run(ctx, tables, cogs, lineages, a.dedup_window)
        extra["config4_dedup_window"] = dd
        log(f"config 4 dedup window: {extra['config4_dedup_window']}")
        if a.tail_window > 0:
            extra["config4_tail"] = config4_tail_run(ctx, tables, cogs, lineages, dtuned,
                                                     extra["config4_dedup_window"], a.tail_window,
                                                     completion=not a.no_completion)
        # dispersal reporting runs the pooled kernel at full residency (its
        # lineages enter and leave the pool every two steps or so)
        dargs = dataclasses.replace(args, pool=DISPERSAL_POOL, block_size=0, blocks_per_cu=0, inner_steps=0,
                                    exec_fill=0)
        extra["dispersal_reporting"] = dispersal_run(ctx, tables, cogs, lineages, dargs, a.dispersal_window,
                                                     a.drain_window, wide_slice=a.dispersal_wide)
        del tables
        extra["config2_full_run"] = config2_run(ctx)
        log(f"config 2: {extra['config2_full_run']}")
        extra["chain_latency"] = chain_latency_run(ctx)
        log(f"chain latency: {extra['chain_latency']}")
        extra["config3_window"] = config3_run(ctx, 2000.0)
        # the same map without dedup: the gathers from its tables (61 MB,
        # resident in the 256 MB last-level cache) over the whole window
        extra["config3_gather_window"] = map_window_run(ctx, 128, 1e-5, 2000.0, nh.DedupCache.none())
        log(f"config 3 without dedup: {extra['config3_gather_window']}")
        log(f"config 3: {extra['config3_window']}")
        extra["gather_window"] = gather_run(ctx, 2000.0)
        log(f"gather: {extra['gather_window']}")
        gc.collect()  # the earlier parts' host arrays: the serving loop polls from the host
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
                               f"DedupCache::None, {a.form} form, delta_t 2, event_slice Relative(2.0)",
                   "lineages": int(len(lineages)), "steps_per_run": int(o0.steps),
                   "speciations_per_run": n_ev, "rounds": int(o0.rounds), "remaining_lineages": n_left},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
