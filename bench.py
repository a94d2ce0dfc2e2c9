"""Benchmark of the KKS phase-field step on one MI355X.

Headline (value): the plain form on 4096 x 4096 interior cells, FP64,
dx = 0.25 nm, periodic, seeded uniform matrix composition in
Ni-(0.0202..0.0269)Nb-(0.2794..0.3288)Cr with a central enriched strip (1/8 of
the width) at Ni-0.169Nb-0.272Cr, nucleation on (Philox seed 42, every
floor(1e-4/dt) steps, fused into the step), dt = 0.1 dx^2 / (4 D_max).
A "step" is one explicit Euler step over the whole grid; steps run four per
launch (the temporally blocked kernel; FP32 storage: three) except around
nucleation steps, which run one per launch with the nucleation pass fused.  value = interior cells x
timed steps / time of the timed loop (HIP events), fields resident in HBM.

Clock state: the GPU's clock takes ~50 launches of sustained load to settle
(the first launches of a cold run are up to 30% slower), so before the warm-up
steps the four-step kernel runs --settle launches on scratch copies of the
fields; the timed loop and the kernel leg then run in the same settled state,
and the line fails (no output, exit 3) if the kernel leg's time per step
exceeds the timed loop's by more than 10% (a short timed loop runs 1-3%
faster per launch than the leg's sustained back-to-back launches; the ratio is
in the line as timing.kernel_leg_over_timed_loop).

roofline: the dominant kernel is the four-step kernel; its algorithmic bytes
per launch (FP64: 4 fields read + 4 written = 64 B per cell, for four
cell-updates) over its average launch duration (HIP events on the launch
stream, --kernel-reps back-to-back launches right after the timed loop); the
three-, two- and one-step kernels are reported beside it.  traffic: HBM bytes per
launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs under profiles/ (scaled
by a copy kernel).
cpu_baseline: the FP64 CPU oracle built -O3 -march=native on the host
(oracle/cpu_bench.py, a child process, OpenMP on every physical core with
OMP_PROC_BIND=close) on a bounded sample of the same workload; beside it one
core, and the 256^2 plain case (1000 steps) in full.

Also reported (same JSON line):
  dropin    -- the reference driver's per-step call sequence through the C++
               drop-in API (tests/cpp/replay_main) on the same grid (no-flux,
               nucleation on, no I/O), grouped four steps per launch (eight
               calls to two launches), three, and one.
  batched   -- 512 independent simulations, square sides log-uniform in
               128..1024, FP32 storage with FP64 arithmetic and reductions, one
               persistent launch per one, two or three steps replayed from a
               hipGraph; tile, steps per launch and launches per graph from the
               autotuner; against loops of plain FP64 launches (four-, three-,
               two- and one-step) over the same instances; phase-fraction error
               against FP64 replays of sampled instances.
  precision -- 8192^2, the headline's initialisation, FP32 storage (with and
               without a composition offset) against FP64, 20,000 steps: max
               field differences, phase fractions and mass drift at 1k/10k/20k
               steps; and the same grid with seeded delta and Laves particles
               (FP64 against FP32 with the offset) so that the phase-fraction
               error is measured where the fractions are not zero.
  serving   -- Poisson sweep requests (side log-uniform 128..2048, 500..20000
               steps, random composition and nucleation seed) with continuous
               batching under an HBM budget, at several arrival rates:
               throughput, p50/p99 latency and the device-busy share; and the
               shortest-remaining-work order against FIFO at the highest rate.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form plain|precision]
    python bench.py --trace-summary <rocprofv3 kernel_trace.csv>
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "phasefield-precipitate-aging_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s
METRIC = "cell-updates/s (% HBM roofline) plain vs batched; p50/p99 sweep latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_traffic(workload):
    """HBM bytes per step-kernel launch from the committed rocprofv3 counter
    runs (profiles/traffic_<workload>.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def load_valu_issue(kernel):
    """VALU issue time per launch (us) of `kernel` at 4096^2 from the committed
    rocprofv3 SQ counter summary (profiles/sq_counters_steps_4096.txt), or None."""
    path = os.path.join(ROOT, "profiles", "sq_counters_steps_4096.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(kernel):
                return float(line.split()[-1])
    return None


def host_cpu():
    """Model name, logical CPUs this process may run on, the physical cores
    among them (distinct (package, core) pairs from /sys) and the sockets."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                if k.strip() == "model name":
                    model = v.strip()
                    break
    except OSError:
        pass
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    cores, sockets = set(), set()
    for c in cpus:
        base = "/sys/devices/system/cpu/cpu%d/topology/" % c
        try:
            with open(base + "physical_package_id") as f:
                pkg = f.read().strip()
            with open(base + "core_id") as f:
                core = f.read().strip()
        except OSError:
            continue
        cores.add((pkg, core))
        sockets.add(pkg)
    return {"model": model, "logical_cpus": len(cpus), "physical_cores": len(cores) or None,
            "sockets": len(sockets) or None, "os_cpu_count": os.cpu_count()}


def cpu_baseline(args):
    """The FP64 CPU oracle built -O3 -march=native for this host, timed in a
    child process (oracle/cpu_bench.py) on every physical core this process
    may use (OMP_PROC_BIND=close) and on one core."""
    import subprocess
    cpu = host_cpu()
    # every physical core, within the 16-CPU share a one-GPU host gives a job
    threads = min(16, cpu["physical_cores"] or cpu["logical_cpus"])
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    cmd = [sys.executable, os.path.join(ROOT, "oracle", "cpu_bench.py"),
           "--k1-size", str(args.size), "--k1-steps", str(args.cpu_steps),
           "--k1-one-steps", str(args.cpu_one_steps), "--k0-steps", "1000"]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError("CPU baseline failed: " + r.stderr[-2000:])
    d = json.loads(r.stdout.strip().splitlines()[-1])
    k1, k1o, k0 = d["k1"], d["k1_one"], d["k0"]
    return {
        "value": k1["cell_updates_per_s"], "unit": "cell-updates/s", "cores": k1["threads"],
        "kind": "port",
        "sample": "%d steps of the %dx%d FP64 plain step (same init, periodic, nucleation "
                  "off) after 1 warm-up step, %.1f s" % (k1["steps"], args.size, args.size,
                                                          k1["seconds"]),
        "build": d["build"] + " (timed); the checker is built -O2 -ffp-contract=off",
        "omp": {"OMP_NUM_THREADS": threads, "OMP_PROC_BIND": "close", "OMP_PLACES": "cores"},
        "host_cpu": cpu,
        "one_core": {"value": k1o["cell_updates_per_s"], "cores": 1,
                     "sample": "%d steps of the same %dx%d step, %.1f s"
                               % (k1o["steps"], args.size, args.size, k1o["seconds"])},
        "k0_256_1000_steps": {
            "grid": [256, 256], "steps": k0["steps"], "bc": "periodic",
            "nucleation": False, "dt": k0["dt"],
            "all_cores": {"value": k0["all_threads"]["cell_updates_per_s"],
                          "cores": k0["all_threads"]["threads"],
                          "seconds": k0["all_threads"]["seconds"]},
            "one_core": {"value": k0["one_thread"]["cell_updates_per_s"], "cores": 1,
                         "seconds": k0["one_thread"]["seconds"]}},
    }


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _params(s):
    from pfpa._lib import PfParams
    P = PfParams()
    P.dt = s.dt
    for k in range(2):
        P.DCr[k], P.DNb[k] = s.DCr[k], s.DNb[k]
        P.kappa[k], P.omega[k], P.Lmob[k] = s.kappa[k], s.omega[k], s.Lmob[k]
    P.alpha = s.alpha
    for k in range(9):
        P.mask[k] = s.mask[k]
    return P


def plain_section(args, dtype_name="f64"):
    import torch

    from pfpa import model as M
    from pfpa._lib import check, lib
    from pfpa.sim import Simulation

    N = args.size
    dtype = torch.float64 if dtype_name == "f64" else torch.float32
    f64 = dtype == torch.float64
    s = M.make_setup(N, N, bc=1)
    f0 = M.init_enriched_band(N, N, 1, seed=42)
    sim = Simulation(s, f0, dtype=dtype, nucleation=True, seed=42)
    del f0
    stream = sim.stream
    P = sim.P
    st = ctypes.c_void_p(stream.cuda_stream)
    # scratch copies of the fields: the clock-settling launches and the
    # kernel leg run on these, so the simulation itself advances exactly
    # warmup + steps steps
    with torch.cuda.stream(stream):
        a = [x.clone() for x in sim.a]
        b = [torch.empty_like(x) for x in a]
    fn = {4: lib().pf_step4_f64 if f64 else lib().pf_step4_f32,
          3: lib().pf_step3_f64 if f64 else lib().pf_step3_f32,
          2: lib().pf_step2_f64 if f64 else lib().pf_step2_f32,
          1: lib().pf_step_f64 if f64 else lib().pf_step_f32}
    K = sim.steps_per_launch  # the dominant kernel: four steps per launch (FP64), three (FP32)

    def launch(k, n):
        for i in range(n):
            x, y = (a, b) if i % 2 == 0 else (b, a)
            args_ = [sim.ctx, s.nx, s.ny, s.bc, ctypes.byref(P), *[_ptr(t) for t in x],
                     *[_ptr(t) for t in y]] + ([None] if k == 1 else []) + [st]
            check(fn[k](*args_), "step")

    def time_kernel(k, reps):
        k0 = torch.cuda.Event(enable_timing=True)
        k1 = torch.cuda.Event(enable_timing=True)
        k0.record(stream)
        launch(k, reps)
        k1.record(stream)
        k1.synchronize()
        return k0.elapsed_time(k1) / 1e3 / reps

    # settle the clock on the scratch copies, then warm up and time the
    # simulation's own steps
    settle_s = time_kernel(K, args.settle) if args.settle > 0 else 0.0
    sim.step(args.warmup)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    sim.step(args.steps)
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1)
    cells = N * N
    value = cells * args.steps / (ms / 1e3)

    # The kernels alone, back to back on the sim's stream right after the
    # timed loop (same clock state): the K-step kernel (the dominant one),
    # the shorter multi-step kernels (what the schedule leaves between
    # nucleation steps) and the one-step kernel (nucleation steps; the drop-in
    # API with grouping off).
    kern_s = {k: time_kernel(k, args.kernel_reps) for k in range(K, 0, -1)}
    kernK_s = kern_s[K]
    # The timed loop is the K-step kernel plus a few one-step nucleation
    # launches.  A short loop (20 steps: five launches) runs 1-3% faster per
    # launch than the kernel leg's 200 back-to-back launches, whose sustained
    # load holds the clock a little lower; a cold clock made the leg up to
    # 30% slower.  A leg more than 10% slower per step than the loop ran in
    # another clock state.
    leg_vs_loop = kernK_s / K * 1e3 / (ms / args.steps)
    if leg_vs_loop > 1.10:
        log("bench.py: the %d-step kernel's time per step (%.2f us, %d back-to-back "
            "launches) exceeds the timed loop's (%.2f us) by more than 10%%: the two "
            "measurements are not in the same clock state"
            % (K, kernK_s / K * 1e6, args.kernel_reps, ms / args.steps * 1e3))
        sys.exit(3)
    bpc = 64 if f64 else 32
    alg_bytes = bpc * cells          # 4 fields read + 4 written per launch
    tag = "fp64" if f64 else "fp32"

    def other(name, k, secs):
        return {"kernel": name, "steps_per_launch": k, "kernel_avg_us": secs * 1e6,
                "achieved": alg_bytes / secs / 1e9, "frac": alg_bytes / secs / HBM_PEAK,
                "traffic": load_traffic("plain%s_%s_%d" % ("" if k == 1 else k, tag, N)),
                "algorithmic_bytes_per_launch": alg_bytes}

    nuc_every = sim.nuc_every
    sim.close()
    tname = "double" if f64 else "float"
    kname = {4: "pf::step4_kernel<%s, 1, 3>" % tname, 3: "pf::stepk_kernel<%s, 1, 3, 3>" % tname,
             2: "pf::stepk_kernel<%s, 1, 2, 3>" % tname}
    valu_us = load_valu_issue(kname[K]) if N == 4096 else None
    words = {1: "one", 2: "two", 3: "three", 4: "four"}
    others = {"%s_step_kernel" % words[k]:
              other("%s (%s fused step%s per launch)" % (kname.get(k, "pf::step_kernel"),
                                                       words[k], "s" if k > 1 else ""), k,
                    kern_s[k]) for k in range(K - 1, 0, -1)}
    return {
        "value": value, "ms_per_step": ms / args.steps, "N": N, "nuc_every": nuc_every,
        "dt": s.dt,
        "timing": {
            "clock_settle": "%d back-to-back %s-step launches on scratch copies of the "
                            "fields before the warm-up (%.1f us each)"
                            % (args.settle, words[K], settle_s * 1e6),
            "kernel_leg": "%d back-to-back launches of each kernel on the scratch copies right "
                          "after the timed loop" % args.kernel_reps,
            "%s_step_kernel_us_per_step" % words[K]: kernK_s / K * 1e6,
            "timed_loop_us_per_step": ms / args.steps * 1e3,
            "kernel_leg_over_timed_loop": leg_vs_loop,
        },
        "roofline": {
            "bound": "hbm", "achieved": alg_bytes / kernK_s / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": alg_bytes / kernK_s / HBM_PEAK,
            "traffic": load_traffic("plain%d_%s_%d" % (K, tag, N)),
            "kernel": "%s (%s fused steps per launch)" % (kname[K], words[K]),
            "kernel_avg_us": kernK_s * 1e6,
            "algorithmic_bytes_per_launch": alg_bytes,
            "cell_updates_per_launch": K * cells,
            "algorithmic_bytes_per_cell_update": bpc / K,
            # the rate per step against the one-step roofline (bpc bytes per
            # cell-update at peak HBM), and the share of the launch the VALU
            # instructions alone need (rocprofv3 SQ_INSTS_VALU, profiles/)
            "cell_update_rate_vs_one_step_roofline": value * bpc / HBM_PEAK,
            "valu_issue_us": valu_us,
            "valu_issue_frac": (valu_us / (kernK_s * 1e6)) if valu_us else None,
            "kernel_us_per_step": kernK_s / K * 1e6,
            "note": "temporal blocking: one read and one write of the fields per launch "
                    "carry %d cell-updates per cell, so frac falls as K rises while the time "
                    "per step falls; the launch is bound by FP64 arithmetic at two waves per "
                    "SIMD (valu_issue_frac), not by HBM" % K,
            **others,
        },
    }


def batched_section(args):
    """512 uneven simulations: batched FP32 vs a loop of plain FP64 launches."""
    import torch

    from pfpa._lib import PfStepItem, check, lib
    from pfpa.batch import PF_STORAGE_F32, PF_STORAGE_F64, Batch, job_bytes, make_desc
    from pfpa.sim import Simulation
    from pfpa.workloads import (autotune_batch, batch_cell_updates, batch_specs, spec_fields,
                                spec_setup, step_counts)

    specs = batch_specs(args.batch_instances, seed=7)
    cells = sum(sp["n"] ** 2 for sp in specs)
    pool = sum(job_bytes(PF_STORAGE_F32, sp["n"], sp["n"]) for sp in specs) + (64 << 20)
    batch = Batch(PF_STORAGE_F32, pool_bytes=pool, max_jobs=len(specs))
    total_steps = args.batch_steps
    # room for the autotuner's launches (at most 3 steps x 10 launches x 45
    # configurations, then the graph candidates), so that every job is still
    # active through the timed run
    target = total_steps + 2000
    plain_inputs = []
    t_gen = time.perf_counter()
    for sp in specs:
        s = spec_setup(sp)
        f = spec_fields(sp)
        # FP32 compositions stored relative to the job's matrix composition
        batch.add(make_desc(s, target, nucleation=True, seed=1000 + sp["index"],
                            instance=sp["index"], x_offset=sp["x0"]), f)
        plain_inputs.append((s, f))
    log("batched: %d instances, %.3g cells, setup %.1f s" % (len(specs), cells,
                                                            time.perf_counter() - t_gen))
    if args.batch_tile:
        # a fixed configuration instead of the autotuner (counter runs)
        rows, depth, ls, g = args.batch_tile
        batch.set_tile(rows, depth)
        batch.set_launch_steps(ls)
        table = [(0.0, rows, depth, ls, g)]
    else:
        best, table = autotune_batch(batch, launches=8)
        _, rows, depth, ls, g = best
    tuned = {"rows": rows, "depth": depth, "launch_steps": ls, "launches_per_graph": g,
             "candidates": len(table),
             "table": [{"rate": r[0], "rows": r[1], "depth": r[2], "launch_steps": r[3],
                        "launches_per_graph": r[4]} for r in sorted(table, reverse=True)]}
    log("batched autotune:", json.dumps(tuned["table"][:6]))
    batch.run(g, g)
    batch.sync()
    before = step_counts(batch)
    launches = max(g, (total_steps // ls) // g * g)
    t0 = time.perf_counter()
    batch.run(launches, g)
    batch.sync()
    el = time.perf_counter() - t0
    done = batch_cell_updates(batch, before)
    batched_rate = done / el
    after = step_counts(batch)
    # phase fractions of sampled jobs after their steps so far
    order = sorted(range(len(specs)), key=lambda k: specs[k]["n"])
    sample = [order[i] for i in (0, len(order) // 3, 2 * len(order) // 3)]
    batched_summ = {k: batch.summarize(k) for k in sample}
    batch.status()
    batch.close()

    # the same mix and configuration with FP64 storage: separates what the
    # persistent batched launch gains (filling the chip with many small grids)
    # from what FP32 storage gains (half the bytes)
    b64 = Batch(PF_STORAGE_F64, pool_bytes=sum(job_bytes(PF_STORAGE_F64, sp["n"], sp["n"])
                                                for sp in specs) + (64 << 20),
                max_jobs=len(specs))
    for sp, (s, f) in zip(specs, plain_inputs):
        b64.add(make_desc(s, target, nucleation=True, seed=1000 + sp["index"],
                          instance=sp["index"]), f)
    # its own strip height (FP64 storage moves twice the bytes and prefers
    # shorter chunks than FP32), the FP32 winner's depth, steps and graphs
    if args.batch_tile:
        rows64 = rows
        b64.set_tile(rows, depth)
        b64.set_launch_steps(ls)
    else:
        best64, _ = autotune_batch(b64, launches=8, rows_set=(32, 48, 64), depth_set=(depth,),
                                   launch_steps_set=(ls,), graph_set=(g,))
        rows64 = best64[1]
    b64.run(g, g)
    b64.sync()
    before64 = step_counts(b64)
    launches64 = max(g, min(launches, 300) // g * g)
    t0 = time.perf_counter()
    b64.run(launches64, g)
    b64.sync()
    fp64_storage_rate = batch_cell_updates(b64, before64) / (time.perf_counter() - t0)
    b64.close()
    log("batched FP64 storage: %.4g cell-updates/s (rows %d)" % (fp64_storage_rate, rows64))

    # loops of plain FP64 launches over the same instances, issued from C, at
    # four, three, two and one steps per launch; the fastest is the comparison
    ctx = ctypes.c_void_p()
    check(lib().pf_ctx_create(ctypes.byref(ctx)), "ctx")
    bufs = []
    prms = []
    for s, f in plain_inputs:
        A = [torch.as_tensor(x).cuda() for x in f]
        B = [torch.empty_like(x) for x in A]
        bufs.append((A, B))
        prms.append(_params(s))
    items = [(PfStepItem * len(bufs))(), (PfStepItem * len(bufs))()]
    for par in range(2):
        for i, ((A, B), (s, _), P) in enumerate(zip(bufs, plain_inputs, prms)):
            it = items[par][i]
            it.nx, it.ny, it.bc = s.nx, s.ny, s.bc
            it.prm = ctypes.pointer(P)
            src, dst = (A, B) if par == 0 else (B, A)
            for k in range(4):
                it.old_fields[k] = src[k].data_ptr()
                it.new_fields[k] = dst[k].data_ptr()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    loop_rates = {}
    for name, fn, per in (("four_step", lib().pf_step4_loop_f64, 4),
                          ("three_step", lib().pf_step3_loop_f64, 3),
                          ("two_step", lib().pf_step2_loop_f64, 2),
                          ("one_step", lib().pf_step_loop_f64, 1)):
        for par in range(2):
            check(fn(ctx, len(bufs), items[par], st), "loop")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.loop_steps):
            check(fn(ctx, len(bufs), items[i % 2], st), "loop")
        torch.cuda.synchronize()
        loop_rates[name] = cells * per * args.loop_steps / (time.perf_counter() - t0)
    loop_rate = max(loop_rates.values())
    lib().pf_ctx_destroy(ctx)
    del bufs

    # FP64 replays of the sampled jobs for the phase-fraction error
    errs = []
    for k in sample:
        s, f = plain_inputs[k]
        sim = Simulation(s, f, dtype=torch.float64, nucleation=True, seed=1000 + k, instance=k)
        sim.step(after[k])
        ref = sim.summarize()
        sim.close()
        errs.append(max(abs(ref[3] - batched_summ[k][3]), abs(ref[4] - batched_summ[k][4])))
    return {
        "instances": len(specs), "cells_per_step": cells, "launches": launches,
        "steps_per_job": after[0] - before[0],
        "dtype": "f32-storage/f64-compute",
        "value": batched_rate, "unit": "cell-updates/s",
        "plain_loop_fp64_value": loop_rate, "plain_loop_fp64_rates": loop_rates,
        "plain_loop_launches": args.loop_steps,
        "speedup_vs_plain_loop": batched_rate / loop_rate,
        "fp64_storage_value": fp64_storage_rate,
        "fp64_storage_launches": launches64,
        "fp64_storage_rows": rows64,
        "speedup_fp64_storage_vs_plain_loop": fp64_storage_rate / loop_rate,
        "autotuned": tuned,
        "phase_fraction_max_abs_error_vs_fp64": max(errs),
        "phase_fraction_sample": {"jobs": sample, "steps": [after[k] for k in sample]},
    }


def mobility_section(args):
    """The variable-mobility Cahn-Hilliard divergence (pf_mobility_divergence_f64)
    at the headline's grid: 24 FP64 input arrays read and 2 written per cell,
    208 B per cell, timed with HIP events over back-to-back launches after
    warm-up launches."""
    import torch

    from pfpa import api
    N = args.size
    shape = (N + 2, N + 2)
    g = torch.Generator(device="cuda").manual_seed(11)

    def field(lo, hi):
        return lo + (hi - lo) * torch.rand(shape, dtype=torch.float64, device="cuda", generator=g)

    conc = [field(0.0, 0.5) for _ in range(6)]
    mob_cr = [field(1e-21, 1e-20) for _ in range(8)]
    mob_nb = [field(1e-21, 1e-20) for _ in range(8)]
    pd, pl = field(0.0, 1.0), field(0.0, 1.0)
    out = [torch.empty_like(pd), torch.empty_like(pd)]
    dx = 0.25e-9
    st = torch.cuda.current_stream()
    for _ in range(20):
        api.chemical_convolution(conc, mob_cr, mob_nb, pd, pl, dx, dx, *out)
    reps = args.kernel_reps
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        api.chemical_convolution(conc, mob_cr, mob_nb, pd, pl, dx, dx, *out)
    e1.record(st)
    e1.synchronize()
    secs = e0.elapsed_time(e1) / 1e3 / reps
    cells = shape[0] * shape[1]
    alg = 26 * 8 * cells
    del conc, mob_cr, mob_nb, pd, pl, out
    return {"kernel": "pf::mobility_divergence_kernel", "grid_padded": list(shape),
            "kernel_avg_us": secs * 1e6, "algorithmic_bytes_per_launch": alg,
            "algorithmic_bytes_per_cell": 208, "achieved": alg / secs / 1e9,
            "frac": alg / secs / HBM_PEAK, "unit": "GB/s", "bound": "hbm",
            "cells_per_s": cells / secs}


def dropin_section(args):
    """The reference driver's call sequence (src/main.cpp:289-321) through the
    C++ drop-in API (tests/cpp/replay_main, which replays it call for call) on
    the plain workload: host wall clock over the last two thirds of the steps,
    until the device has finished them (the final read-out to host memory is
    reported on its own)."""
    import re
    import subprocess
    import tempfile

    import numpy as np

    from pfpa import model as M
    harness = os.path.join(ROOT, "tests", "cpp", "replay_main")
    if not os.path.exists(harness):
        return None
    N = args.size
    s = M.make_setup(N, N, bc=0)
    f0 = M.init_enriched_band(N, N, 0, seed=42)
    out = {}
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        with open(fin, "wb") as fh:
            fh.write(np.array([N + 2, N + 2], dtype=np.int32).tobytes())
            for a in f0:
                fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        del f0
        nuc = M.nuc_interval(s.dt)
        for spl in (4, 3, 1):
            r = subprocess.run([harness, fin, os.path.join(td, "out.bin"), str(args.dropin_steps),
                                repr(s.dt), repr(s.dx), str(nuc), "42", "0", str(spl)],
                               capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError("drop-in replay failed: " + r.stderr[-2000:])
            m = re.search(r"last (\d+) steps in ([0-9.eE+-]+) s; read-out ([0-9.eE+-]+) s",
                          r.stdout)
            n, secs, rsecs = int(m.group(1)), float(m.group(2)), float(m.group(3))
            out["steps_per_launch_%d" % spl] = {"value": N * N * n / secs,
                                                "ms_per_step": secs / n * 1e3, "steps_timed": n,
                                                "read_out_s": rsecs}
    return {"unit": "cell-updates/s", "grid": [N, N], "bc": "no-flux (the driver's)",
            "nucleation_interval": nuc, **out}


def serving_section(args, capacity=None):
    """The serving loop at each rate of --serve-rates (FIFO), and the
    shortest-remaining-work order at the highest rate, on the same seeded
    request streams.  capacity: the batched form's cell-updates/s, for the
    share of it each rate uses."""
    from pfpa.serve import fp64_replay_error, make_requests, serve

    def one(rate, order):
        reqs = make_requests(rate, args.serve_duration, seed=1)
        done = [0]

        def progress(r):   # one line per 100 completions (long runs stay visibly alive)
            done[0] += 1
            if done[0] % 100 == 0:
                log("serving %s %.1f/s: %d of %d requests done" % (order, rate, done[0],
                                                                   len(reqs)))

        rows = args.serve_rows if args.serve_rows == "auto" else int(args.serve_rows)
        finished, stats = serve(reqs, pool_bytes=args.serve_pool_gb << 30,
                                launches_per_graph=args.serve_lpg, rows=rows,
                                depth=3, launch_steps=3, log=progress, order=order)
        # FP64 replays of the requests with the least, the median and the most
        # work (cells x steps)
        by_work = sorted(finished, key=lambda r: r.n * r.n * r.steps)
        sample = [by_work[0], by_work[len(by_work) // 2], by_work[-1]] if by_work else []
        err, rows = fp64_replay_error(finished, sample)
        stats["arrival_rate_per_s"] = rate
        stats["arrival_window_s"] = args.serve_duration
        stats["phase_fraction_max_abs_error_vs_fp64"] = err
        stats["fp64_replays"] = rows
        if capacity:
            stats["throughput_share_of_batched_capacity"] = \
                stats["throughput_cell_updates_per_s"] / capacity
        log("serving:", json.dumps({k: v for k, v in stats.items()
                                    if k != "latency_by_size"}))
        return stats

    rates = sorted(args.serve_rates)
    sweep = [one(r, "fifo") for r in rates]
    out = {"rates": sweep, "batched_capacity_cell_updates_per_s": capacity,
           "pool_bytes": args.serve_pool_gb << 30,
           "requests": "Poisson arrivals over %.0f s; side log-uniform 128..2048, steps "
                       "uniform 500..20000, composition uniform in the matrix or enriched "
                       "range, random nucleation seed (seed 1 at every rate)"
                       % args.serve_duration}
    if not args.no_srw:
        out["srw_at_highest_rate"] = one(rates[-1], "srw")
    return out


def _precision_run(args, s, f0, forms, label):
    """Step the forms {name: (dtype, storage_offset)} of one workload side by
    side to each mark; rows of max |dx| against FP64, phase fractions and
    each form's own mass drift; rates from HIP events around the stepping."""
    import torch

    from pfpa.sim import Simulation
    N = s.N[0]
    sims = {n: Simulation(s, f0, dtype=dt, nucleation=True, seed=42, storage_offset=off)
            for n, (dt, off) in forms.items()}
    m0 = {n: sim.summarize() for n, sim in sims.items()}
    ref = sims["fp64"]
    secs = {n: 0.0 for n in sims}
    rows = []
    done = 0
    for mark in args.precision_marks:
        for n, sim in sims.items():
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(sim.stream)
            sim.step(mark - done)
            e1.record(sim.stream)
            e1.synchronize()
            secs[n] += e0.elapsed_time(e1) / 1e3
        done = mark
        summ = {n: sim.summarize() for n, sim in sims.items()}
        row = {"step": mark}
        for n, sim in sims.items():
            if n != "fp64":
                for k, c in ((0, "Cr"), (1, "Nb")):
                    row["max_abs_dx_%s_%s" % (c, n)] = float(
                        (ref.a[k] - (sim.a[k].double() + sim.xoff[k])).abs().max())
        for n in sims:
            row["delta_frac_" + n] = summ[n][3]
            row["laves_frac_" + n] = summ[n][4]
            if n != "fp64":
                row["phase_frac_abs_err_" + n] = max(abs(summ[n][3] - summ["fp64"][3]),
                                                     abs(summ[n][4] - summ["fp64"][4]))
        for n in sims:
            row["mass_drift_Cr_" + n] = summ[n][0] - m0[n][0]
            row["mass_drift_Nb_" + n] = summ[n][1] - m0[n][1]
        rows.append(row)
        log("precision %s" % label, json.dumps(row))
    offsets = {n: list(sim.xoff) for n, sim in sims.items() if sim.xoff != (0.0, 0.0)}
    for sim in sims.values():
        sim.close()
    rates = {n: N * N * done / secs[n] for n in sims}
    return {"workload": label, "rows": rows, "rates_cell_updates_per_s": rates,
            "offsets": offsets, "steps": done}


def precision_section(args):
    """8192^2 with the headline's initialisation: FP32 storage without and
    with a composition offset (the median composition) against FP64; then
    the same initialisation with seeded delta and Laves particles, FP32 with
    the offset against FP64, so that the phase-fraction error is measured
    where the fractions are not zero."""
    import torch

    from pfpa import model as M
    N = args.precision_size
    s = M.make_setup(N, N, bc=1)
    f0 = M.init_enriched_band(N, N, 1, seed=42)
    plain = _precision_run(args, s, f0, {
        "fp64": (torch.float64, None), "fp32": (torch.float32, None),
        "fp32_offset": (torch.float32, "median")},
        "%dx%d, the headline's initialisation (no precipitates), periodic, nucleation on"
        % (N, N))
    n = args.precision_particles
    M.add_particles(f0, s.dx, 1, seed=7, n_del=n, n_lav=n, radius=5e-9)
    parts = _precision_run(args, s, f0, {
        "fp64": (torch.float64, None), "fp32_offset": (torch.float32, "median")},
        "%dx%d, the same initialisation with %d delta + %d Laves particles (R = 5 nm, "
        "seed 7), periodic, nucleation on" % (N, N, n, n))
    del f0
    return {"grid": [N, N], "dtypes": "FP32 storage, FP64 arithmetic (KKS solve included) and "
                                      "compensated FP64 reductions, against FP64 storage",
            "no_precipitates": plain, "with_precipitates": parts}


def trace_summary(path, last=200):
    """Per (kernel, grid) launch counts and durations from a rocprofv3
    --kernel-trace CSV: all launches, and the last `last` launches of each
    group (bench.py's kernel leg is the last run of each step kernel at the
    headline's grid).  Prints a table."""
    import csv
    groups = {}
    with open(path) as f:
        rd = csv.DictReader(f)
        cols = rd.fieldnames or []
        gcols = [c for c in cols if c.startswith("Grid_Size")]
        for row in rd:
            key = (row["Kernel_Name"],) + tuple(row[c] for c in gcols)
            dur = (int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) / 1e3
            groups.setdefault(key, []).append((int(row["Start_Timestamp"]), dur))
    print("%-60s %-18s %6s %9s %9s %9s %12s" % ("kernel", "grid", "calls", "avg_us", "min_us",
                                                 "max_us", "last%d_avg_us" % last))
    for key, v in sorted(groups.items(), key=lambda kv: -sum(d for _, d in kv[1])):
        v.sort()
        d = [x for _, x in v]
        tail = d[-last:]
        print("%-60s %-18s %6d %9.1f %9.1f %9.1f %12.1f" % (
            key[0][:60], "x".join(key[1:]), len(d), sum(d) / len(d), min(d), max(d),
            sum(tail) / len(tail)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the 4096^2 plain workload's run length (about 0.9 s of GPU time)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=600)
    # the GPU's clock takes ~50 launches (~30 ms) of sustained load to settle:
    # four-step launches on scratch copies of the fields before the warm-up
    ap.add_argument("--settle", type=int, default=200)
    ap.add_argument("--kernel-reps", type=int, default=200,
                    help="back-to-back launches of each step kernel for the roofline")
    ap.add_argument("--form", choices=["plain", "precision"], default="plain")
    ap.add_argument("--size", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--cpu-one-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-only", action="store_true")
    ap.add_argument("--only", choices=["dropin", "batched", "serving", "precision", "mobility"],
                    default=None, help="run one of the other sections alone (no headline line)")
    ap.add_argument("--dropin-steps", type=int, default=900,
                    help="steps of the C++ drop-in replay (the driver's call sequence)")
    ap.add_argument("--batch-instances", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=5000)
    ap.add_argument("--batch-tile", type=int, nargs=4, default=None,
                    metavar=("ROWS", "DEPTH", "STEPS", "LPG"),
                    help="batched form: this tile, steps per launch and launches per graph "
                         "instead of the autotuner's")
    ap.add_argument("--loop-steps", type=int, default=20,
                    help="launches of each plain loop (two-step and one-step)")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[6.0, 16.0, 24.0])
    ap.add_argument("--serve-duration", type=float, default=20.0)
    ap.add_argument("--serve-pool-gb", type=int, default=4)
    ap.add_argument("--serve-lpg", type=int, default=17,
                    help="step launches per captured graph in the serving loop")
    ap.add_argument("--serve-rows", default="auto",
                    help="rows per strip task of the serving batch, or 'auto' (16 while few "
                         "cells run, 64 under load)")
    ap.add_argument("--no-srw", action="store_true",
                    help="skip the shortest-remaining-work serving run")
    ap.add_argument("--precision-size", type=int, default=8192)
    ap.add_argument("--precision-marks", type=int, nargs="+", default=[1000, 10000, 20000])
    ap.add_argument("--precision-particles", type=int, default=256,
                    help="delta and Laves particles (each) of the precision form's second "
                         "workload")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only launch the one- to four-step kernels (FP64, FP32) at --size "
                         "a few times each, for rocprofv3 counter runs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a plain device copy of 4 x 4098^2 FP64 arrays (1.07 GB moved per "
                         "dispatch) to scale rocprofv3 FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--trace-summary", default=None,
                    help="summarize a rocprofv3 kernel_trace.csv per kernel and grid, and exit")
    args = ap.parse_args()

    if args.trace_summary:
        trace_summary(args.trace_summary, last=args.kernel_reps)
        return

    if args.gpus != 1:
        print("bench.py: runs on one GPU; --gpus %d is not supported (the y-slab "
              "sharding across GPUs is a design, not built)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        # a plain device copy of the headline's four FP64 fields: the bytes
        # per dispatch scale the rocprofv3 FETCH_SIZE / WRITE_SIZE counters,
        # and its time (HIP events after the clock has settled) is the
        # streaming ceiling the step kernels are compared with
        import torch
        n = 4098 * 4098
        src = [torch.ones(n, dtype=torch.float64, device="cuda") for _ in range(4)]
        dst = [torch.empty_like(x) for x in src]
        for _ in range(50):
            for a, b in zip(dst, src):
                a.copy_(b)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        reps = 50
        e0.record()
        for _ in range(reps):
            for a, b in zip(dst, src):
                a.copy_(b)
        e1.record()
        e1.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / (4 * reps)
        moved = 2 * n * 8
        print(json.dumps({"calibration": "torch copy_ of 4 x %d FP64" % n,
                          "bytes_read_per_dispatch": n * 8, "bytes_written_per_dispatch": n * 8,
                          "dispatches_timed": 4 * reps, "us_per_dispatch": us,
                          "copy_GB_per_s": moved / us / 1e3,
                          "copy_frac_of_peak": moved / (us * 1e-6) / HBM_PEAK}))
        return

    if args.kernels_only:
        # one-, two-, three- and four-step launches at --size, FP64 then
        # FP32, for counter collection (rocprofv3 --pmc); prints one JSON line
        import torch

        from pfpa import model as M
        from pfpa.sim import Simulation
        N = args.size or 4096
        s = M.make_setup(N, N, bc=1)
        f0 = M.init_enriched_band(N, N, 1, seed=42)
        for dt in (torch.float64, torch.float32):
            sim = Simulation(s, f0, dtype=dt)
            for k, n in ((1, 1), (2, 20), (3, 21), (4, 20)):
                sim.steps_per_launch = k
                sim.step(n)
            torch.cuda.synchronize()
            sim.close()
        print(json.dumps({"kernels_only": [N, N], "launches": {
            "one_step": 1, "two_step": 10, "three_step": 7, "four_step": 5}}))
        return

    if args.form == "precision":
        args.only = "precision"
    if args.only:
        # one section alone (for profiling), one JSON line
        args.size = args.size or 4096
        fn = {"dropin": dropin_section, "batched": batched_section,
              "serving": serving_section, "precision": precision_section,
              "mobility": mobility_section}[args.only]
        print(json.dumps({args.only: fn(args)}))
        return

    args.size = args.size or 4096
    t_all = time.perf_counter()
    plain = plain_section(args)
    log("plain: %.4g cell-updates/s, %.1f us/launch %s" %
        (plain["value"], plain["roofline"]["kernel_avg_us"], plain["roofline"]["kernel"]))
    out = {
        "metric": METRIC,
        "value": plain["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": plain["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "plain form, %dx%d interior, dx=0.25nm, periodic, nucleation on "
                        "(Philox seed 42, every %d steps, fused), dt=%.4g s"
                        % (args.size, args.size, plain["nuc_every"], plain["dt"]),
            "grid": [args.size, args.size],
            "form": "plain",
            "parallelism": "single GPU",
        },
        "roofline": plain["roofline"],
        "timing": plain["timing"],
    }
    if not args.plain_only:
        out["dropin"] = dropin_section(args)
        log("dropin:", json.dumps(out["dropin"]))
        out["mobility_divergence"] = mobility_section(args)
        log("mobility:", json.dumps(out["mobility_divergence"]))
        out["batched"] = batched_section(args)
        log("batched:", json.dumps({k: v for k, v in out["batched"].items()
                                    if k != "autotuned"}))
        if not args.no_precision:
            out["precision"] = precision_section(args)
        out["serving"] = serving_section(args, capacity=out["batched"]["value"])
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args)
        cpu["gpu_over_cpu_same_workload"] = plain["value"] / cpu["value"]
        log("cpu:", json.dumps(cpu))
    out["cpu_baseline"] = cpu
    out["bench_wall_s"] = time.perf_counter() - t_all
    print(json.dumps(out))


if __name__ == "__main__":
    main()
