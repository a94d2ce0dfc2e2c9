"""Benchmark of the KKS phase-field step on one MI355X.

Headline (value): the plain form on 4096 x 4096 interior cells, FP64,
dx = 0.25 nm, periodic, seeded uniform matrix composition in
Ni-(0.0202..0.0269)Nb-(0.2794..0.3288)Cr with a central enriched strip (1/8 of
the width) at Ni-0.169Nb-0.272Cr, nucleation on (Philox seed 42, every
floor(1e-4/dt) steps, fused into the step), dt = 0.1 dx^2 / (4 D_max).
A "step" is one explicit Euler step over the whole grid; steps run three per
launch (the temporally blocked kernel) except around nucleation steps, which
run one per launch with the nucleation pass fused.  value = interior cells x
timed steps / time of the timed loop (HIP events), fields resident in HBM.
roofline: the dominant kernel is the three-step kernel; its algorithmic bytes
per launch (FP64: 4 fields read + 4 written = 64 B per cell, for three
cell-updates) over its average launch duration (HIP events on the launch
stream); the two-step and one-step kernels are reported beside it.  traffic:
HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs under
profiles/ (scaled by a copy kernel).
cpu_baseline: the FP64 CPU oracle (OpenMP, all host threads) on a bounded
sample of the same workload.

Also reported (same JSON line):
  dropin   -- the reference driver's per-step call sequence through the C++
              drop-in API (tests/cpp/replay_main) on the same grid (no-flux,
              nucleation on, no I/O), grouped three steps per launch and one.
  batched  -- 512 independent simulations, square sides log-uniform in
              128..1024, FP32 storage with FP64 arithmetic and reductions, one
              persistent launch per one, two or three steps replayed from a
              hipGraph; tile, steps per launch and launches per graph from the
              autotuner; against loops of plain FP64 launches (three-, two-
              and one-step) over the same instances; phase-fraction error
              against FP64 replays of sampled instances.
  serving  -- Poisson sweep requests (side log-uniform 128..2048, 500..20000
              steps, random composition and nucleation seed) with continuous
              batching under an HBM budget; throughput and p50/p99 latency.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form plain|precision]
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
    """Model name, logical CPUs visible and socket count of the host."""
    model, sockets = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "model name" and model == "unknown":
                    model = v.strip()
                elif k == "physical id":
                    sockets.add(v.strip())
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "sockets": len(sockets) or None}


def cpu_baseline(N, steps, warmup=1):
    """Time the FP64 oracle (all host threads via OpenMP) on N x N for `steps`."""
    import oracle as O
    from pfpa import model as M
    s = M.make_setup(N, N, bc=1)
    f0 = M.init_enriched_band(N, N, 1, seed=42)
    P = O.make_params(s.dt, s.DCr, s.DNb, s.kappa, s.omega, s.Lmob, s.alpha, s.mask)
    O.run(f0, P, 1, warmup)
    t0 = time.perf_counter()
    O.run(f0, P, 1, steps)
    el = time.perf_counter() - t0
    return N * N * steps / el, O.num_threads(), el


def cpu_baseline_single(N, steps):
    """The same on one thread (a shorter sample); restores the thread count."""
    import oracle as O
    from pfpa import model as M
    s = M.make_setup(N, N, bc=1)
    f0 = M.init_enriched_band(N, N, 1, seed=42)
    P = O.make_params(s.dt, s.DCr, s.DNb, s.kappa, s.omega, s.Lmob, s.alpha, s.mask)
    threads = O.num_threads()
    O.set_threads(1)
    try:
        t0 = time.perf_counter()
        O.run(f0, P, 1, steps)
        el = time.perf_counter() - t0
    finally:
        O.set_threads(threads)
    return N * N * steps / el, el


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
    s = M.make_setup(N, N, bc=1)
    f0 = M.init_enriched_band(N, N, 1, seed=42)
    sim = Simulation(s, f0, dtype=dtype, nucleation=True, seed=42)
    stream = sim.stream
    sim.step(args.warmup)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    sim.step(args.steps)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    cells = N * N
    value = cells * args.steps / (ms / 1e3)

    # The kernels alone, back to back on the sim's stream: the three-step kernel
    # (the dominant one), the two-step kernel (what the schedule leaves between
    # nucleation steps) and the one-step kernel (nucleation steps; the drop-in
    # API with grouping off).
    P = sim.P
    a, b = sim.a, sim.b
    st = ctypes.c_void_p(stream.cuda_stream)
    f64 = dtype == torch.float64
    reps = max(20, min(200, args.steps))

    def time_kernel(fn, nuc_arg):
        k0 = torch.cuda.Event(enable_timing=True)
        k1 = torch.cuda.Event(enable_timing=True)
        k0.record(stream)
        for i in range(reps):
            x, y = (a, b) if i % 2 == 0 else (b, a)
            args_ = [sim.ctx, s.nx, s.ny, s.bc, ctypes.byref(P), *[_ptr(t) for t in x],
                     *[_ptr(t) for t in y]] + ([None] if nuc_arg else []) + [st]
            check(fn(*args_), "step")
        k1.record(stream)
        torch.cuda.synchronize()
        return k0.elapsed_time(k1) / 1e3 / reps

    kern3_s = time_kernel(lib().pf_step3_f64 if f64 else lib().pf_step3_f32, False)
    kern2_s = time_kernel(lib().pf_step2_f64 if f64 else lib().pf_step2_f32, False)
    kern1_s = time_kernel(lib().pf_step_f64 if f64 else lib().pf_step_f32, True)
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
    valu_us = load_valu_issue("pf::stepk_kernel<%s, 1, 3, 3>" % ("double" if f64 else "float")) \
        if N == 4096 else None
    return {
        "value": value, "ms_per_step": ms / args.steps, "N": N, "nuc_every": nuc_every,
        "dt": s.dt,
        "roofline": {
            "bound": "hbm", "achieved": alg_bytes / kern3_s / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": alg_bytes / kern3_s / HBM_PEAK,
            "traffic": load_traffic("plain3_%s_%d" % (tag, N)),
            "kernel": "pf::stepk_kernel<K=3> (three fused steps per launch)",
            "kernel_avg_us": kern3_s * 1e6,
            "algorithmic_bytes_per_launch": alg_bytes,
            "cell_updates_per_launch": 3 * cells,
            "algorithmic_bytes_per_cell_update": bpc / 3,
            # the rate per step against the one-step roofline (bpc bytes per
            # cell-update at peak HBM), and the share of the launch the VALU
            # instructions alone need (rocprofv3 SQ_INSTS_VALU, profiles/)
            "cell_update_rate_vs_one_step_roofline": value * bpc / HBM_PEAK,
            "valu_issue_us": valu_us,
            "valu_issue_frac": (valu_us / (kern3_s * 1e6)) if valu_us else None,
            "two_step_kernel": other("pf::stepk_kernel<K=2> (two fused steps per launch)", 2,
                                     kern2_s),
            "one_step_kernel": other("pf::step_kernel (one fused step per launch)", 1, kern1_s),
        },
    }


def batched_section(args):
    """512 uneven simulations: batched FP32 vs a loop of plain FP64 launches."""
    import torch

    from pfpa._lib import PfStepItem, check, lib
    from pfpa.batch import PF_STORAGE_F32, Batch, job_bytes, make_desc
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

    # loops of plain FP64 launches over the same instances, issued from C, at
    # three, two and one steps per launch; the fastest is the comparison
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
    for name, fn, per in (("three_step", lib().pf_step3_loop_f64, 3),
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
        "autotuned": tuned,
        "phase_fraction_max_abs_error_vs_fp64": max(errs),
        "phase_fraction_sample": {"jobs": sample, "steps": [after[k] for k in sample]},
    }


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
        for spl in (3, 1):
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


def serving_section(args):
    from pfpa.serve import fp64_replay_error, make_requests, serve
    reqs = make_requests(args.serve_rate, args.serve_duration, seed=1)
    done = [0]

    def progress(r):   # one line per 100 completions (long runs stay visibly alive)
        done[0] += 1
        if done[0] % 100 == 0:
            log("serving: %d of %d requests done" % (done[0], len(reqs)))

    finished, stats = serve(reqs, pool_bytes=args.serve_pool_gb << 30,
                            launches_per_graph=args.serve_lpg, rows=args.serve_rows,
                            depth=3, launch_steps=3, log=progress)
    sample = sorted(finished, key=lambda r: r.n * r.n * r.steps)[:3]
    err, _ = fp64_replay_error(finished, sample)
    stats["arrival_rate_per_s"] = args.serve_rate
    stats["arrival_window_s"] = args.serve_duration
    stats["phase_fraction_max_abs_error_vs_fp64"] = err
    stats["fp64_replayed"] = len(sample)
    return stats


def precision_section(args):
    """FP32 storage, without and with a composition offset (the median
    composition), vs FP64 on 8192^2 (same init as the plain case): max field
    differences, phase fractions and mass drift at 1k/10k/20k steps."""
    import torch

    from pfpa import model as M
    from pfpa.sim import Simulation
    N = args.size
    s = M.make_setup(N, N, bc=1)
    f0 = M.init_enriched_band(N, N, 1, seed=42)
    s64 = Simulation(s, f0, dtype=torch.float64, nucleation=True, seed=42)
    s32 = Simulation(s, f0, dtype=torch.float32, nucleation=True, seed=42)
    s32o = Simulation(s, f0, dtype=torch.float32, nucleation=True, seed=42,
                      storage_offset="median")
    del f0
    m0 = s64.summarize()
    rows = []
    done = 0
    secs = {64: 0.0, 32: 0.0, "32o": 0.0}   # GPU time of each form's stepping (HIP events)

    def timed_steps(sim, n, key):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(sim.stream)
        sim.step(n)
        e1.record(sim.stream)
        e1.synchronize()
        secs[key] += e0.elapsed_time(e1) / 1e3

    for mark in args.precision_marks:
        timed_steps(s64, mark - done, 64)
        timed_steps(s32, mark - done, 32)
        timed_steps(s32o, mark - done, "32o")
        done = mark

        def max_dx(sim, k):
            return float((s64.a[k] - (sim.a[k].double() + sim.xoff[k])).abs().max())

        r64, r32, r32o = s64.summarize(), s32.summarize(), s32o.summarize()
        rows.append({"step": mark, "max_abs_dx_Cr": max_dx(s32, 0),
                     "max_abs_dx_Nb": max_dx(s32, 1),
                     "max_abs_dx_Cr_offset": max_dx(s32o, 0),
                     "max_abs_dx_Nb_offset": max_dx(s32o, 1),
                     "delta_frac_64": r64[3], "delta_frac_32": r32[3],
                     "delta_frac_32_offset": r32o[3],
                     "laves_frac_64": r64[4], "laves_frac_32": r32[4],
                     "laves_frac_32_offset": r32o[4],
                     "mass_drift_Cr_64": r64[0] - m0[0], "mass_drift_Cr_32": r32[0] - m0[0],
                     "mass_drift_Cr_32_offset": r32o[0] - m0[0],
                     "mass_drift_Nb_64": r64[1] - m0[1], "mass_drift_Nb_32": r32[1] - m0[1],
                     "mass_drift_Nb_32_offset": r32o[1] - m0[1]})
        log("precision", json.dumps(rows[-1]))
    offset = list(s32o.xoff)
    s64.close()
    s32.close()
    s32o.close()
    rates = {"fp64": N * N * done / secs[64], "fp32_storage": N * N * done / secs[32],
             "fp32_storage_offset": N * N * done / secs["32o"], "offset": offset,
             "unit": "cell-updates/s", "steps": done}
    log("precision rates", json.dumps(rates))
    return rows, rates


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the 4096^2 plain workload's run length (about 0.9 s of GPU time)
    ap.add_argument("--steps", type=int, default=10000)
    # the GPU's clock takes about 30 ms of sustained load to settle (the first
    # launches of a cold run are up to 30% slower): 600 steps are ~55 ms
    ap.add_argument("--warmup", type=int, default=600)
    ap.add_argument("--form", choices=["plain", "precision"], default="plain")
    ap.add_argument("--size", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-only", action="store_true")
    ap.add_argument("--only", choices=["dropin", "batched", "serving"], default=None,
                    help="run one of the other sections alone (no headline line)")
    ap.add_argument("--dropin-steps", type=int, default=900,
                    help="steps of the C++ drop-in replay (the driver's call sequence)")
    ap.add_argument("--batch-instances", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=5000)
    ap.add_argument("--loop-steps", type=int, default=20,
                    help="launches of each plain loop (two-step and one-step)")
    ap.add_argument("--serve-rate", type=float, default=6.0)
    ap.add_argument("--serve-duration", type=float, default=20.0)
    ap.add_argument("--serve-pool-gb", type=int, default=4)
    ap.add_argument("--serve-lpg", type=int, default=17,
                    help="step launches per captured graph in the serving loop")
    ap.add_argument("--serve-rows", type=int, default=16,
                    help="rows per strip task of the serving batch")
    ap.add_argument("--precision-marks", type=int, nargs="+", default=[1000, 10000, 20000])
    ap.add_argument("--kernels-only", action="store_true",
                    help="only launch the one-, two- and three-step kernels (FP64, FP32) at --size "
                         "a few times each, for rocprofv3 counter runs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a plain device copy of 4 x 4098^2 FP64 arrays (1.07 GB moved per "
                         "dispatch) to scale rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: runs on one GPU; --gpus %d is not supported (the y-slab "
              "sharding across GPUs is a design, not built)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        import torch
        n = 4098 * 4098
        src = [torch.ones(n, dtype=torch.float64, device="cuda") for _ in range(4)]
        dst = [torch.empty_like(x) for x in src]
        for _ in range(5):
            for a, b in zip(dst, src):
                a.copy_(b)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "torch copy_ of 4 x %d FP64" % n,
                          "bytes_read_per_dispatch": n * 8, "bytes_written_per_dispatch": n * 8}))
        return

    if args.kernels_only:
        # one-, two- and three-step launches at --size, FP64 then FP32, for
        # counter collection (rocprofv3 --pmc); prints one JSON line
        import torch

        from pfpa import model as M
        from pfpa.sim import Simulation
        N = args.size or 4096
        s = M.make_setup(N, N, bc=1)
        f0 = M.init_enriched_band(N, N, 1, seed=42)
        for dt in (torch.float64, torch.float32):
            sim = Simulation(s, f0, dtype=dt)
            for k, n in ((1, 1), (2, 20), (3, 21)):
                sim.steps_per_launch = k
                sim.step(n)
            torch.cuda.synchronize()
            sim.close()
        print(json.dumps({"kernels_only": [N, N], "launches": {"one_step": 1, "two_step": 10,
                                                               "three_step": 7}}))
        return

    if args.only:
        # one section alone (for profiling), one JSON line
        args.size = args.size or 4096
        fn = {"dropin": dropin_section, "batched": batched_section,
              "serving": serving_section}[args.only]
        print(json.dumps({args.only: fn(args)}))
        return

    if args.form == "precision":
        args.size = args.size or 8192
        rows, rates = precision_section(args)
        print(json.dumps({"form": "precision", "grid": [args.size, args.size], "rows": rows,
                          "rates": rates}))
        return

    args.size = args.size or 4096
    t_all = time.perf_counter()
    plain = plain_section(args)
    log("plain: %.4g cell-updates/s, %.1f us/step kernel" %
        (plain["value"], plain["roofline"]["kernel_avg_us"]))
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
    }
    if not args.plain_only:
        out["dropin"] = dropin_section(args)
        log("dropin:", json.dumps(out["dropin"]))
        out["batched"] = batched_section(args)
        log("batched:", json.dumps(out["batched"]))
        out["serving"] = serving_section(args)
        log("serving:", json.dumps(out["serving"]))
    cpu = None
    if not args.no_cpu:
        cpu_N = 4096 if args.size >= 4096 else args.size
        v, cores, el = cpu_baseline(cpu_N, args.cpu_steps)
        cpu = {"value": v, "unit": "cell-updates/s", "cores": cores, "kind": "port",
               "sample": "%d steps of the %dx%d FP64 plain step (same init, periodic) after "
                         "1 warm-up step, %.1f s" % (args.cpu_steps, cpu_N, cpu_N, el),
               "host_cpu": host_cpu()}
        v1, el1 = cpu_baseline_single(cpu_N, 2)
        cpu["single_thread"] = {"value": v1, "unit": "cell-updates/s", "cores": 1,
                                "sample": "2 steps of the same %dx%d step, %.1f s"
                                          % (cpu_N, cpu_N, el1)}
    out["cpu_baseline"] = cpu
    out["bench_wall_s"] = time.perf_counter() - t_all
    print(json.dumps(out))


if __name__ == "__main__":
    main()
