"""Benchmark of the KKS phase-field step on one MI355X.

Default workload (the plain form): 4096 x 4096 interior cells, FP64,
dx = 0.25 nm, periodic, seeded uniform matrix composition in
Ni-(0.0202..0.0269)Nb-(0.2794..0.3288)Cr with a central enriched strip
(1/8 of the width) at Ni-0.169Nb-0.272Cr, nucleation on (Philox seed 42,
every floor(1e-4/dt) steps, fused into the step), dt = 0.1 dx^2 / (4 D_max).

A "step" is one fused launch over the whole grid (the nucleation pass rides
along on its steps).  value = interior cells x timed steps / wall time of the
timed loop, fields resident in HBM.  roofline: algorithmic bytes of the step
kernel (64 B per cell-update in FP64: 4 fields read + 4 written) over its
average launch duration measured with HIP events on the launch stream.
cpu_baseline: the FP64 CPU oracle (OpenMP) on this host, on a bounded sample
of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form plain|precision]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "phasefield-precipitate-aging_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s


def load_traffic(workload):
    """HBM bytes per step-kernel launch from the committed rocprofv3 counter
    runs (profiles/traffic_<workload>.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def cpu_baseline(N, steps, warmup=1):
    """Time the FP64 oracle (all host cores via OpenMP) on N x N for `steps`."""
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--form", choices=["plain", "precision"], default="plain")
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a plain device copy of 4 x 4098^2 FP64 arrays (1.07 GB moved per "
                         "dispatch) to scale rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        n = 4098 * 4098
        src = [torch.ones(n, dtype=torch.float64, device="cuda") for _ in range(4)]
        dst = [torch.empty_like(x) for x in src]
        for _ in range(5):
            for a, b in zip(dst, src):
                a.copy_(b)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "torch copy_ of 4 x %d FP64", "bytes_read_per_dispatch":
                          n * 8, "bytes_written_per_dispatch": n * 8}))
        return

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is described in DESIGN.md, not built)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import ctypes

    import torch

    from pfpa import model as M
    from pfpa._lib import PfParams, check, lib
    from pfpa.sim import Simulation

    N = args.size
    dtype = torch.float64 if args.form == "plain" else torch.float32
    s = M.make_setup(N, N, bc=1)
    f0 = M.init_enriched_band(N, N, 1, seed=42)
    sim = Simulation(s, f0, dtype=dtype, nucleation=True, seed=42)
    stream = sim.stream

    sim.step(args.warmup)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    sim.step(args.steps)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1)
    cells = N * N
    value = cells * args.steps / (ms / 1e3)

    # dominant kernel alone: back-to-back step launches (no nucleation steps)
    P = sim.P
    a, b = sim.a, sim.b
    pt = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    st = ctypes.c_void_p(stream.cuda_stream)
    fn = lib().pf_step_f64 if dtype == torch.float64 else lib().pf_step_f32
    reps = max(20, min(200, args.steps))
    k0 = torch.cuda.Event(enable_timing=True)
    k1 = torch.cuda.Event(enable_timing=True)
    k0.record(stream)
    for i in range(reps):
        x, y = (a, b) if i % 2 == 0 else (b, a)
        check(fn(sim.ctx, s.nx, s.ny, s.bc, ctypes.byref(P), *[pt(t) for t in x],
                 *[pt(t) for t in y], None, st), "step")
    k1.record(stream)
    torch.cuda.synchronize()
    kern_s = k0.elapsed_time(k1) / 1e3 / reps
    bpc = 64 if dtype == torch.float64 else 32
    alg_bytes = bpc * cells
    achieved = alg_bytes / kern_s
    workload = "plain_fp64_%d" % N if args.form == "plain" else "precision_fp32_%d" % N
    traffic = load_traffic(workload)

    mass = sim.summarize()
    sim.close()

    cpu = None
    if not args.no_cpu:
        cpu_N = 4096 if N >= 4096 else N
        v, cores, el = cpu_baseline(cpu_N, args.cpu_steps)
        cpu = {"value": v, "unit": "cell-updates/s", "cores": cores, "kind": "port",
               "sample": "%d steps of the %dx%d FP64 plain step (same init, periodic) after "
                         "1 warm-up step, %.1f s" % (args.cpu_steps, cpu_N, cpu_N, el)}

    out = {
        "metric": "cell-updates/s (% HBM roofline) plain vs batched; p50/p99 sweep latency",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if dtype == torch.float64 else "f32-storage/f64-compute",
        "data": "synthetic",
        "config": {
            "workload": "%s form, %dx%d interior, dx=0.25nm, periodic, nucleation on "
                        "(Philox seed 42, every %d steps, fused), dt=%.4g s"
                        % (args.form, N, N, sim.nuc_every, s.dt),
            "grid": [N, N],
            "form": args.form,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "kernel": "pf::step_kernel (fused step)",
            "kernel_avg_us": kern_s * 1e6,
            "algorithmic_bytes_per_launch": alg_bytes,
        },
        "cpu_baseline": cpu,
        "wall_s": wall,
        "mean_fields": mass,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
