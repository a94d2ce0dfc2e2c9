"""Serving loop: composition / size sweep requests arrive as a Poisson stream
and are run with continuous batching on one GPU.

Each request is one simulation (square grid, step count, composition,
nucleation seed).  The loop admits waiting requests into the batch while the
device pool (the HBM budget) has room, replays the batched step graph, polls
job states, retires finished jobs (their summary is read, their memory freed)
and admits more -- requests that arrive while others run join as slots and
memory free up, and their work overlaps on the device.  Completion latency
is measured from arrival to retirement.
"""
import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import model as M
from .batch import PF_STORAGE_F32, Batch, job_bytes, make_desc


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the run
    n: int                  # grid side
    steps: int
    x0: tuple               # (x_Cr, x_Nb) of the matrix
    seed: int
    admitted: float = -1.0
    done: float = -1.0
    summary: list = field(default_factory=list)
    job: int = -1
    fields: list = None     # final fields, when serve(read_fields=True)


def make_requests(rate, duration, seed=1, n_lo=128, n_hi=2048, steps_lo=500, steps_hi=20000,
                  max_requests=None):
    """Poisson arrivals at `rate` per second over `duration` seconds; grid
    side log-uniform in [n_lo, n_hi], steps uniform in [steps_lo, steps_hi],
    composition uniform in the matrix or enriched range, random nucleation
    seed."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration or (max_requests and len(reqs) >= max_requests):
            break
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        steps = int(rng.integers(steps_lo, steps_hi + 1))
        if rng.uniform() < 0.5:
            x0 = (rng.uniform(0.26, 0.29), rng.uniform(0.15, 0.18))
        else:
            x0 = (rng.uniform(*M.MATRIX_CR), rng.uniform(*M.MATRIX_NB))
        reqs.append(Request(len(reqs), t, n, steps, x0, int(rng.integers(1 << 62))))
    return reqs


def request_fields(r, dx=0.25e-9):
    # a uniform matrix with seeded noise and two seeded particles, so the
    # phase fractions of the answer are not trivially zero
    radius = min(5e-9, r.n * dx / 12)
    return M.init_matrix_with_particles(r.n, r.n, dx, 1, seed=r.seed % (1 << 31), xCr0=r.x0[0],
                                        xNb0=r.x0[1], noise=1e-3, n_del=1, n_lav=1,
                                        radius=radius)


def request_desc(r, dx=0.25e-9):
    s = M.make_setup(r.n, r.n, dx=dx, bc=1)
    return s, make_desc(s, r.steps, nucleation=True, seed=r.seed, instance=r.rid)


def request_offset(r):
    """The composition offset a request's FP32 arrays are stored relative to:
    its matrix composition (known from the request, so admission computes
    nothing for it)."""
    return (float(r.x0[0]), float(r.x0[1]))


def serve(requests, pool_bytes, launches_per_graph=17, rows=16, depth=3, launch_steps=3,
          log=None, storage=PF_STORAGE_F32, time_limit=None, read_fields=False,
          offset=True):
    """Run the request stream; returns (finished requests, stats).  Each device
    round is one replay of a graph of `launches_per_graph` step launches, each
    advancing every running job by up to `launch_steps` steps (1, 2 or 3).
    read_fields: also keep each request's final fields (r.fields, absolute
    compositions).  offset (FP32 storage): store each request's compositions
    relative to its matrix composition (request_offset), which keeps more bits
    of each step's increment."""
    batch = Batch(storage, pool_bytes=pool_bytes, max_jobs=max(16, len(requests)))
    batch.set_tile(rows, depth)
    batch.set_launch_steps(launch_steps)
    pending = list(requests)       # sorted by arrival
    waiting = []
    running = {}                   # job id -> request
    finished = []
    rejected = []                  # larger than the whole pool
    # rid -> future of (desc, fields), built on a worker thread while the loop
    # keeps the device fed (a 2048^2 request's fields take ~0.2 s on the host)
    prepared = {}
    workers = ThreadPoolExecutor(max_workers=1)
    cell_updates = 0
    t0 = time.perf_counter()

    def build(r):
        s, d = request_desc(r)
        f = request_fields(r)
        if offset and storage == PF_STORAGE_F32:
            d.x_offset[0], d.x_offset[1] = request_offset(r)
        return d, f

    def prepare(r):
        if r.rid not in prepared:
            prepared[r.rid] = workers.submit(build, r)
        return prepared[r.rid]

    while pending or waiting or running:
        now = time.perf_counter() - t0
        if time_limit and now > time_limit:
            break
        while pending and pending[0].arrival <= now:
            waiting.append(pending.pop(0))
        # admit in arrival order; a request that does not fit lets smaller ones
        # pass, and one whose fields are still being built waits for a later
        # round while the device has work
        still = []
        for r in waiting:
            fut = prepare(r)
            if running and not fut.done():
                still.append(r)
                continue
            d, f = fut.result()
            jid = batch.add(d, f)
            if jid is None:
                if not running:
                    # the pool is empty and still too small: it never fits
                    del prepared[r.rid]
                    rejected.append(r)
                else:
                    still.append(r)
                continue
            del prepared[r.rid]
            r.job = jid
            r.admitted = time.perf_counter() - t0
            running[jid] = r
        waiting = still
        if running:
            batch.run(launches_per_graph, launches_per_graph)   # asynchronous
            # start building the next admissions' fields
            now = time.perf_counter() - t0
            for r in waiting[:4] + [q for q in pending[:8] if q.arrival <= now + 0.5]:
                prepare(r)
            states = batch.poll_all()
            for jid, steps, active in states:
                r = running.get(jid)
                if r is None or active:
                    continue
                r.summary = batch.summarize(jid)
                if read_fields:
                    r.fields = batch.read(jid)
                r.done = time.perf_counter() - t0
                cell_updates += r.n * r.n * r.steps
                batch.remove(jid)
                del running[jid]
                finished.append(r)
                if log:
                    log(r)
        elif pending:
            # nothing on the device: get the next arrival ready and wait for it
            prepare(pending[0])
            time.sleep(max(0.0, min(0.01, pending[0].arrival - (time.perf_counter() - t0))))
    wall = time.perf_counter() - t0
    workers.shutdown(wait=True, cancel_futures=True)
    batch.status()
    batch.close()
    lat = np.array([r.done - r.arrival for r in finished]) if finished else np.zeros(1)
    stats = {
        "requests": len(requests),
        "completed": len(finished),
        "rejected": len(rejected),
        "wall_s": wall,
        "cell_updates": cell_updates,
        "throughput_cell_updates_per_s": cell_updates / wall if wall > 0 else 0.0,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_max_s": float(lat.max()),
        "launches_per_graph": launches_per_graph,
        "launch_steps": launch_steps,
        "pool_bytes": pool_bytes,
    }
    return finished, stats


def fp64_replay_error(requests, sample, dx=0.25e-9):
    """Replay `sample` finished requests alone in the plain FP64 path and
    return the largest |delta| and |Laves| phase-fraction difference."""
    import torch  # noqa: F401
    from .sim import Simulation
    worst = 0.0
    rows = []
    for r in sample:
        s, _ = request_desc(r, dx)
        sim = Simulation(s, request_fields(r, dx), dtype=torch.float64, nucleation=True,
                         seed=r.seed, instance=r.rid)
        sim.step(r.steps)
        ref = sim.summarize()
        sim.close()
        e = max(abs(ref[3] - r.summary[3]), abs(ref[4] - r.summary[4]))
        rows.append({"rid": r.rid, "n": r.n, "steps": r.steps, "delta64": ref[3],
                     "delta32": r.summary[3], "laves64": ref[4], "laves32": r.summary[4],
                     "err": e})
        worst = max(worst, e)
    return worst, rows
