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
from .batch import PF_STORAGE_F32, Batch, make_desc


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
          offset=True, order="fifo", fill_cells=16 << 20, max_hold_s=2.0, max_bypass_s=1.0,
          auto_rows=(16, 64), auto_cells=(12 << 20, 24 << 20), auto_graphs=(1, 3)):
    """Run the request stream; returns (finished requests, stats).  Each device
    round is one replay of a graph of `launches_per_graph` step launches, each
    advancing every running job by up to `launch_steps` steps (1, 2 or 3).
    read_fields: also keep each request's final fields (r.fields, absolute
    compositions).  offset (FP32 storage): store each request's compositions
    relative to its matrix composition (request_offset), which keeps more bits
    of each step's increment.

    order "fifo": waiting requests are admitted in arrival order.
    order "srw" (shortest remaining work first): waiting requests are admitted
    in order of work (cells x steps), and every round only the running jobs
    with the least remaining work run, as many as fill `fill_cells` cells per
    launch; the others are held for the round (pf_batch_hold), except that a
    job held for `max_hold_s` in a row runs from then on until it is done.
    Results do not depend on the order: a held job resumes where it stopped.

    rows: the strip-task height, or "auto": `auto_rows[0]` (short tasks, more
    parallelism for a few small jobs) while the running jobs hold fewer than
    `auto_cells[1]` cells, `auto_rows[1]` (less halo recomputation, the batched
    form's best throughput) from there until they drop below `auto_cells[0]`;
    in the loaded state a round also replays the graph `auto_graphs[1]` times
    instead of `auto_graphs[0]`, so the host's polling between rounds takes a
    smaller share of the device's time.  The tile never changes a job's
    results.
    In both orders a request that does not fit in the pool lets the ones
    behind it pass, but once it has waited `max_bypass_s` since it first did
    not fit, nothing behind it is admitted until it fits (a large request is
    not starved by a stream of small ones)."""
    if order not in ("fifo", "srw"):
        raise ValueError("order must be 'fifo' or 'srw'")
    import torch
    batch = Batch(storage, pool_bytes=pool_bytes, max_jobs=max(16, len(requests)))
    auto = rows == "auto"
    cur_rows = auto_rows[0] if auto else int(rows)
    graphs = auto_graphs[0] if auto else 1     # graph replays per round
    batch.set_tile(cur_rows, depth)
    batch.set_launch_steps(launch_steps)
    ext = torch.cuda.ExternalStream(batch.stream())
    pending = list(requests)       # sorted by arrival
    waiting = []
    running = {}                   # job id -> request
    progress = {}                  # job id -> steps done at the last poll
    held_since = {}                # job id -> time it was first held (in a row)
    promoted = set()               # jobs held for max_hold_s: never held again
    finished = []
    rejected = []                  # larger than the whole pool
    # rid -> future of (desc, fields), built on a worker thread while the loop
    # keeps the device fed (a 2048^2 request's fields take ~0.2 s on the host)
    prepared = {}
    workers = ThreadPoolExecutor(max_workers=1)
    cell_updates = 0
    busy = 0.0                     # device time of the rounds (HIP events on the batch stream)
    rounds = held_rounds = tile_switches = 0
    blocked_since = {}             # rid -> time it first did not fit
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

    def admit(now):
        """Admit waiting requests while they fit; a request whose fields are
        still being built waits for a later round while the device has work."""
        nonlocal waiting
        if order == "srw":
            waiting.sort(key=lambda q: (q.n * q.n * q.steps, q.arrival))
        still = []
        barrier = False            # a request passed over too long blocks the rest
        for r in waiting:
            if barrier:
                still.append(r)
                continue
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
                    first = blocked_since.setdefault(r.rid, now)
                    if now - first > max_bypass_s:
                        barrier = True
                continue
            del prepared[r.rid]
            blocked_since.pop(r.rid, None)
            r.job = jid
            r.admitted = time.perf_counter() - t0
            running[jid] = r
            progress[jid] = 0
        waiting = still

    while pending or waiting or running:
        now = time.perf_counter() - t0
        if time_limit and now > time_limit:
            break
        while pending and pending[0].arrival <= now:
            waiting.append(pending.pop(0))
        # Under load (the auto tile's loaded state) admissions upload their
        # fields while the round runs and join the next round; otherwise they
        # join this one.
        loaded = auto and cur_rows == auto_rows[1]
        if not loaded or not running:
            admit(now)
        if running and auto:
            cells_running = sum(r.n * r.n for r in running.values())
            want = cur_rows
            if cur_rows == auto_rows[0] and cells_running >= auto_cells[1]:
                want = auto_rows[1]
            elif cur_rows == auto_rows[1] and cells_running < auto_cells[0]:
                want = auto_rows[0]
            if want != cur_rows:
                batch.set_tile(want, depth)
                cur_rows = want
                graphs = auto_graphs[1] if want == auto_rows[1] else auto_graphs[0]
                tile_switches += 1
        if running:
            if order == "srw":
                now = time.perf_counter() - t0
                ranked = sorted(running, key=lambda j: (running[j].n ** 2 *
                                                        (running[j].steps - progress[j]),
                                                        running[j].arrival))
                cells = 0
                for j in ranked:
                    # a job held too long runs from then on until it is done
                    if j in held_since and now - held_since[j] > max_hold_s:
                        promoted.add(j)
                    run = cells < fill_cells or j in promoted
                    if run:
                        cells += running[j].n ** 2
                        if j in held_since:
                            batch.hold(j, False)
                            del held_since[j]
                    else:
                        if j not in held_since:
                            batch.hold(j, True)
                            held_since[j] = now
                        held_rounds += 1
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(ext)
            batch.run(graphs * launches_per_graph, launches_per_graph)   # asynchronous
            e1.record(ext)
            rounds += 1
            now = time.perf_counter() - t0
            if loaded:
                while pending and pending[0].arrival <= now:
                    waiting.append(pending.pop(0))
                admit(now)
            # start building the next admissions' fields
            for r in waiting[:4] + [q for q in pending[:8] if q.arrival <= now + 0.5]:
                prepare(r)
            states = batch.poll_all()     # waits for the round
            busy += e0.elapsed_time(e1) / 1e3
            for jid, steps, active in states:
                r = running.get(jid)
                if r is None:
                    continue
                progress[jid] = steps
                if active:
                    continue
                r.summary = batch.summarize(jid)
                if read_fields:
                    r.fields = batch.read(jid)
                r.done = time.perf_counter() - t0
                cell_updates += r.n * r.n * r.steps
                batch.remove(jid)
                del running[jid]
                progress.pop(jid, None)
                held_since.pop(jid, None)
                promoted.discard(jid)
                finished.append(r)
                if log:
                    log(r)
        elif pending:
            # nothing on the device: get the next arrival ready and wait for it
            prepare(pending[0])
            time.sleep(max(0.0, min(0.01, pending[0].arrival - (time.perf_counter() - t0))))
    wall = time.perf_counter() - t0
    workers.shutdown(wait=True, cancel_futures=True)
    # nucleation events wider than one cell are counted, not applied, in batch
    # mode (they need a domain of a few cells); report them with the results
    wide = batch.status(raise_on_wide=False)
    batch.close()
    lat = np.array([r.done - r.arrival for r in finished]) if finished else np.zeros(1)
    by_size = {}
    for name, sel in (("side<=512", lambda r: r.n <= 512), ("side>512", lambda r: r.n > 512)):
        ls = np.array([r.done - r.arrival for r in finished if sel(r)])
        by_size[name] = ({"count": int(ls.size), "p50_s": float(np.percentile(ls, 50)),
                          "p99_s": float(np.percentile(ls, 99))} if ls.size else {"count": 0})
    window = max((r.arrival for r in requests), default=0.0)
    stats = {
        "requests": len(requests),
        "completed": len(finished),
        "rejected": len(rejected),
        "order": order,
        "wall_s": wall,
        "cell_updates": cell_updates,
        "throughput_cell_updates_per_s": cell_updates / wall if wall > 0 else 0.0,
        "offered_cell_updates_per_s": (sum(r.n * r.n * r.steps for r in requests) / window
                                       if window > 0 else None),
        "device_busy_s": busy,
        "device_busy_share": busy / wall if wall > 0 else 0.0,
        "rounds": rounds,
        "held_job_rounds": held_rounds,
        "rows": rows,
        "tile_switches": tile_switches,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_max_s": float(lat.max()),
        "latency_by_size": by_size,
        "wide_nucleation_events": int(wide),
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
