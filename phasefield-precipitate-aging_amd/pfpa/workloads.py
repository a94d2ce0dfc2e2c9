"""Seeded synthetic workloads of the benchmark and an autotuner for the
batched form."""
import math
import time

import numpy as np

from . import model as M


def batch_specs(count=512, lo=128, hi=1024, seed=7):
    """`count` independent square simulations: side log-uniform in [lo, hi],
    composition uniform in the matrix or in the enriched range, 0-8 seeded
    delta/Laves particles each.  Returns light specs; fields() materialises."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(count):
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        if rng.uniform() < 0.5:
            x0 = (rng.uniform(0.26, 0.29), rng.uniform(0.15, 0.18))   # enriched
        else:
            x0 = (rng.uniform(*M.MATRIX_CR), rng.uniform(*M.MATRIX_NB))
        npart = int(rng.integers(0, 9))
        nd = int(rng.integers(0, npart + 1))
        specs.append(dict(n=n, x0=x0, n_del=nd, n_lav=npart - nd,
                          seed=int(rng.integers(1 << 30)), index=k))
    return specs


def spec_setup(spec, dx=0.25e-9):
    return M.make_setup(spec["n"], spec["n"], dx=dx, bc=1)


def spec_fields(spec, dx=0.25e-9):
    n = spec["n"]
    radius = min(5e-9, n * dx / 12)
    try:
        return M.init_matrix_with_particles(n, n, dx, 1, seed=spec["seed"], xCr0=spec["x0"][0],
                                            xNb0=spec["x0"][1], noise=1e-3,
                                            n_del=spec["n_del"], n_lav=spec["n_lav"],
                                            radius=radius)
    except RuntimeError:
        return M.init_matrix_with_particles(n, n, dx, 1, seed=spec["seed"], xCr0=spec["x0"][0],
                                            xNb0=spec["x0"][1], noise=1e-3, n_del=0, n_lav=0,
                                            radius=radius)


def batch_cell_updates(batch, before):
    """Cell-updates done since `before` (a dict job -> steps from
    step_counts()), from the jobs' own step counters."""
    total = 0
    for j, steps, _ in batch.poll_all():
        ny, nx = batch.shapes[j]
        total += (steps - before.get(j, 0)) * (nx - 2) * (ny - 2)
    return total


def step_counts(batch):
    return {j: steps for j, steps, _ in batch.poll_all()}


def autotune_batch(batch, launches=8, rows_set=(8, 16, 32, 48, 64), depth_set=(2, 3, 6),
                   launch_steps_set=(1, 2, 3), graph_set=(1, 8)):
    """Measure the current job mix's cell-updates/s for each tile shape
    (rows, prefetch depth), time steps per launch (one, or two or three with
    the temporally blocked march) and launches per captured graph; leave the
    batch on the fastest.  Rates come from the jobs' step counters, so launches that
    advance some jobs by one step (nucleation steps, last steps) are counted
    exactly.  The jobs advance during tuning (their targets must allow it).
    Returns (best, table), entries (rate, rows, depth, launch_steps, per_graph)."""
    table = []

    def measure(n, per_graph):
        batch.sync()
        before = step_counts(batch)
        t0 = time.perf_counter()
        batch.run(n, per_graph)
        batch.sync()
        el = time.perf_counter() - t0
        return batch_cell_updates(batch, before) / el

    for ls in launch_steps_set:
        batch.set_launch_steps(ls)
        for rows in rows_set:
            for depth in depth_set:
                batch.set_tile(rows, depth)
                batch.commit()
                batch.run(2, 0)            # warm: caches
                table.append((measure(launches, 0), rows, depth, ls, 1))
    best = max(table)
    _, rows, depth, ls, _ = best
    batch.set_launch_steps(ls)
    batch.set_tile(rows, depth)
    batch.commit()
    for g in graph_set:
        if g <= 1:
            continue
        batch.run(g, g)                    # capture
        rate = measure(g * max(1, launches // g), g)
        table.append((rate, rows, depth, ls, g))
        if rate > best[0]:
            best = (rate, rows, depth, ls, g)
    return best, table
