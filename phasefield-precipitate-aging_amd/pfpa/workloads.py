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


def autotune_batch(batch, steps=8, rows_set=(4, 8, 16, 32), depth_set=(2, 3, 6),
                   graph_set=(1, 8)):
    """Time `steps` steps of the current job mix for each tile configuration
    and steps-per-graph; leave the batch on the fastest.  The jobs advance
    during tuning (their targets must allow it).  Returns (best, table)."""
    table = []
    for rows in rows_set:
        for depth in depth_set:
            batch.set_tile(rows, depth)
            batch.commit()
            batch.run(2, 0)            # warm: compile, caches
            batch.sync()
            t0 = time.perf_counter()
            batch.run(steps, 0)
            batch.sync()
            table.append(((time.perf_counter() - t0) / steps, rows, depth, 1))
    table.sort()
    _, rows, depth, _ = table[0]
    batch.set_tile(rows, depth)
    batch.commit()
    best = table[0]
    for spg in graph_set:
        if spg <= 1:
            continue
        batch.run(spg, spg)            # capture
        batch.sync()
        t0 = time.perf_counter()
        batch.run(spg * max(1, steps // spg), spg)
        batch.sync()
        dt = (time.perf_counter() - t0) / (spg * max(1, steps // spg))
        table.append((dt, rows, depth, spg))
        if dt < best[0]:
            best = (dt, rows, depth, spg)
    return best, table
