"""Host-side logic of the package (no GPU): the launch plan of the stepping
loop, the batched and serving workloads."""
import math

import numpy as np
import pytest

from pfpa import model as M
from pfpa.serve import make_requests
from pfpa.sim import plan_launches
from pfpa.workloads import batch_specs


@pytest.mark.parametrize("max_steps", [1, 2, 3, 4])
@pytest.mark.parametrize("nuc_every", [0, 1, 2, 3, 4, 7, 891])
@pytest.mark.parametrize("t0", [0, 1, 5, 890])
def test_plan_launches(t0, max_steps, nuc_every):
    """Every step exactly once, in order; groups never hold a nucleation step;
    every nucleation step runs alone with the pass; groups are as long as the
    rules allow (a shorter group is followed by a nucleation step or the end)."""
    for n in (0, 1, 2, 3, 5, 17, 100):
        plan = plan_launches(t0, n, max_steps, nuc_every)
        assert sum(k for k, _ in plan) == n
        t = t0
        for i, (k, nuc) in enumerate(plan):
            due = [nuc_every > 0 and (t + j) % nuc_every == 0 for j in range(k)]
            assert 1 <= k <= max_steps
            if k > 1:
                assert not any(due) and not nuc
            else:
                assert nuc == due[0]
            if k < max_steps and i + 1 < len(plan):
                # cut short only by a nucleation step within reach
                reach = min(max_steps, n - (t - t0))
                assert any(nuc_every > 0 and (t + j) % nuc_every == 0 for j in range(reach))
            t += k


def _plan_brute_force(t0, n, max_steps, nuc_every):
    """Per launch, the longest group of 4, 3 or 2 steps with no nucleation step
    in it, else one step (with the pass when it is a nucleation step)."""
    def due(t):
        return nuc_every > 0 and t % nuc_every == 0
    out, t, left = [], t0, n
    while left > 0:
        k = 1
        for cand in (4, 3, 2):
            if cand <= max_steps and left >= cand and not any(due(t + j) for j in range(cand)):
                k = cand
                break
        out.append((k, k == 1 and due(t)))
        t += k
        left -= k
    return out


@pytest.mark.parametrize("max_steps", [1, 2, 3, 4])
def test_plan_launches_against_brute_force(max_steps):
    for nuc_every in (0, 1, 2, 3, 4, 5, 7, 891):
        for t0 in (0, 1, 2, 5, 889, 890, 891):
            for n in (0, 1, 2, 3, 4, 7, 50, 2000):
                assert plan_launches(t0, n, max_steps, nuc_every) == \
                    _plan_brute_force(t0, n, max_steps, nuc_every), (t0, n, nuc_every)


def test_plan_launches_schedule_of_the_benchmark():
    """Nucleation every 891 steps from step 0: 1000 steps are one nucleation
    launch, three-step launches and the groups the schedule leaves."""
    plan = plan_launches(0, 1000, 3, 891)
    assert plan[0] == (1, True)
    nuc_steps = []
    t = 0
    for k, nuc in plan:
        if nuc:
            nuc_steps.append(t)
        t += k
    assert nuc_steps == [0, 891]
    assert sum(1 for k, _ in plan if k == 3) >= 330


def test_default_steps_per_launch():
    """Four steps per launch for FP64 storage by default, and for FP32 below
    2560^2 interior cells (three from there); an explicit request is kept where
    the grid and mask allow it, else cut to three (interiors below 4x4) or one
    (below 3x3, or not a symmetric five-point mask)."""
    import torch

    from pfpa.sim import default_steps_per_launch as spl

    big, five, four = M.make_setup(64, 48, bc=1), M.make_setup(3, 9, bc=0), M.make_setup(4, 4, bc=1)
    assert spl(big, torch.float64) == 4 and spl(big, torch.float32) == 4
    huge, edge = M.make_setup(8192, 8192, bc=1), M.make_setup(2560, 2560, bc=1)
    assert spl(huge, torch.float64) == 4 and spl(huge, torch.float32) == 3
    assert spl(edge, torch.float32) == 3 and spl(M.make_setup(2560, 2559, bc=1), torch.float32) == 4
    for k in (1, 2, 3, 4):
        assert spl(big, torch.float64, k) == k and spl(big, torch.float32, k) == k
    assert spl(four, torch.float64) == 4
    assert spl(five, torch.float64) == 3 and spl(five, torch.float64, 4) == 3
    assert spl(five, torch.float64, 2) == 2
    tiny = M.make_setup(2, 9, bc=0)
    assert spl(tiny, torch.float64) == 1
    nine = M.make_setup(64, 48, bc=1)
    nine.mask = [1.0] * 4 + [-8.0] + [1.0] * 4
    assert spl(nine, torch.float64) == 1
    for bad in (0, 5, -1):
        with pytest.raises(ValueError):
            spl(big, torch.float64, bad)


def test_dropin_pending_launches():
    """The launches that run p pending drop-in steps: they add up to p, their
    number has the parity of p (so the result lands in the pointer set the
    caller's swaps put it in), four-step launches come in pairs and only when
    allowed, and a full group is two four-step launches (eight steps) or one
    three-step launch."""
    from pfpa.api import group_size, pending_launches
    for four in (False, True):
        for p in range(0, 40):
            ks = pending_launches(p, four)
            assert sum(ks) == p
            assert len(ks) % 2 == p % 2
            assert all(k in ((1, 3, 4) if four else (1, 3)) for k in ks)
            assert ks.count(4) % 2 == 0
            assert ks == sorted(ks, reverse=True)
            assert ks.count(1) <= 2
    assert pending_launches(8, True) == [4, 4]
    assert pending_launches(7, True) == [3, 3, 1]
    assert pending_launches(3, False) == [3]
    assert group_size(4, 6, 6) == 8 and group_size(4, 5, 100) == 3
    assert group_size(3, 100, 100) == 3


def test_batch_specs():
    specs = batch_specs(512, 128, 1024, seed=7)
    assert len(specs) == 512
    assert specs == batch_specs(512, 128, 1024, seed=7)      # seeded
    n = np.array([sp["n"] for sp in specs])
    assert n.min() >= 128 and n.max() <= 1024
    # log-uniform: median near the geometric mean sqrt(128 * 1024) = 362
    assert 300 < np.median(n) < 440
    cells = (n.astype(float) ** 2).sum()
    assert 1.1e8 < cells < 1.6e8                                # E = 2.48e5 per instance
    for sp in specs:
        assert 0 <= sp["n_del"] + sp["n_lav"] <= 8
        cr, nb = sp["x0"]
        enriched = 0.26 <= cr <= 0.29 and 0.15 <= nb <= 0.18
        matrix = (M.MATRIX_CR[0] <= cr <= M.MATRIX_CR[1] and
                  M.MATRIX_NB[0] <= nb <= M.MATRIX_NB[1])
        assert enriched or matrix


def test_requests_poisson():
    rate, window = 6.0, 200.0
    reqs = make_requests(rate, window, seed=3)
    assert reqs == sorted(reqs, key=lambda r: r.arrival)
    arr = np.array([r.arrival for r in reqs])
    assert arr.min() >= 0 and arr.max() <= window
    # count ~ Poisson(rate * window): within 5 sigma
    assert abs(len(reqs) - rate * window) < 5 * math.sqrt(rate * window)
    gaps = np.diff(arr)
    assert abs(gaps.mean() - 1 / rate) < 0.1 / rate
    for r in reqs:
        assert 128 <= r.n <= 2048 and 500 <= r.steps <= 20000
    assert len({r.seed for r in reqs}) == len(reqs)


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU; --gpus N > 1 exits with a message before any
    GPU work."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stdout + r.stderr
