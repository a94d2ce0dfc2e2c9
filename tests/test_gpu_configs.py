"""The batched (512 instances, sides 128..1024) and serving (sides 128..2048)
workloads at their real sizes on the GPU, and an FP64-storage batch against
the CPU oracle directly.

- The 512-instance mix the benchmark runs (batch_specs(512, seed=7)) plus one
  1024^2 job, FP32 storage with per-job composition offsets, three steps per
  launch replayed from a captured hipGraph: the smallest, the median and the
  1024^2 job are bit-equal to the plain FP32 path run alone.
- An FP64-storage batch (a 1024^2 no-flux job among others) against
  oracle.run at the FP64 field tolerance |a - b| <= 1e-12 max(|a|, |b|, s_f).
- Serving requests drawn from the serving ranges (a 2048^2 and a 128^2
  request among others): each result is the same bits alone and in the mix.
"""
import numpy as np
import pytest
import torch

from pfpa import model as M
from pfpa.batch import PF_STORAGE_F32, PF_STORAGE_F64, Batch, job_bytes, make_desc
from pfpa.serve import Request, serve
from pfpa.sim import Simulation
from pfpa.workloads import batch_specs, spec_fields, spec_setup

pytestmark = pytest.mark.gpu

TOL = 1e-12


def mixed_err(a, b):
    a = a[1:-1, 1:-1]
    b = b[1:-1, 1:-1]
    sf = max(np.abs(b).max(), 1e-300)
    return float((np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), sf)).max())


def test_batched_512_instance_mix_equals_plain(gpu):
    specs = batch_specs(512, seed=7)
    # the top of the size range: one 1024^2 job with particles
    specs.append(dict(n=1024, x0=(0.30, 0.024), n_del=3, n_lav=3, seed=12345,
                      index=len(specs)))
    pool = sum(job_bytes(PF_STORAGE_F32, sp["n"], sp["n"]) for sp in specs) + (64 << 20)
    b = Batch(PF_STORAGE_F32, pool_bytes=pool, max_jobs=len(specs))
    b.set_tile(48, 3)
    b.set_launch_steps(3)
    inputs = {}
    order = sorted(range(len(specs)), key=lambda k: specs[k]["n"])
    picks = {"smallest": order[0], "median": order[len(order) // 2], "n1024": len(specs) - 1}
    print("picked jobs", {k: (v, specs[v]["n"]) for k, v in picks.items()}, flush=True)
    ids = {}
    for sp in specs:
        s = spec_setup(sp)
        f = spec_fields(sp)
        jid = b.add(make_desc(s, 1000, nucleation=True, seed=1000 + sp["index"],
                              instance=sp["index"], x_offset=sp["x0"]), f)
        assert jid is not None
        ids[sp["index"]] = jid
        if sp["index"] in picks.values():
            inputs[sp["index"]] = (s, f)
    assert specs[picks["n1024"]]["n"] == 1024
    # 8 launches: two replays of a captured graph of four launches
    b.run(8, 4)
    b.sync()
    for name, k in picks.items():
        steps, active = b.poll(ids[k])
        assert active and steps >= 16, (name, steps)
        got = b.read(ids[k])
        s, f = inputs[k]
        sim = Simulation(s, f, dtype=torch.float32, nucleation=True, seed=1000 + k, instance=k,
                         storage_offset=specs[k]["x0"])
        sim.step(steps)
        want = sim.fields()
        sim.close()
        print(name, "job", k, "n", specs[k]["n"], "steps", steps, flush=True)
        for c in range(4):
            assert np.array_equal(got[c].astype(np.float64), want[c]), (name, c)
        # the fields moved: particles relax, the noise diffuses
        assert np.abs(want[0] - f[0])[1:-1, 1:-1].max() > 0.0
    b.status()
    b.close()


def test_batched_fp64_storage_against_oracle(gpu, oracle):
    jobs = [(1024, 1024, 0, 30), (300, 170, 1, 31), (129, 129, 0, 32)]
    pool = sum(job_bytes(PF_STORAGE_F64, Nx, Ny) for Nx, Ny, _, _ in jobs) + (16 << 20)
    b = Batch(PF_STORAGE_F64, pool_bytes=pool)
    b.set_tile(48, 3)
    b.set_launch_steps(3)
    cases = []
    for k, (Nx, Ny, bc, steps) in enumerate(jobs):
        s = M.make_setup(Nx, Ny, bc=bc)
        f = M.init_matrix_with_particles(Nx, Ny, s.dx, bc, seed=20 + k, n_del=2, n_lav=2,
                                         radius=min(5e-9, min(Nx, Ny) * s.dx / 12))
        jid = b.add(make_desc(s, steps, instance=k), f)
        cases.append((jid, s, f, steps, bc))
    b.run(3, 3)        # one captured graph of three launches
    b.run(20, 0)       # then eager launches past every target
    for jid, s, f, steps, bc in cases:
        assert b.poll(jid) == (steps, False)
        got = b.read(jid)
        P = oracle.make_params(s.dt, s.DCr, s.DNb, s.kappa, s.omega, s.Lmob, s.alpha, s.mask)
        want = oracle.run(f, P, bc, steps)
        for c in range(4):
            e = mixed_err(got[c], want[c])
            print("job", s.N, "bc", bc, "steps", steps, "field", c, "mixed err %.3e" % e,
                  flush=True)
            assert e <= TOL
        assert np.abs(want[2] - f[2]).max() > 1e-4   # the particles' interfaces moved
    b.close()


def test_serving_requests_at_full_size_alone_vs_mix(gpu):
    # (side, steps) from the serving ranges: sides 128..2048, 500..20000 steps
    shapes = [(2048, 520), (128, 500), (700, 611), (300, 777), (1500, 505), (181, 1001)]
    rng = np.random.default_rng(17)
    reqs = []
    for rid, (n, steps) in enumerate(shapes):
        x0 = (rng.uniform(*M.MATRIX_CR), rng.uniform(*M.MATRIX_NB)) if rid % 2 else \
            (rng.uniform(0.26, 0.29), rng.uniform(0.15, 0.18))
        reqs.append(Request(rid, 0.002 * rid, n, steps, x0, int(rng.integers(1 << 62))))
    pool = 6 << 30
    mixed, stats = serve(reqs, pool, launches_per_graph=5, rows=16, read_fields=True)
    assert stats["completed"] == len(reqs) and stats["rejected"] == 0
    by_id = {r.rid: r for r in mixed}
    # the large request shared the device with others
    big = by_id[0]
    assert any(r.admitted < big.done and r.done > big.admitted for r in mixed if r.rid != 0)
    for rid in (0, 1, 3):
        r0 = reqs[rid]
        alone, _ = serve([Request(r0.rid, 0.0, r0.n, r0.steps, r0.x0, r0.seed)], pool,
                         launches_per_graph=5, rows=16, read_fields=True)
        a, m = alone[0], by_id[rid]
        print("request", rid, "n", r0.n, "steps", r0.steps, flush=True)
        for c in range(4):
            assert np.array_equal(a.fields[c], m.fields[c]), (rid, c)
        assert a.summary == m.summary
