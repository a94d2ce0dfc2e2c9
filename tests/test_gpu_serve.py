"""The serving loop (pfpa.serve): requests admitted by HBM budget with
continuous batching give the same bits as each request served alone and as
the plain FP32-storage path run alone."""
import numpy as np
import pytest
import torch

from pfpa.batch import PF_STORAGE_F32, job_bytes
from pfpa.serve import make_requests, request_desc, request_fields, request_offset, serve
from pfpa.sim import Simulation

pytestmark = pytest.mark.gpu


def test_serving_alone_vs_mix():
    reqs = make_requests(rate=400.0, duration=0.03, seed=5, n_lo=64, n_hi=150, steps_lo=20,
                         steps_hi=70)
    assert len(reqs) >= 6
    # room for about three of the largest at once: later arrivals wait and
    # are admitted as earlier ones retire
    pool = 3 * job_bytes(PF_STORAGE_F32, 150, 150) + (1 << 20)
    mixed, stats = serve(reqs, pool, launches_per_graph=5, rows=16, read_fields=True)
    assert stats["completed"] == len(reqs)
    assert max(r.admitted for r in mixed) > min(r.done for r in mixed)   # admissions overlap
    by_id = {r.rid: r for r in mixed}
    for rid in (0, len(reqs) // 2, len(reqs) - 1):
        r0 = reqs[rid]
        alone, _ = serve([type(r0)(r0.rid, 0.0, r0.n, r0.steps, r0.x0, r0.seed)], pool,
                         launches_per_graph=5, rows=16, read_fields=True)
        a, m = alone[0], by_id[rid]
        for k in range(4):
            assert np.array_equal(a.fields[k], m.fields[k]), (rid, k)
        assert a.summary == m.summary
        # and the plain FP32-storage path alone, with the same composition offset
        # (the serving loop stores each request relative to its matrix composition)
        s, _ = request_desc(r0)
        assert request_offset(r0) == (r0.x0[0], r0.x0[1]) and r0.x0[0] > 0.2
        sim = Simulation(s, request_fields(r0), dtype=torch.float32, nucleation=True,
                         seed=r0.seed, instance=r0.rid, storage_offset=request_offset(r0))
        sim.step(r0.steps)
        want = sim.fields()
        sim.close()
        for k in range(4):
            assert np.array_equal(m.fields[k].astype(np.float64), want[k]), (rid, k)


def test_request_larger_than_the_pool_is_rejected():
    """A request that cannot fit even in the empty pool is rejected (the loop
    ends) while the others are served."""
    reqs = make_requests(rate=400.0, duration=0.02, seed=7, n_lo=64, n_hi=96, steps_lo=10,
                         steps_hi=20)
    assert len(reqs) >= 2
    big = type(reqs[0])(len(reqs), reqs[-1].arrival, 600, 10, reqs[0].x0, 3)
    pool = 2 * job_bytes(PF_STORAGE_F32, 96, 96) + (1 << 20)
    assert job_bytes(PF_STORAGE_F32, 600, 600) > pool
    done, stats = serve(reqs + [big], pool, launches_per_graph=3, rows=16)
    assert stats["rejected"] == 1 and stats["completed"] == len(reqs)
    assert big not in done


def test_hold_keeps_state_and_resumes_same_bits():
    """A held job takes no part in the launches (its step count stays) and,
    released, ends with the same bits as a job never held."""
    from pfpa import model as M
    from pfpa.batch import Batch, make_desc
    s = M.make_setup(90, 70, bc=1)
    f = M.init_matrix_with_particles(90, 70, s.dx, 1, seed=4, radius=2e-9)
    so = M.make_setup(64, 64, bc=0)
    fo = M.init_matrix_with_particles(64, 64, so.dx, 0, seed=5, radius=2e-9)

    def run(hold):
        b = Batch(PF_STORAGE_F32, pool_bytes=64 << 20)
        b.set_launch_steps(3)
        ja = b.add(make_desc(s, 40, nucleation=True, seed=9, instance=1), f)
        jo = b.add(make_desc(so, 200, instance=2), fo)
        b.run(2, 0)
        if hold:
            at = b.poll(ja)[0]
            b.hold(ja, True)
            b.run(5, 5)                      # graph replays while held
            assert b.poll(ja)[0] == at       # no progress while held
            assert b.poll(jo)[0] > at        # the other job ran
            b.hold(ja, False)
        b.run(30, 3)
        assert b.poll(ja) == (40, False)
        out = b.read(ja)
        b.close()
        return out

    a, h = run(False), run(True)
    for k in range(4):
        assert np.array_equal(a[k], h[k]), k


def test_serving_shortest_remaining_work_same_results():
    """Shortest-remaining-work order holds the longer jobs (some rounds run
    only the shorter ones) and every result is the same bits as in FIFO."""
    reqs = make_requests(rate=300.0, duration=1.0, seed=9, n_lo=64, n_hi=200, steps_lo=20,
                         steps_hi=120, max_requests=10)
    assert len(reqs) == 10
    for r in reqs:          # all waiting at once, so several run together
        r.arrival = 0.0
    pool = 1 << 30
    fifo, st_f = serve(reqs, pool, launches_per_graph=3, rows=16, read_fields=True)
    srw, st_s = serve(reqs, pool, launches_per_graph=3, rows=16, read_fields=True,
                      order="srw", fill_cells=20000, max_hold_s=0.05)
    assert st_f["completed"] == st_s["completed"] == len(reqs)
    assert st_s["held_job_rounds"] > 0 and st_f["held_job_rounds"] == 0
    assert 0.0 < st_s["device_busy_share"] <= 1.0
    by_f = {r.rid: r for r in fifo}
    for r in srw:
        for k in range(4):
            assert np.array_equal(r.fields[k], by_f[r.rid].fields[k]), (r.rid, k)
        assert r.summary == by_f[r.rid].summary


def test_serving_auto_tile_same_results():
    """The adaptive strip height (short tasks while few cells run, tall ones
    under load) switches the batch's tile between rounds and changes no bits."""
    reqs = make_requests(rate=300.0, duration=1.0, seed=13, n_lo=64, n_hi=200, steps_lo=20,
                         steps_hi=120, max_requests=10)
    for r in reqs:
        r.arrival = 0.0
    pool = 1 << 30
    fixed, _ = serve(reqs, pool, launches_per_graph=3, rows=16, read_fields=True)
    auto, st = serve(reqs, pool, launches_per_graph=3, rows="auto", read_fields=True,
                     auto_rows=(8, 48), auto_cells=(20000, 60000))
    assert st["tile_switches"] >= 1 and st["completed"] == len(reqs)
    by_f = {r.rid: r for r in fixed}
    for r in auto:
        for k in range(4):
            assert np.array_equal(r.fields[k], by_f[r.rid].fields[k]), (r.rid, k)
        assert r.summary == by_f[r.rid].summary
