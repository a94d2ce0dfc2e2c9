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
