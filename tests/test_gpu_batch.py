"""Batched form: every job gives the same bits as the plain path run alone,
alone or in any mix, eager or replayed from a hipGraph."""
import numpy as np
import pytest
import torch

from pfpa import model as M
from pfpa.batch import PF_STORAGE_F32, PF_STORAGE_F64, Batch, make_desc
from pfpa.sim import Simulation

pytestmark = pytest.mark.gpu

NUC_DX = 4.83e-6   # nucleation-pass mesh spacing that puts P inside (0, 1) on 66 x 50 grids


def rand_fields(Nx, Ny, bc, seed, particles=True):
    rng = np.random.default_rng(seed)
    xCr = 0.3 + 0.01 * rng.standard_normal((Ny + 2, Nx + 2))
    xNb = 0.02 + 0.005 * rng.standard_normal((Ny + 2, Nx + 2))
    pd = np.zeros((Ny + 2, Nx + 2))
    pl = np.zeros((Ny + 2, Nx + 2))
    if particles:
        pd[2:Ny // 3, 2:Nx // 3] = 1.0
        pl[Ny // 2:Ny // 2 + 4, Nx // 2:Nx // 2 + 5] = 0.7
    return [M.fill_ghosts(a, bc) for a in (xCr, xNb, pd, pl)]


def nuc_fields(Nx, Ny, bc, seed):
    rng = np.random.default_rng(seed)
    xCr = rng.uniform(0.15, 0.3, (Ny + 2, Nx + 2))
    xNb = rng.uniform(0.19, 0.24, (Ny + 2, Nx + 2))
    z = np.zeros((Ny + 2, Nx + 2))
    return [M.fill_ghosts(a, bc) for a in (xCr, xNb, z, z.copy())]


def plain(setup, fields, steps, dtype, **kw):
    sim = Simulation(setup, fields, dtype=dtype, **kw)
    sim.step(steps)
    out = sim.fields()
    sim.close()
    return out


JOBS = [(64, 64, 1), (130, 71, 0), (37, 200, 1), (200, 33, 0), (1, 5, 1), (63, 62, 0)]


@pytest.mark.parametrize("launch_steps", [1, 2, 3])
@pytest.mark.parametrize("storage", [PF_STORAGE_F64, PF_STORAGE_F32])
def test_batch_equals_plain_alone(gpu, storage, launch_steps):
    """Each job (odd and even targets, tiny and ragged grids) against the
    plain path run alone, with one, two and three steps per launch."""
    dtype = torch.float64 if storage == PF_STORAGE_F64 else torch.float32
    b = Batch(storage, pool_bytes=256 << 20)
    b.set_tile(8, 3)
    b.set_launch_steps(launch_steps)
    ids = []
    setups = []
    for k, (Nx, Ny, bc) in enumerate(JOBS):
        s = M.make_setup(Nx, Ny, bc=bc)
        f = rand_fields(Nx, Ny, bc, 100 + k)
        ids.append(b.add(make_desc(s, steps=17 + k, instance=k), f))
        setups.append((s, f, 17 + k))
    b.run(40)          # more launches than any job needs: each stops at its own target
    for jid, (s, f, n) in zip(ids, setups):
        steps, active = b.poll(jid)
        assert steps == n and not active
        got = b.read(jid)
        for spl in (1, 2, 3):
            want = plain(s, f, n, dtype, steps_per_launch=spl)
            for k in range(4):
                assert np.array_equal(got[k].astype(np.float64), want[k]), (jid, spl, k)
    b.close()


@pytest.mark.parametrize("launch_steps", [1, 2, 3])
def test_alone_vs_mix_graph_and_eager(gpu, launch_steps):
    """FP32 job X with nucleation on its schedule: alone (eager) == in a mix
    (graph replays, admissions and a removal in between, other jobs at other
    points of their schedules) == the plain FP32 path."""
    s = M.make_setup(64, 48, bc=1)
    fx = nuc_fields(64, 48, 1, 7)
    steps = 29   # the last step (28) is a nucleation step: embedded cells hold exactly 1
    desc = dict(nucleation=True, seed=77, instance=5, nuc_every=4, nuc_dx=NUC_DX)
    alone = Batch(PF_STORAGE_F32, pool_bytes=64 << 20)
    alone.set_launch_steps(launch_steps)
    jx = alone.add(make_desc(s, steps, **desc), fx)
    alone.run(steps)
    a = alone.read(jx)
    alone.close()

    mix = Batch(PF_STORAGE_F32, pool_bytes=256 << 20)
    mix.set_tile(4, 6)
    mix.set_launch_steps(launch_steps)
    others = []
    for k, (Nx, Ny, bc) in enumerate(JOBS):
        so = M.make_setup(Nx, Ny, bc=bc)
        others.append(mix.add(make_desc(so, 1000, instance=50 + k), rand_fields(Nx, Ny, bc, k)))
    mix.run(7, 3)                          # others advance before X arrives
    jx2 = mix.add(make_desc(s, steps, **desc), fx)
    mix.run(5, 4)
    mix.remove(others[1])                  # retire one while X runs
    so = M.make_setup(150, 90, bc=0)
    mix.add(make_desc(so, 1000, instance=99), rand_fields(150, 90, 0, 3))
    mix.run(steps, 4)
    assert mix.poll(jx2) == (steps, False)
    m = mix.read(jx2)
    mix.close()
    want = plain(s, fx, steps, torch.float32, nucleation=True, seed=77, instance=5, nuc_every=4,
                 nuc_dx=NUC_DX)
    nucleated = (np.asarray(want[2]) == 1.0) | (np.asarray(want[3]) == 1.0)
    assert nucleated[1:-1, 1:-1].sum() > 5   # cells embedded by the last pass
    for k in range(4):
        assert np.array_equal(a[k], m[k]), k
        assert np.array_equal(a[k].astype(np.float64), want[k]), k


def test_graph_replays_match_eager(gpu):
    s = M.make_setup(120, 70, bc=0)
    f = rand_fields(120, 70, 0, 11)
    res = []
    for spg, ls in ((0, 1), (4, 1), (0, 2), (4, 2), (0, 3), (4, 3)):
        b = Batch(PF_STORAGE_F64, pool_bytes=64 << 20)
        b.set_launch_steps(ls)
        j = b.add(make_desc(s, 80), f)
        for r in range(5):                 # eager host work between replays
            b.run(16 // ls + 1, spg)
            b.poll(j)
        assert b.poll(j) == (80, False)
        res.append(b.read(j))
        b.close()
    for r in res[1:]:
        for k in range(4):
            assert np.array_equal(res[0][k], r[k])


@pytest.mark.parametrize("launch_steps", [1, 2, 3])
def test_batch_nucleation_matches_plain(gpu, launch_steps):
    """Nucleation schedule, random stream and embeds: identical to the plain
    fused path, with P inside (0, 1)."""
    for storage, dtype in ((PF_STORAGE_F64, torch.float64), (PF_STORAGE_F32, torch.float32)):
        s = M.make_setup(64, 48, bc=1)
        f = nuc_fields(64, 48, 1, 3)
        b = Batch(storage, pool_bytes=64 << 20)
        b.set_launch_steps(launch_steps)
        j = b.add(make_desc(s, 7, nucleation=True, seed=123, instance=4, nuc_every=3,
                            nuc_dx=NUC_DX), f)
        b.run(7)
        assert b.poll(j) == (7, False)
        got = b.read(j)
        assert b.status() == 0
        b.close()
        want = plain(s, f, 7, dtype, nucleation=True, seed=123, instance=4, nuc_every=3,
                     nuc_dx=NUC_DX)
        nucleated = (np.asarray(want[2]) == 1.0) | (np.asarray(want[3]) == 1.0)
        assert nucleated[1:-1, 1:-1].sum() > 50
        for k in range(4):
            assert np.array_equal(got[k].astype(np.float64), want[k]), (storage, k)


@pytest.mark.parametrize("launch_steps", [1, 3])
def test_batch_offset_matches_plain(gpu, launch_steps):
    """FP32 jobs storing x - x_offset (with the nucleation schedule and P in
    (0, 1)) hold the same bits as the plain FP32 path with the same storage
    offset; reads and summaries come back in absolute compositions."""
    off = (0.22, 0.21)
    s = M.make_setup(64, 48, bc=1)
    f = nuc_fields(64, 48, 1, 3)
    b = Batch(PF_STORAGE_F32, pool_bytes=64 << 20)
    b.set_launch_steps(launch_steps)
    kw = dict(nucleation=True, seed=123, instance=4, nuc_every=3, nuc_dx=NUC_DX)
    j = b.add(make_desc(s, 7, x_offset=off, **kw), f)
    b.run(7)
    assert b.poll(j) == (7, False)
    raw, got, summ = b.read(j, stored=True), b.read(j), b.summarize(j)
    b.close()
    sim = Simulation(s, f, dtype=torch.float32, storage_offset=off, **kw)
    sim.step(7)
    want_raw = [x.reshape(s.ny, s.nx).cpu().numpy() for x in sim.a]
    want, want_summ = sim.fields(), sim.summarize()
    sim.close()
    nucleated = (want[2] == 1.0) | (want[3] == 1.0)
    assert nucleated[1:-1, 1:-1].sum() > 50
    for k in range(4):
        assert np.array_equal(raw[k], want_raw[k]), k
        assert np.array_equal(np.asarray(got[k], dtype=np.float64), want[k]), k
    for k in range(5):
        assert summ[k] == pytest.approx(want_summ[k], rel=1e-12, abs=1e-15), k
    # FP64 storage takes no offset
    b = Batch(PF_STORAGE_F64, pool_bytes=64 << 20)
    with pytest.raises(Exception):
        b.add(make_desc(s, 7, x_offset=off), f)
    b.close()


def test_admission_by_pool_budget(gpu):
    from pfpa.batch import job_bytes
    per = job_bytes(PF_STORAGE_F32, 128, 128)
    b = Batch(PF_STORAGE_F32, pool_bytes=per * 3 + 1000)
    s = M.make_setup(128, 128, bc=1)
    f = rand_fields(128, 128, 1, 1, particles=False)
    ids = [b.add(make_desc(s, 5), f) for _ in range(3)]
    assert None not in ids
    assert b.add(make_desc(s, 5), f) is None      # no room
    b.remove(ids[1])
    j = b.add(make_desc(s, 5), f)                 # room again
    assert j is not None
    b.run(5)
    assert all(not b.poll(i)[1] for i in (ids[0], ids[2], j))
    b.close()


def test_batch_summarize(gpu, oracle):
    s = M.make_setup(200, 150, bc=0)
    f = M.init_matrix_with_particles(200, 150, s.dx, 0, seed=2, radius=2e-9)
    b = Batch(PF_STORAGE_F64, pool_bytes=64 << 20)
    j = b.add(make_desc(s, 3), f)
    b.run(3)
    got = b.summarize(j)
    want = oracle.summarize(b.read(j))
    for g, w in zip(got, want):
        assert g == pytest.approx(w, rel=1e-14, abs=1e-16)
    b.close()


def test_batch_errors(gpu):
    """Invalid arguments are rejected with an error and leave the batch usable."""
    from pfpa import PfError
    with pytest.raises(PfError):
        Batch(PF_STORAGE_F64, pool_bytes=1000)            # pool under 1 MiB
    with pytest.raises(PfError):
        Batch(7, pool_bytes=64 << 20)                     # unknown storage
    b = Batch(PF_STORAGE_F64, pool_bytes=64 << 20, max_jobs=2)
    for bad in ((0, 3), (8, 4), (1 << 17, 3)):
        with pytest.raises(PfError):
            b.set_tile(*bad)
    for bad in (0, 4):
        with pytest.raises(PfError):
            b.set_launch_steps(bad)
    s = M.make_setup(40, 30, bc=0)
    f = rand_fields(40, 30, 0, 1)
    d = make_desc(s, 5)
    d.bc = 7
    with pytest.raises(PfError):
        b.add(d, f)
    d = make_desc(s, 5)
    d.prm.mask[0] = 1.0                                   # a nine-point mask
    with pytest.raises(PfError):
        b.add(d, f)
    with pytest.raises(ValueError):
        b.add(make_desc(s, 5), rand_fields(41, 30, 0, 1))  # fields of another size
    j0 = b.add(make_desc(s, 5), f)
    j1 = b.add(make_desc(s, 7), f)
    with pytest.raises(PfError):
        b.add(make_desc(s, 5), f)                         # job table full
    for call in (lambda: b.remove(999), lambda: b.poll(999), lambda: b.summarize(999),
                 lambda: b.run(-1)):
        with pytest.raises(PfError):
            call()
    b.run(8)
    assert b.poll(j0) == (5, False) and b.poll(j1) == (7, False)
    want = plain(s, f, 7, torch.float64)
    got = b.read(j1)
    for k in range(4):
        assert np.array_equal(got[k], want[k]), k
    b.close()


def test_autotuner_leaves_a_consistent_batch(gpu):
    """autotune_batch over a small grid of candidates: the best entry is one of
    the measured ones, the batch is left on it, and the jobs' results after
    tuning and a further run still equal the plain path at their step counts."""
    from pfpa.workloads import autotune_batch
    b = Batch(PF_STORAGE_F32, pool_bytes=128 << 20)
    setups = []
    for k, (Nx, Ny, bc) in enumerate(JOBS[:4]):
        s = M.make_setup(Nx, Ny, bc=bc)
        f = rand_fields(Nx, Ny, bc, 300 + k)
        setups.append((b.add(make_desc(s, 400, instance=k), f), s, f))
    best, table = autotune_batch(b, launches=2, rows_set=(8, 32), depth_set=(3,),
                                 launch_steps_set=(1, 3), graph_set=(1, 4))
    assert best in table and best[0] == max(t[0] for t in table)
    assert len(table) == 2 * 2 + 1
    b.run(3, best[4] if best[4] > 1 else 0)
    for jid, s, f in setups:
        steps, active = b.poll(jid)
        assert active and 0 < steps < 400
        want = plain(s, f, steps, torch.float32)
        got = b.read(jid)
        for k in range(4):
            assert np.array_equal(got[k].astype(np.float64), want[k]), (jid, k)
    b.close()


def test_admission_while_launches_run(gpu):
    """pf_batch_add does not wait for the launches in flight: a job admitted
    while they run (into the slot and pool region a retired job freed) uploads
    its fields beside them and joins at the next commit.  Both the running job
    and the new one end with the same bits as each run alone."""
    sa = M.make_setup(120, 90, bc=1)
    sb = M.make_setup(140, 100, bc=0)
    sc = M.make_setup(110, 96, bc=1)
    fa, fb, fc = (rand_fields(120, 90, 1, 31), rand_fields(140, 100, 0, 32),
                  rand_fields(110, 96, 1, 33))

    def alone(s, f, steps, inst):
        b = Batch(PF_STORAGE_F32, pool_bytes=64 << 20)
        b.set_launch_steps(3)
        j = b.add(make_desc(s, steps, instance=inst), f)
        b.run(steps, 0)
        out = b.read(j)
        b.close()
        return out

    b = Batch(PF_STORAGE_F32, pool_bytes=64 << 20)
    b.set_launch_steps(3)
    ja = b.add(make_desc(sa, 9, instance=1), fa)
    jb = b.add(make_desc(sb, 300, instance=2), fb)
    b.run(5, 0)
    assert b.poll(ja) == (9, False)
    b.remove(ja)                     # frees slot and region
    b.run(40, 4)                     # asynchronous: B runs
    jc = b.add(make_desc(sc, 60, instance=3), fc)   # admitted while B's launches run
    assert jc == ja                  # the freed slot is reused
    b.run(200, 4)
    assert b.poll(jb) == (300, False) and b.poll(jc) == (60, False)
    got_b, got_c = b.read(jb), b.read(jc)
    b.close()
    for got, want in ((got_b, alone(sb, fb, 300, 2)), (got_c, alone(sc, fc, 60, 3))):
        for k in range(4):
            assert np.array_equal(got[k], want[k]), k


def test_task_claims_with_retired_jobs(gpu):
    """More strip tasks than the persistent grid has waves (each wave takes
    one task directly, then claims more from its XCD's counter), with half
    the jobs retiring part-way: the other jobs, and a retired one, are the
    same bits as the plain FP32 path run alone."""
    n, count = 254, 40
    tasks_per_job = ((n + 57) // 58) * ((n + 7) // 8)      # 58-column strips, 8-row chunks
    assert count * tasks_per_job > 256 * 4 * 4             # more tasks than any grid's waves
    b = Batch(PF_STORAGE_F32, pool_bytes=512 << 20)
    b.set_tile(8, 3)
    b.set_launch_steps(3)
    jobs = []
    for k in range(count):
        s = M.make_setup(n, n, bc=k % 2)
        f = rand_fields(n, n, k % 2, 500 + k)
        steps = 7 if k % 2 else 20                          # odd jobs retire after 3 launches
        jobs.append((b.add(make_desc(s, steps, instance=k), f), s, f, steps))
    b.run(8, 4)                                             # graph replays after the retirements
    for k in (0, 1, count - 2):
        jid, s, f, steps = jobs[k]
        assert b.poll(jid) == (steps, False)
        got = b.read(jid)
        want = plain(s, f, steps, torch.float32)
        for q in range(4):
            assert np.array_equal(got[q].astype(np.float64), want[q]), (k, q)
    b.close()
