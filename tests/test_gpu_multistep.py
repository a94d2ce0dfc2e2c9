"""The temporally blocked kernels, pf_step2_*, pf_step3_* and pf_step4_*,
through the C-ABI.

Each must give what two (three, four) one-step launches give: the same interior
values and the same ghost-ring images, for both boundary modes, both storage
types, ragged grids (strips and chunks that do not divide the grid, chunks of
one row, tiny grids where the images sit inside the first chunk) and every
tile shape.  Their parity with the oracle over long runs is covered by the
Simulation-driven tests, which group steps through them.
"""
import ctypes

import numpy as np
import pytest
import torch

from pfpa import model as M
from pfpa._lib import PfParams, check, lib
from pfpa.sim import Simulation

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _params(s, mask=None):
    P = PfParams()
    P.dt = s.dt
    for k in range(2):
        P.DCr[k], P.DNb[k] = s.DCr[k], s.DNb[k]
        P.kappa[k], P.omega[k], P.Lmob[k] = s.kappa[k], s.omega[k], s.Lmob[k]
    P.alpha = s.alpha
    m = s.mask if mask is None else mask
    for k in range(9):
        P.mask[k] = m[k]
    return P


def _fields(Nx, Ny, bc, seed):
    rng = np.random.default_rng(seed)
    xCr = 0.3 + 0.01 * rng.standard_normal((Ny + 2, Nx + 2))
    xNb = 0.02 + 0.005 * rng.standard_normal((Ny + 2, Nx + 2))
    pd = np.clip(rng.uniform(-0.3, 1.3, (Ny + 2, Nx + 2)), 0, 1)
    pl = np.clip(rng.uniform(-0.5, 1.0, (Ny + 2, Nx + 2)), 0, 1) * (pd < 0.3)
    return [M.fill_ghosts(a, bc) for a in (xCr, xNb, pd, pl)]


def _mixed_err(a, b):
    sf = max(np.abs(b).max(), 1e-300)
    return float((np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), sf)).max())


@pytest.fixture(scope="module")
def ctx(gpu):
    h = ctypes.c_void_p()
    check(lib().pf_ctx_create(ctypes.byref(h)), "pf_ctx_create")
    yield h
    torch.cuda.synchronize()
    lib().pf_ctx_destroy(h)


# strip widths 64 - 2k: 60, 58 and 56 output columns
SHAPES = [(3, 3), (4, 4), (4, 5), (5, 4), (56, 9), (57, 58), (59, 7), (60, 61), (61, 62),
          (62, 3), (120, 9), (121, 33), (200, 131), (37, 300)]

MULTI = {2: ("pf_step2_f64", "pf_step2_f32", "pf_ctx_set_tile2"),
         3: ("pf_step3_f64", "pf_step3_f32", "pf_ctx_set_tile3"),
         4: ("pf_step4_f64", "pf_step4_f32", "pf_ctx_set_tile4")}


@pytest.mark.parametrize("k", [2, 3, 4])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("shape", SHAPES)
def test_k_steps_in_one_launch(ctx, shape, bc, dtype, k):
    Nx, Ny = shape
    if k == 4 and min(Nx, Ny) < 4:
        pytest.skip("the four-step kernel needs a 4x4 interior (rejected, see the error test)")
    nx, ny = Nx + 2, Ny + 2
    s = M.make_setup(Nx, Ny, bc=bc, dt_rule=4e-7)
    P = _params(s)
    f0 = [torch.as_tensor(f).to(device="cuda", dtype=dtype) for f in
          _fields(Nx, Ny, bc, 1000 * Nx + Ny)]
    one = lib().pf_step_f64 if dtype == torch.float64 else lib().pf_step_f32
    f64 = dtype == torch.float64
    multi = getattr(lib(), MULTI[k][0 if f64 else 1])
    set_tile = getattr(lib(), MULTI[k][2])
    cur = f0
    for _ in range(k):
        nxt = [torch.full_like(x, np.nan) for x in f0]
        check(one(ctx, nx, ny, bc, ctypes.byref(P), *map(_p, cur), *map(_p, nxt), None, None),
              "step")
        cur = nxt
    want = [x.double().cpu().numpy() for x in cur]
    outs = []
    for rows, depth in ((0, 0), (1, 2), (2, 3), (5, 6), (6, 3), (7, 2), (64, 3)):
        check(set_tile(ctx, rows, depth), "tile")
        out = [torch.full_like(x, np.nan) for x in f0]
        check(multi(ctx, nx, ny, bc, ctypes.byref(P), *map(_p, f0), *map(_p, out), None),
              "step%d" % k)
        got = [x.double().cpu().numpy() for x in out]
        for f in range(4):
            # interior and ghost ring alike
            assert np.all(np.isfinite(got[f])), (rows, depth, f)
            assert _mixed_err(got[f], want[f]) <= (TOL if f64 else 0.0), \
                (rows, depth, f, _mixed_err(got[f], want[f]))
        outs.append(got)
    check(set_tile(ctx, 0, 0), "tile")
    for o in outs[1:]:
        for f in range(4):
            assert np.array_equal(o[f], outs[0][f])


@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("spl", [2, 3, 4])
def test_fused_run_matches_one_step_run(gpu, oracle, bc, spl):
    """A Simulation grouping steps through the multi-step kernels, with the
    reference's nucleation schedule (odd interval, so groups straddle the
    nucleation steps), against one launch per step and against the oracle."""
    Nx, Ny = 150, 97
    s = M.make_setup(Nx, Ny, bc=bc)
    f0 = M.init_matrix_with_particles(Nx, Ny, s.dx, bc, seed=3, radius=2.5e-9)
    a = Simulation(s, f0, nucleation=True, nuc_every=7, steps_per_launch=spl)
    b = Simulation(s, f0, nucleation=True, nuc_every=7, steps_per_launch=1)
    assert a.steps_per_launch == spl and b.steps_per_launch == 1
    for n in (1, 2, 5, 40, 153):
        a.step(n)
        b.step(n)
    ga, gb = a.fields(), b.fields()
    want = oracle.run(f0, oracle.make_params(s.dt, s.DCr, s.DNb, s.kappa, s.omega, s.Lmob,
                                             s.alpha, s.mask), bc, a.steps_done)
    for k in range(4):
        assert _mixed_err(ga[k], gb[k]) <= TOL, k
        assert _mixed_err(ga[k][1:-1, 1:-1], want[k][1:-1, 1:-1]) <= TOL, k
    a.close()
    b.close()


@pytest.mark.parametrize("k", [2, 3, 4])
def test_multistep_errors(ctx, k):
    s = M.make_setup(8, 8, bc=0)
    P = _params(s)
    f = [torch.zeros(10 * 10, dtype=torch.float64, device="cuda") for _ in range(8)]
    two = getattr(lib(), MULTI[k][0])
    set_tile = getattr(lib(), MULTI[k][2])
    # in place
    assert two(ctx, 10, 10, 0, ctypes.byref(P), *map(_p, f[:4]), *map(_p, f[:4]), None) != 0
    # nine-point mask
    m9 = list(s.mask)
    m9[0] = m9[2] = m9[6] = m9[8] = 1.0
    P9 = _params(s, m9)
    assert two(ctx, 10, 10, 0, ctypes.byref(P9), *map(_p, f[:4]), *map(_p, f[4:]), None) != 0
    assert b"five-point" in lib().pf_last_error()
    # five-point but not symmetric (west and east weights differ)
    m5 = list(s.mask)
    m5[3] *= 1.25
    m5[4] -= 0.25 * s.mask[3]
    P5 = _params(s, m5)
    assert two(ctx, 10, 10, 0, ctypes.byref(P5), *map(_p, f[:4]), *map(_p, f[4:]), None) != 0
    assert b"symmetric five-point" in lib().pf_last_error()
    # interior smaller than 3x3 (four steps: 4x4)
    assert two(ctx, 4, 10, 0, ctypes.byref(P), *map(_p, f[:4]), *map(_p, f[4:]), None) != 0
    small = two(ctx, 5, 10, 0, ctypes.byref(P), *map(_p, f[:4]), *map(_p, f[4:]), None)
    assert (small != 0) == (k == 4)
    if k == 4:
        assert b"4x4" in lib().pf_last_error()
    # bad boundary mode, bad tile
    assert two(ctx, 10, 10, 7, ctypes.byref(P), *map(_p, f[:4]), *map(_p, f[4:]), None) != 0
    assert set_tile(ctx, 8, 4) != 0
    assert set_tile(ctx, -1, 3) != 0
    assert two(ctx, 10, 10, 0, ctypes.byref(P), *map(_p, f[:4]), *map(_p, f[4:]), None) == 0
    torch.cuda.synchronize()


def test_step_loops_equal_single_calls(gpu):
    """pf_step_loop_f64 / pf_step2_loop_f64 / pf_step3_loop_f64 /
    pf_step4_loop_f64 (one launch per grid of a list, the batched form's
    comparison baseline) give the same bits
    as the single-grid entry points on each grid; a bad list is rejected."""
    from pfpa._lib import PfParams, PfStepItem, check, lib
    grids = [(40, 30, 0), (129, 17, 1), (7, 64, 1)]
    items = (PfStepItem * len(grids))()
    keep = []
    singles = []
    for i, (Nx, Ny, bc) in enumerate(grids):
        s = M.make_setup(Nx, Ny, bc=bc, dt_rule=4e-7)
        f = _fields(Nx, Ny, bc, 70 + i)
        P = PfParams()
        P.dt = s.dt
        for k in range(2):
            P.DCr[k], P.DNb[k] = s.DCr[k], s.DNb[k]
            P.kappa[k], P.omega[k], P.Lmob[k] = s.kappa[k], s.omega[k], s.Lmob[k]
        P.alpha = s.alpha
        for k in range(9):
            P.mask[k] = s.mask[k]
        A = [torch.as_tensor(x).cuda() for x in f]
        B = [torch.empty_like(x) for x in A]
        it = items[i]
        it.nx, it.ny, it.bc = s.nx, s.ny, s.bc
        it.prm = ctypes.pointer(P)
        for k in range(4):
            it.old_fields[k] = A[k].data_ptr()
            it.new_fields[k] = B[k].data_ptr()
        keep.append((P, A, B))
        singles.append((s, P, A))
    ctx = ctypes.c_void_p()
    check(lib().pf_ctx_create(ctypes.byref(ctx)), "ctx")
    for name, single in (("pf_step_loop_f64", "pf_step_f64"), ("pf_step2_loop_f64", "pf_step2_f64"),
                         ("pf_step3_loop_f64", "pf_step3_f64"), ("pf_step4_loop_f64", "pf_step4_f64")):
        check(getattr(lib(), name)(ctx, len(grids), items, None), name)
        torch.cuda.synchronize()
        for (s, P, A), (_, _, B) in zip(singles, keep):
            C = [torch.empty_like(x) for x in A]
            args = [ctx, s.nx, s.ny, s.bc, ctypes.byref(P), *[_p(x) for x in A], *[_p(x) for x in C]]
            if single == "pf_step_f64":
                args.append(None)
            check(getattr(lib(), single)(*args, None), single)
            torch.cuda.synchronize()
            for k in range(4):
                assert torch.equal(B[k], C[k]), (name, s.N, k)
        assert getattr(lib(), name)(ctx, -1, items, None) != 0
        assert getattr(lib(), name)(ctx, 2, None, None) != 0
    lib().pf_ctx_destroy(ctx)


@pytest.mark.parametrize("spl", [3, 4])
def test_tune_tile_leaves_fields_and_results(gpu, spl):
    """Simulation.tune_tile (tune_tile3 at three steps per launch) measures on
    scratch copies: the fields do not move, the chosen height is one of the
    candidates, and stepping afterwards gives the same bits as an untuned
    simulation."""
    s = M.make_setup(300, 200, bc=1)
    f0 = M.init_matrix_with_particles(300, 200, s.dx, 1, seed=4, radius=2e-9)
    tuned = Simulation(s, f0, steps_per_launch=spl)
    before = tuned.fields()
    tune = tuned.tune_tile3 if spl == 3 else tuned.tune_tile
    rows, times = tune(rows_set=(6, 16, 48), launches=4)
    assert rows in (6, 16, 48) and set(times) == {6, 16, 48}
    after = tuned.fields()
    for k in range(4):
        assert np.array_equal(before[k], after[k])
    tuned.step(31)
    plain = Simulation(s, f0, steps_per_launch=spl)
    plain.step(31)
    for x, y in zip(tuned.fields(), plain.fields()):
        assert np.array_equal(x, y)
    if spl == 4:
        with pytest.raises(ValueError):
            tuned.tune_tile3()
        # the default candidates include the kernel's own default height
        rows, times = tuned.tune_tile(launches=2)
        default = lib().pf_default_rows4(s.nx, s.ny, 0)
        assert default in times and rows in times and max(times) <= 200
    tuned.close()
    plain.close()


def test_concurrent_simulations_on_two_threads(gpu):
    """Two simulations stepped at the same time from two host threads, each
    with its own context and stream, give the same bits as each run alone
    (the library keeps no shared mutable state outside a context)."""
    import threading
    cases = [(M.make_setup(150, 90, bc=0), 21), (M.make_setup(97, 130, bc=1), 22)]
    inputs = [M.init_matrix_with_particles(s.N[0], s.N[1], s.dx, s.bc, seed=seed,
                                           radius=2e-9) for s, seed in cases]
    alone = []
    alone_summ = []
    for (s, _), f in zip(cases, inputs):
        sim = Simulation(s, f, nucleation=True, seed=5)
        sim.step(40)
        alone.append(sim.fields())
        alone_summ.append(sim.summarize())
        sim.close()
    out = [None, None]
    summ = [None, None]
    errors = []

    def worker(i):
        try:
            s, _ = cases[i]
            stream = torch.cuda.Stream()
            sim = Simulation(s, inputs[i], nucleation=True, seed=5, stream=stream)
            for _ in range(10):
                sim.step(4)
            # queued behind the steps on the simulation's own stream: the
            # result buffer is made there too, not on the current stream
            summ[i] = sim.summarize()
            out[i] = sim.fields()
            sim.close()
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for i in range(2):
        for k in range(4):
            assert np.array_equal(out[i][k], alone[i][k]), (i, k)
        assert summ[i] == alone_summ[i], i
