"""FP32 storage with a composition offset (pf_ctx_set_storage_offset): the
arrays hold x - offset, every FP32 entry point works in those units, and the
step, nucleation pass, summaries and energy report what the same fields
stored without an offset give, to FP32 rounding.

Nucleation decisions: bit-exact against the oracle given the absolute
compositions the stored values stand for (stored + offset in FP64).
Multi-step and fused paths: the same bits as one launch per step and the
separate pass.  Accuracy: against the FP64 plain form, the offset keeps
the composition error in the matrix at least 4x below FP32 storage without
one."""
import ctypes

import numpy as np
import pytest
import torch

from pfpa import model as M
from pfpa._lib import PfNucParams, PfParams, check, lib
from pfpa.sim import Simulation

pytestmark = pytest.mark.gpu

OFF = (0.25, 0.25)


def _dev(x, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x)).to(device="cuda", dtype=dtype)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _nuc_struct(a, seed, call):
    N = PfNucParams()
    N.sigma_del, N.sigma_lav = a["sigma_del"], a["sigma_lav"]
    N.unit_a, N.ifce_width = a["unit_a"], 2.5e-9
    N.dx = N.dy = a["dx"]
    N.dz = a["dz"]
    N.dt = a["dt"]
    N.seed, N.call, N.instance = seed, call, 0
    return N


def _params(s):
    P = PfParams()
    P.dt = s.dt
    for k in range(2):
        P.DCr[k], P.DNb[k] = s.DCr[k], s.DNb[k]
        P.kappa[k], P.omega[k], P.Lmob[k] = s.kappa[k], s.omega[k], s.Lmob[k]
    P.alpha = s.alpha
    for k in range(9):
        P.mask[k] = s.mask[k]
    return P


def _stored_case(oracle, seed):
    """The synthetic nucleation case stored with offset 0.25, and the absolute
    compositions the stored values stand for (stored + 0.25 in FP64, as the
    kernels form them); those round back to the stored values."""
    from test_oracle import synthetic_nucleation_case
    fields, a = synthetic_nucleation_case(oracle, nx=70, ny=45, seed=seed)
    stored = [(fields[k] - OFF[k]).astype(np.float32) for k in range(2)]
    stored += [f.astype(np.float32) for f in fields[2:]]
    absolute = [stored[k].astype(np.float64) + OFF[k] for k in range(2)]
    absolute += [f.astype(np.float64) for f in stored[2:]]
    for k in range(2):
        assert np.array_equal((absolute[k] - OFF[k]).astype(np.float32), stored[k])
    return stored, absolute, a


def test_offset_nucleation_pass_bit_exact(gpu, oracle):
    """pf_nucleation_f32 on offset storage decides as the oracle does on the
    absolute compositions and embeds xe - offset."""
    stored, absolute, a = _stored_case(oracle, 13)
    ny, nx = stored[0].shape
    want, nev, _ = oracle.nucleation(absolute, a["sigma_del"], a["sigma_lav"], a["unit_a"],
                                     a["dx"] * nx, a["dx"] * ny, a["dz"], a["dt"], seed=5,
                                     call=2)
    assert nev > 50
    ctx = ctypes.c_void_p()
    check(lib().pf_ctx_create(ctypes.byref(ctx)), "ctx")
    check(lib().pf_ctx_set_storage_offset(ctx, *OFF), "offset")
    d = [_dev(x, torch.float32) for x in stored]
    N = _nuc_struct(a, 5, 2)
    check(lib().pf_nucleation_f32(ctx, nx, ny, *[_p(x) for x in d], ctypes.byref(N), None),
          "nucleation")
    torch.cuda.synchronize()
    for k in range(4):
        got = d[k].cpu().numpy()
        expect = (want[k] - (OFF[k] if k < 2 else 0.0)).astype(np.float32)
        assert np.array_equal(got, expect), k
    lib().pf_ctx_destroy(ctx)


def test_offset_fused_nucleation_matches_separate_pass(gpu, oracle):
    """pf_step_f32 with the nucleation pass fused == pf_step_f32 then
    pf_nucleation_f32, on offset storage, and both differ from the same run
    without an offset by FP32 rounding only."""
    stored, absolute, a = _stored_case(oracle, 17)
    ny, nx = stored[0].shape
    s = M.make_setup(nx - 2, ny - 2, bc=1, dt_rule=1e-7)
    stored = [M.fill_ghosts(f.copy(), 1) for f in stored]
    P = _params(s)
    N = _nuc_struct(a, 99, 4)
    ctx = ctypes.c_void_p()
    check(lib().pf_ctx_create(ctypes.byref(ctx)), "ctx")
    check(lib().pf_ctx_set_storage_offset(ctx, *OFF), "offset")
    src = [_dev(x, torch.float32) for x in stored]
    fused = [torch.empty_like(x) for x in src]
    sep = [torch.empty_like(x) for x in src]
    check(lib().pf_step_f32(ctx, nx, ny, 1, ctypes.byref(P), *[_p(x) for x in src],
                            *[_p(x) for x in fused], ctypes.byref(N), None), "fused")
    check(lib().pf_step_f32(ctx, nx, ny, 1, ctypes.byref(P), *[_p(x) for x in src],
                            *[_p(x) for x in sep], None, None), "step")
    check(lib().pf_nucleation_f32(ctx, nx, ny, *[_p(x) for x in sep], ctypes.byref(N), None),
          "nucleation")
    # the same on absolute FP32 storage (offset 0)
    check(lib().pf_ctx_set_storage_offset(ctx, 0.0, 0.0), "offset 0")
    src0 = [_dev(M.fill_ghosts(x.copy(), 1), torch.float32) for x in absolute]
    plain = [torch.empty_like(x) for x in src0]
    check(lib().pf_step_f32(ctx, nx, ny, 1, ctypes.byref(P), *[_p(x) for x in src0],
                            *[_p(x) for x in plain], ctypes.byref(N), None), "plain")
    torch.cuda.synchronize()
    embedded = 0
    for k in range(4):
        f = fused[k].cpu().numpy()[1:-1, 1:-1]
        g = sep[k].cpu().numpy()[1:-1, 1:-1]
        assert np.array_equal(f, g), k
        h = plain[k].cpu().numpy()[1:-1, 1:-1].astype(np.float64)
        fa = f.astype(np.float64) + (OFF[k] if k < 2 else 0.0)
        assert np.abs(fa - h).max() <= 6e-8, k   # two FP32 roundings near 0.3
        if k >= 2:
            embedded += int((f == 1.0).sum())
    assert embedded > 50
    lib().pf_ctx_destroy(ctx)


@pytest.mark.parametrize("spl", [3, 4])
@pytest.mark.parametrize("bc", [0, 1])
def test_offset_multistep_equals_one_step(gpu, bc, spl):
    """Three and four steps per launch on offset storage == one launch per
    step, with the nucleation schedule on (fused nucleation launches in
    between)."""
    Nx, Ny = 150, 97
    s = M.make_setup(Nx, Ny, bc=bc)
    f0 = M.init_matrix_with_particles(Nx, Ny, s.dx, bc, seed=3, radius=2.5e-9)
    a = Simulation(s, f0, dtype=torch.float32, nucleation=True, nuc_every=7,
                   steps_per_launch=spl, storage_offset="mean")
    assert a.steps_per_launch == spl
    b = Simulation(s, f0, dtype=torch.float32, nucleation=True, nuc_every=7,
                   steps_per_launch=1, storage_offset="mean")
    assert a.xoff == b.xoff and 0.25 < a.xoff[0] < 0.35 and 0.02 < a.xoff[1] < 0.1
    for n in (1, 5, 40, 60):
        a.step(n)
        b.step(n)
    for x, y in zip(a.a, b.a):
        assert torch.equal(x, y)
    a.close()
    b.close()


def test_offset_accuracy_and_reports(gpu):
    """Against the FP64 plain form after 1000 steps: the RMS composition error
    in the matrix (phi < 1e-6) is at least 4x smaller with the median offset;
    summaries and energy agree with the FP64 run to FP32 accuracy."""
    Nx = Ny = 256
    s = M.make_setup(Nx, Ny, bc=1)
    f0 = M.init_matrix_with_particles(Nx, Ny, s.dx, 1, seed=1)
    ref = Simulation(s, f0, dtype=torch.float64)
    plain = Simulation(s, f0, dtype=torch.float32)
    off = Simulation(s, f0, dtype=torch.float32, storage_offset="median")
    for sim in (ref, plain, off):
        sim.step(1000)
    fr, fp, fo = ref.fields(), plain.fields(), off.fields()
    matrix = (fr[2] < 1e-6) & (fr[3] < 1e-6)
    matrix[0, :] = matrix[-1, :] = matrix[:, 0] = matrix[:, -1] = False
    assert matrix.sum() > Nx * Ny // 2
    for k in range(2):
        ep = float(np.sqrt(np.mean((fp[k] - fr[k])[matrix] ** 2)))
        eo = float(np.sqrt(np.mean((fo[k] - fr[k])[matrix] ** 2)))
        print("RMS x error in the matrix, component %d: no offset %.3e, offset %.3e"
              % (k, ep, eo))
        assert eo * 4 <= ep, (k, ep, eo)
    sr, so = ref.summarize(), off.summarize()
    # mean x_Cr, x_Nb (conserved up to the FP32 rounding of the increments in the
    # precipitates, whose x - offset is not small); gamma, delta, Laves fractions
    for k in range(5):
        assert abs(so[k] - sr[k]) <= (1e-7 if k < 2 else 1e-6), (k, so[k], sr[k])
    er, eo_ = ref.energy(), off.energy()
    assert abs(eo_[0] - er[0]) <= 1e-5 * abs(er[0])
    for sim in (ref, plain, off):
        sim.close()


def test_offset_wide_nucleation_events(gpu, oracle):
    """The deferred-event path (scan over the whole grid, overlapping embeds of
    radius ~1.7 cells, gamma rings around them) on offset storage: the oracle's
    result on the absolute compositions, less the offset."""
    Nx, Ny = 38, 30
    nx, ny = Nx + 2, Ny + 2
    rng = np.random.default_rng(21)
    absolute = [M.fill_ghosts(rng.uniform(0.15, 0.3, (ny, nx)), 0),
                M.fill_ghosts(rng.uniform(0.19, 0.24, (ny, nx)), 0),
                np.zeros((ny, nx)), np.zeros((ny, nx))]
    stored = [(absolute[k] - OFF[k]).astype(np.float32) for k in range(2)]
    stored += [f.astype(np.float32) for f in absolute[2:]]
    absolute = [stored[k].astype(np.float64) + OFF[k] for k in range(2)] + absolute[2:]
    a = dict(sigma_del=0.0105, sigma_lav=0.0108, unit_a=0.352e-9, dx=2e-13, dz=2e-13, dt=3e38)
    want, nev, _ = oracle.nucleation(absolute, a["sigma_del"], a["sigma_lav"], a["unit_a"],
                                     a["dx"] * nx, a["dx"] * ny, a["dz"], a["dt"],
                                     seed=77, call=2)
    assert nev > 20
    ctx = ctypes.c_void_p()
    check(lib().pf_ctx_create(ctypes.byref(ctx)), "ctx")
    check(lib().pf_ctx_set_storage_offset(ctx, *OFF), "offset")
    d = [_dev(x, torch.float32) for x in stored]
    N = _nuc_struct(a, 77, 2)
    check(lib().pf_nucleation_f32(ctx, nx, ny, *[_p(x) for x in d], ctypes.byref(N), None),
          "nucleation")
    torch.cuda.synchronize()
    for k in range(4):
        expect = (want[k] - (OFF[k] if k < 2 else 0.0)).astype(np.float32)
        assert np.array_equal(d[k].cpu().numpy(), expect), k
    wide = ctypes.c_uint64()
    over = ctypes.c_int()
    check(lib().pf_nucleation_status(ctx, ctypes.byref(wide), ctypes.byref(over)), "status")
    assert wide.value == nev and over.value == 0
    lib().pf_ctx_destroy(ctx)


def test_offset_nine_point_mask(gpu):
    """The general nine-term Laplacian (one step per launch) on offset
    storage: within FP32 rounding of the same run without an offset."""
    Nx, Ny = 96, 70
    s = M.make_setup(Nx, Ny, bc=0, dt_rule=5e-8)
    s.mask = M.nine_point_mask(s.dx, 1.3 * s.dx)
    f0 = M.init_matrix_with_particles(Nx, Ny, s.dx, 0, seed=8, radius=2e-9)
    a = Simulation(s, f0, dtype=torch.float32, storage_offset="median")
    b = Simulation(s, f0, dtype=torch.float32)
    c = Simulation(s, f0, dtype=torch.float64)
    assert a.steps_per_launch == 1
    for sim in (a, b, c):
        sim.step(25)
    fa, fb, fc = a.fields(), b.fields(), c.fields()
    for k in range(4):
        ea = np.abs(fa[k] - fc[k])[1:-1, 1:-1].max()
        eb = np.abs(fb[k] - fc[k])[1:-1, 1:-1].max()
        assert ea <= max(eb, 1e-7), (k, ea, eb)
    for sim in (a, b, c):
        sim.close()


def test_offset_errors(gpu):
    ctx = ctypes.c_void_p()
    check(lib().pf_ctx_create(ctypes.byref(ctx)), "ctx")
    assert lib().pf_ctx_set_storage_offset(ctx, float("nan"), 0.0) != 0
    assert lib().pf_ctx_set_storage_offset(ctx, 0.0, float("inf")) != 0
    assert lib().pf_ctx_set_storage_offset(ctx, 0.3, 0.02) == 0
    lib().pf_ctx_destroy(ctx)
    s = M.make_setup(16, 16, bc=1)
    f0 = M.init_matrix_with_particles(16, 16, s.dx, 1, seed=1, n_del=0, n_lav=0)
    with pytest.raises(ValueError):
        Simulation(s, f0, dtype=torch.float64, storage_offset="mean")
    with pytest.raises(ValueError):
        Simulation(s, f0, dtype=torch.float32, storage_offset="mode")
