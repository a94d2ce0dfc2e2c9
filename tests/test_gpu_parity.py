"""GPU parity: libpfpa.so (HIP, gfx950) against the CPU oracle and the
golden vectors, through the C-ABI.

Field tolerance (FP64): |a - b| <= 1e-12 * max(|a|, |b|, s_f) per cell, with
s_f = max |field| over the grid (phi is ~0 away from particles, so a pure
relative error is meaningless there).  Nucleation decisions: bit-exact.
"""
import ctypes

import numpy as np
import pytest
import torch

import pfpa
from pfpa import model as M
from pfpa._lib import PfNucParams, PfParams, check, lib
from pfpa.sim import Simulation, iter_launches

pytestmark = pytest.mark.gpu

TOL = 1e-12


def mixed_err(a, b):
    a = a[1:-1, 1:-1]
    b = b[1:-1, 1:-1]
    sf = max(np.abs(b).max(), 1e-300)
    return float((np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), sf)).max())


def oracle_params(O, s):
    return O.make_params(s.dt, s.DCr, s.DNb, s.kappa, s.omega, s.Lmob, s.alpha, s.mask)


def _dev(x, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x)).to(device="cuda", dtype=dtype)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


# ------------------------------------------------------------------ pointwise


def test_device_thermo_golden(gpu, thermo_golden):
    G = thermo_golden
    xs = _dev(G["p_in"])
    out = torch.zeros(2 * len(G["p_in"]), dtype=torch.float64, device="cuda")
    check(lib().pf_thermo_eval(0, len(G["p_in"]), _p(xs), _p(out), None), "eval")
    o = out.cpu().numpy().reshape(-1, 2)
    assert np.allclose(o[:, 0], G["p_out"], rtol=1e-14, atol=1e-17)
    assert np.allclose(o[:, 1], G["pPrime_out"], rtol=1e-14, atol=1e-17)

    fin = G["fict_in"]
    out = torch.zeros(6 * len(fin), dtype=torch.float64, device="cuda")
    check(lib().pf_thermo_eval(1, len(fin), _p(_dev(fin)), _p(out), None), "eval")
    o = out.cpu().numpy().reshape(-1, 6)
    phys = np.abs(fin[:, 2] + fin[:, 3] + fin[:, 4] - 1.0) < 1e-12
    names = ["fict_%s_%s" % (ph, c) for ph in ("gam", "del", "lav") for c in ("Cr", "Nb")]
    for k, n in enumerate(names):
        err = np.abs(o[:, k] - G[n])
        assert err[phys].max() < 1e-12, n
        assert np.all(err[~phys] <= 1e-11 * np.maximum(1.0, np.abs(G[n][~phys]))), n

    comp = G["comp_in"]
    out = torch.zeros(13 * len(comp), dtype=torch.float64, device="cuda")
    check(lib().pf_thermo_eval(2, len(comp), _p(_dev(comp)), _p(out), None), "eval")
    o = out.cpu().numpy().reshape(-1, 13)
    for a, ph in enumerate(("gam", "del", "lav")):
        assert np.abs(o[:, a] - G["g_" + ph]).max() <= 1e-14 * np.abs(G["g_" + ph]).max()
        for k, c in enumerate(("Cr", "Nb")):
            ref = G["dg_%s_dx%s" % (ph, c)]
            assert np.abs(o[:, 3 + 2 * a + k] - ref).max() <= 1e-14 * np.abs(ref).max()
    scale = np.maximum(np.abs(G["D_CrCr"]), np.abs(G["D_NbNb"]))
    for k, n in enumerate(("D_CrCr", "D_CrNb", "D_NbCr", "D_NbNb")):
        assert np.all(np.abs(o[:, 9 + k] - G[n]) <= 1e-12 * np.maximum(np.abs(G[n]), 1e-3 * scale))


def test_device_nucleation_probability(gpu, oracle):
    rng = np.random.default_rng(4)
    rows = []
    for _ in range(500):
        xCr, xNb = rng.uniform(0.15, 0.35), rng.uniform(0.01, 0.25)
        dG = oracle.g(0, xCr, xNb)
        rows.append((xCr, xNb, dG, oracle.D(0, 0, xCr, xNb), oracle.D(1, 1, xCr, xNb),
                     rng.uniform(0.008, 0.14), 0.25 * 0.352e-9 ** 3, 10 ** rng.uniform(3, 12),
                     10 ** rng.uniform(-26, -18), 10 ** rng.uniform(-3, 12)))
    rows = np.array(rows)
    out = torch.zeros(2 * len(rows), dtype=torch.float64, device="cuda")
    check(lib().pf_thermo_eval(3, len(rows), _p(_dev(rows)), _p(out), None), "eval")
    o = out.cpu().numpy().reshape(-1, 2)
    for r, (R, P) in zip(rows, o):
        R0, P0 = oracle.nucleation_probability(*r)
        assert R == pytest.approx(R0, rel=1e-14)
        assert P == pytest.approx(P0, rel=1e-12, abs=1e-300)


def test_philox_gpu(gpu, oracle):
    ff = 0xffffffff
    kat = [([0, 0, 0, 0, 0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
           ([ff] * 6, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
           ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822, 0x299f31d0],
            [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]
    rng = np.random.default_rng(0)
    rand = rng.integers(0, 2 ** 32, size=(1000, 6), dtype=np.uint64)
    tup = np.array([k[0] for k in kat] + rand.tolist(), dtype=np.uint32)
    inp = torch.as_tensor(tup.view(np.int32)).cuda()
    out = torch.zeros(4 * len(tup), dtype=torch.int32, device="cuda")
    check(lib().pf_philox_eval(len(tup), _p(inp), _p(out), None), "philox")
    o = out.cpu().numpy().view(np.uint32).reshape(-1, 4)
    for i, (_, want) in enumerate(kat):
        assert o[i].tolist() == want
    for i in range(3, len(tup)):
        assert o[i].tolist() == oracle.philox(tup[i, :4].tolist(), tup[i, 4:].tolist())


# ------------------------------------------------------------------ step


def test_step_golden_fixture(gpu, step_golden):
    """GPU vs the literal replay of the reference's kernel sequence on the
    reference's compiled thermodynamics (no-flux, 14 x 11, 4 steps)."""
    G = step_golden
    F = G["fields"]
    s = M.Setup(N=(F.shape[3] - 2, F.shape[2] - 2), dx=float(G["dx"][0]), dt=float(G["dt"][0]),
                bc=0, DCr=list(G["DCr"]), DNb=list(G["DNb"]), Lmob=list(G["Lmob"]),
                kappa=list(G["kappa"]), omega=list(G["omega"]), alpha=float(G["alpha"][0]),
                mask=list(G["mask"]))
    sim = Simulation(s, [F[0][k] for k in range(4)])
    for step in range(1, F.shape[0]):
        sim.step(1)
        got = sim.fields()
        for k in range(4):
            assert mixed_err(got[k], F[step][k]) <= TOL, (step, k)
            full = np.abs(got[k] - F[step][k])
            full[0, 0] = full[0, -1] = full[-1, 0] = full[-1, -1] = 0.0
            assert full.max() <= TOL * max(np.abs(F[step][k]).max(), 1.0)
    sim.close()


def test_step_golden_fixture_periodic(gpu, step_golden):
    """GPU in periodic mode vs the literal replay with periodic ghost rings
    (14 x 11, 4 steps): interior and ghost ring (the kernel writes each ghost
    cell with the value of the interior cell it wraps to, corners included)."""
    G = step_golden
    F = G["fields_periodic"]
    s = M.Setup(N=(F.shape[3] - 2, F.shape[2] - 2), dx=float(G["dx"][0]), dt=float(G["dt"][0]),
                bc=1, DCr=list(G["DCr"]), DNb=list(G["DNb"]), Lmob=list(G["Lmob"]),
                kappa=list(G["kappa"]), omega=list(G["omega"]), alpha=float(G["alpha"][0]),
                mask=list(G["mask"]))
    sim = Simulation(s, [F[0][k] for k in range(4)], steps_per_launch=1)
    for step in range(1, F.shape[0]):
        sim.step(1)
        got = sim.fields()
        for k in range(4):
            assert mixed_err(got[k], F[step][k]) <= TOL, (step, k)
            full = np.abs(got[k] - F[step][k])
            assert full.max() <= TOL * max(np.abs(F[step][k]).max(), 1.0), (step, k)
    sim.close()
    # and all four steps in one four-step launch (the FP64 default), and in
    # one three-step plus one one-step launch
    for spl in (4, 3):
        sim = Simulation(s, [F[0][k] for k in range(4)], steps_per_launch=spl)
        assert sim.steps_per_launch == spl
        sim.step(4)
        got = sim.fields()
        for k in range(4):
            assert mixed_err(got[k], F[4][k]) <= TOL, (spl, k)
        sim.close()


@pytest.mark.parametrize("bc", [1, 0])
def test_parity_256_1000_steps(gpu, oracle, bc):
    """The 256^2 plain case (2 delta + 2 Laves particles in a noisy matrix),
    1000 steps, against the oracle."""
    s = M.make_setup(256, 256, bc=bc)
    f0 = M.init_matrix_with_particles(256, 256, s.dx, bc, seed=1)
    sim = Simulation(s, f0)
    sim.step(1000)
    got = sim.fields()
    want = oracle.run(f0, oracle_params(oracle, s), bc, 1000)
    for k in range(4):
        e = mixed_err(got[k], want[k])
        print("bc", bc, "field", k, "mixed err %.3e" % e)
        assert e <= TOL
    # the run moved the fields (not a trivially static case)
    assert np.abs(want[2] - f0[2]).max() > 1e-3
    sim.close()


@pytest.mark.parametrize("shape", [(1, 1), (3, 2), (61, 5), (62, 63), (63, 64), (124, 9),
                                   (200, 131), (37, 300)])
@pytest.mark.parametrize("bc", [0, 1])
def test_ragged_shapes(gpu, oracle, shape, bc):
    """Edge and ragged sizes: strips and row chunks that do not divide the
    grid, single cells, tall and wide grids."""
    Nx, Ny = shape
    s = M.make_setup(Nx, Ny, bc=bc, dt_rule=4e-7)
    rng = np.random.default_rng(Nx * 1000 + Ny)
    xCr = 0.3 + 0.01 * rng.standard_normal((Ny + 2, Nx + 2))
    xNb = 0.02 + 0.005 * rng.standard_normal((Ny + 2, Nx + 2))
    pd = np.clip(rng.uniform(-0.3, 1.3, (Ny + 2, Nx + 2)), 0, 1)
    pl = np.clip(rng.uniform(-0.5, 1.0, (Ny + 2, Nx + 2)), 0, 1) * (pd < 0.3)
    f0 = [M.fill_ghosts(a, bc) for a in (xCr, xNb, pd, pl)]
    want = oracle.run(f0, oracle_params(oracle, s), bc, 5)
    results = []
    for rows, remap, depth in ((0, 1, 0), (1, 0, 2), (3, 1, 6), (16, 1, 2), (64, 0, 3),
                               (7, 1, 6)):
        sim = Simulation(s, f0)
        sim.set_tile(rows, remap, depth)
        sim.step(5)
        got = sim.fields()
        for k in range(4):
            assert mixed_err(got[k], want[k]) <= TOL, (rows, remap, depth, k)
            # ghost ring holds the boundary image of the interior
            g = M.fill_ghosts(got[k].copy(), bc)
            assert np.array_equal(g, got[k]), (rows, remap, depth, k)
        results.append(got)
        sim.close()
    # tile shape does not change a single bit
    for r in results[1:]:
        for k in range(4):
            assert np.array_equal(r[k], results[0][k])


def test_fp32_precision_form(gpu, oracle):
    """FP32 storage with FP64 arithmetic against the FP64 oracle."""
    s = M.make_setup(256, 256, bc=1)
    f0 = M.init_matrix_with_particles(256, 256, s.dx, 1, seed=1)
    sim = Simulation(s, f0, dtype=torch.float32)
    sim.step(300)
    got = sim.fields()
    want = oracle.run(f0, oracle_params(oracle, s), 1, 300)
    dx = [np.abs(got[k] - want[k])[1:-1, 1:-1].max() for k in range(4)]
    print("fp32 storage max abs error after 300 steps:", dx)
    assert dx[0] < 5e-6 and dx[1] < 5e-6
    assert dx[2] < 5e-5 and dx[3] < 5e-5
    fr32 = sim.summarize()
    fr64 = oracle.summarize(want)
    assert abs(fr32[3] - fr64[3]) < 1e-4 and abs(fr32[4] - fr64[4]) < 1e-4
    sim.close()


# ------------------------------------------------------------------ nucleation


def _nuc_struct(a, nx, ny, seed, call):
    N = PfNucParams()
    N.sigma_del, N.sigma_lav = a["sigma_del"], a["sigma_lav"]
    N.unit_a, N.ifce_width = a["unit_a"], 2.5e-9
    N.dx = N.dy = a["dx"]
    N.dz = a["dz"]
    N.dt = a["dt"]
    N.seed, N.call, N.instance = seed, call, 0
    return N


def test_nucleation_decisions_bit_exact(gpu, oracle):
    """Identical fields and random stream: identical decisions and embeds."""
    from test_oracle import synthetic_nucleation_case
    for seed, call in ((7, 3), (42, 0), (2 ** 40 + 5, 2 ** 33 + 1)):
        fields, a = synthetic_nucleation_case(oracle, nx=70, ny=45, seed=seed % 1000)
        ny, nx = fields[0].shape
        want, nev, flags = oracle.nucleation(fields, a["sigma_del"], a["sigma_lav"], a["unit_a"],
                                             a["dx"] * nx, a["dx"] * ny, a["dz"], a["dt"],
                                             seed=seed, call=call)
        assert nev > 50
        ctx = ctypes.c_void_p()
        check(lib().pf_ctx_create(ctypes.byref(ctx)), "ctx")
        d = [_dev(x) for x in fields]
        N = _nuc_struct(a, nx, ny, seed, call)
        check(lib().pf_nucleation_f64(ctx, nx, ny, *[_p(x) for x in d], ctypes.byref(N), None),
              "nucleation")
        torch.cuda.synchronize()
        for k in range(4):
            got = d[k].cpu().numpy()
            assert np.array_equal(got, want[k]), k
        wide = ctypes.c_uint64()
        over = ctypes.c_int()
        check(lib().pf_nucleation_status(ctx, ctypes.byref(wide), ctypes.byref(over)), "status")
        assert wide.value == 0 and over.value == 0
        lib().pf_ctx_destroy(ctx)


def test_nucleation_decisions_fp32_storage(gpu, oracle):
    """pf_nucleation_f32: decisions on FP32-stored fields equal the oracle's on
    the same (rounded) values; embeds are the equilibrium compositions rounded
    to FP32."""
    from test_oracle import synthetic_nucleation_case
    fields, a = synthetic_nucleation_case(oracle, nx=70, ny=45, seed=11)
    fields = [f.astype(np.float32).astype(np.float64) for f in fields]
    ny, nx = fields[0].shape
    want, nev, _ = oracle.nucleation(fields, a["sigma_del"], a["sigma_lav"], a["unit_a"],
                                     a["dx"] * nx, a["dx"] * ny, a["dz"], a["dt"], seed=9,
                                     call=4)
    assert nev > 50
    ctx = ctypes.c_void_p()
    check(lib().pf_ctx_create(ctypes.byref(ctx)), "ctx")
    d = [_dev(x, torch.float32) for x in fields]
    N = _nuc_struct(a, nx, ny, 9, 4)
    check(lib().pf_nucleation_f32(ctx, nx, ny, *[_p(x) for x in d], ctypes.byref(N), None),
          "nucleation")
    torch.cuda.synchronize()
    for k in range(4):
        got = d[k].cpu().numpy()
        assert np.array_equal(got, want[k].astype(np.float32)), k
    lib().pf_ctx_destroy(ctx)


def test_fused_nucleation_matches_separate_pass(gpu, oracle):
    """pf_step with nucleation fused == step, then the standalone pass == the
    oracle's step followed by its nucleation pass."""
    from test_oracle import synthetic_nucleation_case
    fields, a = synthetic_nucleation_case(oracle, nx=90, ny=70, seed=5)
    ny, nx = fields[0].shape
    s = M.make_setup(nx - 2, ny - 2, bc=1, dt_rule=1e-7)
    fields = [M.fill_ghosts(f.copy(), 1) for f in fields]
    P = PfParams()
    P.dt = s.dt
    for k in range(2):
        P.DCr[k], P.DNb[k] = s.DCr[k], s.DNb[k]
        P.kappa[k], P.omega[k], P.Lmob[k] = s.kappa[k], s.omega[k], s.Lmob[k]
    P.alpha = s.alpha
    for k in range(9):
        P.mask[k] = s.mask[k]
    N = _nuc_struct(a, nx, ny, 99, 4)
    ctx = ctypes.c_void_p()
    check(lib().pf_ctx_create(ctypes.byref(ctx)), "ctx")
    src = [_dev(x) for x in fields]
    fused = [torch.empty_like(x) for x in src]
    sep = [torch.empty_like(x) for x in src]
    check(lib().pf_step_f64(ctx, nx, ny, 1, ctypes.byref(P), *[_p(x) for x in src],
                            *[_p(x) for x in fused], ctypes.byref(N), None), "fused")
    check(lib().pf_step_f64(ctx, nx, ny, 1, ctypes.byref(P), *[_p(x) for x in src],
                            *[_p(x) for x in sep], None, None), "step")
    check(lib().pf_nucleation_f64(ctx, nx, ny, *[_p(x) for x in sep], ctypes.byref(N), None),
          "nucleation")
    torch.cuda.synchronize()
    stepped = oracle.step(fields, oracle_params(oracle, s), 1)
    want, nev, _ = oracle.nucleation(stepped, a["sigma_del"], a["sigma_lav"], a["unit_a"],
                                     a["dx"] * nx, a["dx"] * ny, a["dz"], a["dt"], seed=99,
                                     call=4)
    assert nev > 50
    for k in range(4):
        f = fused[k].cpu().numpy()
        g = sep[k].cpu().numpy()
        assert np.array_equal(f[1:-1, 1:-1], g[1:-1, 1:-1]), k
        assert mixed_err(f, want[k]) <= TOL, k
    lib().pf_ctx_destroy(ctx)


def test_benchmark_nucleation_is_inert(gpu):
    """At the 4096^2 parameters (dx*nx passed as the mesh spacing) P rounds to
    zero, as in the reference: a nucleation step changes nothing."""
    s = M.make_setup(512, 512, bc=1)
    f0 = M.init_enriched_band(512, 512, 1, seed=42)
    a = Simulation(s, f0, nucleation=True)
    b = Simulation(s, f0, nucleation=False)
    a.step(3)
    b.step(3)
    for x, y in zip(a.fields(), b.fields()):
        assert np.array_equal(x, y)
    a.close()
    b.close()


# ------------------------------------------------------------------ reductions, dataviz


def test_summarize_matches_oracle(gpu, oracle):
    s = M.make_setup(300, 257, bc=0)
    f0 = M.init_matrix_with_particles(300, 257, s.dx, 0, seed=4)
    sim = Simulation(s, f0)
    got = sim.summarize()
    want = oracle.summarize(f0)
    for g, w in zip(got, want):
        assert g == pytest.approx(w, rel=1e-14, abs=1e-16)
    # deterministic: the same answer every time
    assert sim.summarize() == got
    sim.close()


def test_dataviz(gpu):
    rng = np.random.default_rng(1)
    nx, ny = 37, 29
    f = [rng.uniform(0, 1, (ny, nx)) for _ in range(4)]
    d = [_dev(x) for x in f]
    cNi = torch.zeros(nx * ny, dtype=torch.float64, device="cuda")
    phi = torch.zeros(nx * ny, dtype=torch.float64, device="cuda")
    check(lib().pf_dataviz_f64(nx, ny, *[_p(x) for x in d], _p(cNi), _p(phi), None), "dataviz")
    torch.cuda.synchronize()
    p = lambda x: x ** 3 * (6 * x * x - 15 * x + 10)  # noqa: E731
    assert np.allclose(cNi.cpu().numpy().reshape(ny, nx), 1 - f[0] - f[1], rtol=0, atol=1e-15)
    assert np.allclose(phi.cpu().numpy().reshape(ny, nx), p(f[2]) + p(f[3]), rtol=1e-14)


# ------------------------------------------------------------------ size-independent properties


def test_mass_conservation_4096(gpu):
    """At the benchmark size: sum of each composition over the interior is
    conserved by the explicit step (periodic)."""
    s = M.make_setup(4096, 4096, bc=1)
    f0 = M.init_enriched_band(4096, 4096, 1, seed=42)
    sim = Simulation(s, f0, nucleation=True)
    m0 = sim.summarize()
    sim.step(200)
    m1 = sim.summarize()
    assert m1[0] == pytest.approx(m0[0], rel=1e-13)
    assert m1[1] == pytest.approx(m0[1], rel=1e-13)
    sim.close()


def test_parity_4096_benchmark_workload(gpu, oracle):
    """The benchmark's own workload (4096^2 periodic, enriched band, nucleation
    on: step 0 is a nucleation step, whose P rounds to 0) for 8 steps, which the
    stepping loop runs as one fused nucleation launch, one four-step launch and
    one three-step launch: every cell against the oracle."""
    s = M.make_setup(4096, 4096, bc=1)
    f0 = M.init_enriched_band(4096, 4096, 1, seed=42)
    sim = Simulation(s, f0, nucleation=True)
    assert sim.steps_per_launch == 4
    assert list(iter_launches(0, 8, 4, sim.nuc_every)) == [(1, True), (4, False), (3, False)]
    sim.step(8)
    got = sim.fields()
    sim.close()
    want = oracle.run(f0, oracle_params(oracle, s), 1, 8)
    for k in range(4):
        assert mixed_err(got[k], want[k]) <= TOL, k


def test_largest_grid_fp32(gpu):
    """The largest square grid the 32-bit cell index allows (46340^2 with the
    ghost ring, 2.1e9 cells; FP32 storage, 69 GB of fields): a perturbation
    at the last interior cell, whose periodic neighbours wrap to the start of
    the arrays, evolves over a three-step and a one-step launch, and over one
    four-step launch, exactly as the same perturbation at the corner of a 64^2
    grid."""
    def run(n, pos, four):
        s = M.make_setup(n - 2, n - 2, bc=1)
        P = PfParams()
        P.dt = s.dt
        for k in range(2):
            P.DCr[k], P.DNb[k] = s.DCr[k], s.DNb[k]
            P.kappa[k], P.omega[k], P.Lmob[k] = s.kappa[k], s.omega[k], s.Lmob[k]
        P.alpha = s.alpha
        for k in range(9):
            P.mask[k] = s.mask[k]
        vals = (0.30, 0.02, 0.0, 0.0)
        a = [torch.full((n * n,), v, dtype=torch.float32, device="cuda") for v in vals]
        y, x = pos
        bump = (0.05, 0.01, 0.3, 0.2)
        for k in range(4):
            a[k].view(n, n)[y, x] += bump[k]
            a[k].view(n, n)[0, :] = a[k].view(n, n)[n - 2, :]   # periodic ghost ring
            a[k].view(n, n)[n - 1, :] = a[k].view(n, n)[1, :]
            a[k].view(n, n)[:, 0] = a[k].view(n, n)[:, n - 2]
            a[k].view(n, n)[:, n - 1] = a[k].view(n, n)[:, 1]
        b = [torch.empty_like(t) for t in a]
        ctx = ctypes.c_void_p()
        check(lib().pf_ctx_create(ctypes.byref(ctx)), "ctx")
        if four:
            check(lib().pf_step4_f32(ctx, n, n, 1, ctypes.byref(P), *map(_p, a), *map(_p, b),
                                     None), "step4")
            a, b = b, a
        else:
            check(lib().pf_step3_f32(ctx, n, n, 1, ctypes.byref(P), *map(_p, a), *map(_p, b),
                                     None), "step3")
            check(lib().pf_step_f32(ctx, n, n, 1, ctypes.byref(P), *map(_p, b), *map(_p, a),
                                    None, None), "step")
        torch.cuda.synchronize()
        lib().pf_ctx_destroy(ctx)
        # the 11 x 11 interior neighbourhood of the last interior cell, wrapped
        idx = [(n - 2 - 5 + i - 1) % (n - 2) + 1 for i in range(11)]
        rows = torch.tensor(idx, device="cuda")
        out = [t.view(n, n).index_select(0, rows).index_select(1, rows).cpu().numpy()
               for t in a]
        far = [float(t.view(n, n)[(n - 2) // 2, (n - 2) // 2]) for t in a]
        del a, b
        torch.cuda.empty_cache()
        return out, far

    big = 46340
    assert big * big < 2 ** 31 <= (big + 1) * (big + 1)
    small_out, small_far = run(64, (62, 62), False)
    small4_out, small4_far = run(64, (62, 62), True)
    for k in range(4):
        assert np.array_equal(small4_out[k], small_out[k]), k
    for four in (False, True):
        big_out, big_far = run(big, (big - 2, big - 2), four)
        for k in range(4):
            assert np.array_equal(big_out[k], small_out[k]), (four, k)
            assert big_far[k] == small_far[k], (four, k)
    assert np.abs(small_out[2] - small_out[2][0, 0]).max() > 0   # the bump has spread


def test_precision_form_8192(gpu):
    """The precision workload's size (8192^2, the plain workload's
    initialisation, periodic): FP32 storage and FP64 from the same fields for
    30 steps (FP32: ten three-step launches; FP64: seven four-step launches and
    a two-step one).  FP64 conserves each composition sum to
    rounding; FP32 storage stays within a few FP32 ulps of x per step of FP64
    and conserves the sums to FP32 rounding."""
    N = 8192
    s = M.make_setup(N, N, bc=1)
    f0 = M.init_enriched_band(N, N, 1, seed=42)
    s64 = Simulation(s, f0, dtype=torch.float64)
    s32 = Simulation(s, f0, dtype=torch.float32)
    del f0
    m64, m32 = s64.summarize(), s32.summarize()
    s64.step(30)
    s32.step(30)
    e64, e32 = s64.summarize(), s32.summarize()
    for k in range(2):
        assert e64[k] == pytest.approx(m64[k], rel=1e-13)
        assert e32[k] == pytest.approx(m32[k], rel=1e-7)
    worst = 0.0
    for k in range(2):
        a = s64.a[k].reshape(s.ny, s.nx)[1:-1, 1:-1]
        b = s32.a[k].reshape(s.ny, s.nx)[1:-1, 1:-1].double()
        worst = max(worst, float((a - b).abs().max()))
    assert worst < 2e-6, worst
    s64.close()
    s32.close()


def test_errors_are_reported(gpu):
    L = lib()
    ctx = ctypes.c_void_p()
    check(L.pf_ctx_create(ctypes.byref(ctx)), "ctx")
    P = PfParams()
    t = torch.zeros(16, dtype=torch.float64, device="cuda")
    rc = L.pf_step_f64(ctx, 2, 4, 0, ctypes.byref(P), *[_p(t)] * 8, None, None)
    assert rc == 1 and b"interior" in L.pf_last_error()
    rc = L.pf_step_f64(ctx, 4, 4, 0, ctypes.byref(P), *[_p(t)] * 8, None, None)
    assert rc == 1 and b"in place" in L.pf_last_error()
    rc = L.pf_step_f64(ctx, 4, 4, 7, ctypes.byref(P), *[_p(t)] * 8, None, None)
    assert rc == 1
    # a new field on an old field of another name, two new fields on one buffer, or a
    # new field overlapping an old one by part of its length
    old = [torch.zeros(16, dtype=torch.float64, device="cuda") for _ in range(4)]
    new = [torch.zeros(16, dtype=torch.float64, device="cuda") for _ in range(4)]
    big = torch.zeros(64, dtype=torch.float64, device="cuda")
    for outs, what in (([new[0], old[2], new[2], new[3]], b"in place"),
                       ([new[0], new[1], new[1], new[3]], b"each other"),
                       ([big[8:24], big[24:40], big[40:56], new[3]], None)):
        for ins in (old, [big[0:16]] + old[1:]):
            rc = L.pf_step_f64(ctx, 4, 4, 0, ctypes.byref(P), *map(_p, ins), *map(_p, outs),
                               None, None)
            if what is None and ins is old:
                assert rc == 0, L.pf_last_error()   # disjoint views of one buffer are fine
            else:
                assert rc == 1
                assert (what or b"in place") in L.pf_last_error()
    torch.cuda.synchronize()
    with pytest.raises(pfpa.PfError):
        check(L.pf_thermo_eval(9, 1, _p(t), _p(t), None), "eval")
    L.pf_ctx_destroy(ctx)


# ------------------------------------------------------------------ less common paths


@pytest.mark.parametrize("bc", [0, 1])
def test_nine_point_mask(gpu, oracle, bc):
    """A nine-point Laplacian mask (src/mesh.cpp:198-213) selects the variant
    that also shuffles the diagonal neighbours."""
    Nx, Ny = 150, 77
    dx = 0.25e-9
    s = M.make_setup(Nx, Ny, bc=bc, dt_rule=5e-8)
    m = [1 / (6 * dx * dx), 4 / (6 * dx * dx), 1 / (6 * dx * dx),
         4 / (6 * dx * dx), -10 * (2 * dx * dx) / (6 * dx ** 4), 4 / (6 * dx * dx),
         1 / (6 * dx * dx), 4 / (6 * dx * dx), 1 / (6 * dx * dx)]
    s.mask = m
    f0 = M.init_matrix_with_particles(Nx, Ny, dx, bc, seed=6, radius=2e-9)
    sim = Simulation(s, f0)
    sim.step(25)
    got = sim.fields()
    sim.close()
    want = oracle.run(f0, oracle_params(oracle, s), bc, 25)
    for k in range(4):
        assert mixed_err(got[k], want[k]) <= TOL, k


def _masks(dx, dy):
    """The reference's five- and nine-point masks (src/mesh.cpp:187-213)."""
    five = [0.0, 1 / (dy * dy), 0.0,
            1 / (dx * dx), -2 * (dx * dx + dy * dy) / (dx * dx * dy * dy), 1 / (dx * dx),
            0.0, 1 / (dy * dy), 0.0]
    c = 1 / (6 * dx * dy)
    nine = [c, 4 / (6 * dy * dy), c,
            4 / (6 * dx * dx), -10 * (dx * dx + dy * dy) / (6 * dx * dx * dy * dy),
            4 / (6 * dx * dx),
            c, 4 / (6 * dy * dy), c]
    return five, nine


@pytest.mark.parametrize("kind", ["five", "nine", "lopsided"])
@pytest.mark.parametrize("bc", [0, 1])
def test_anisotropic_masks(gpu, oracle, bc, kind):
    """Masks for dy = 1.3 dx.  The reference's nine-point mask then does not
    sum to zero (its entries sum to -2 (dx - dy)^2 / (6 dx^2 dy^2)), so the
    mask sum of a uniform field is not zero and the constant part of
    q = dt D x_gamma contributes to every update; the five-point mask sums to
    zero.  One step per launch and, for the five-point mask, three.
    "lopsided": a five-point mask with unequal west and east weights, which the
    folded symmetric Laplacian cannot take: it runs the general nine-term sum,
    one step per launch, whatever steps per launch are asked for."""
    Nx, Ny = 96, 70
    dx = 0.25e-9
    dy = 1.3 * dx
    s = M.make_setup(Nx, Ny, bc=bc, dt_rule=5e-8)
    five, ninem = _masks(dx, dy)
    if kind == "lopsided":
        five = list(five)
        five[3] *= 1.2
        five[5] *= 0.8
    s.mask = ninem if kind == "nine" else five
    assert M.symmetric_five_point(s.mask) == (kind == "five")
    if kind == "nine":
        assert abs(sum(ninem)) > 1e-3 * max(abs(v) for v in ninem)
    f0 = M.init_matrix_with_particles(Nx, Ny, dx, bc, seed=8, radius=2e-9)
    want = oracle.run(f0, oracle_params(oracle, s), bc, 25)
    for spl in ((1,) if kind == "nine" else (1, 3, 4)):
        sim = Simulation(s, f0, steps_per_launch=spl)
        assert sim.steps_per_launch == (spl if kind == "five" else 1)
        sim.step(25)
        got = sim.fields()
        sim.close()
        for k in range(4):
            assert mixed_err(got[k], want[k]) <= TOL, (spl, k)


def test_wide_nucleation_events(gpu, oracle):
    """Scan radius over the whole grid and embeds of radius ~1.7 cells that
    overlap: the deferred-event path, in the oracle's order."""
    Nx, Ny = 38, 30
    nx, ny = Nx + 2, Ny + 2
    rng = np.random.default_rng(21)
    fields = [M.fill_ghosts(rng.uniform(0.15, 0.3, (ny, nx)), 0),
              M.fill_ghosts(rng.uniform(0.19, 0.24, (ny, nx)), 0),
              np.zeros((ny, nx)), np.zeros((ny, nx))]
    a = dict(sigma_del=0.0105, sigma_lav=0.0108, unit_a=0.352e-9, dx=2e-13, dz=2e-13, dt=3e38)
    want, nev, flags = oracle.nucleation(fields, a["sigma_del"], a["sigma_lav"], a["unit_a"],
                                         a["dx"] * nx, a["dx"] * ny, a["dz"], a["dt"],
                                         seed=77, call=2)
    assert nev > 20
    ctx = ctypes.c_void_p()
    check(lib().pf_ctx_create(ctypes.byref(ctx)), "ctx")
    d = [_dev(x) for x in fields]
    N = _nuc_struct(a, nx, ny, 77, 2)
    check(lib().pf_nucleation_f64(ctx, nx, ny, *[_p(x) for x in d], ctypes.byref(N), None),
          "nucleation")
    torch.cuda.synchronize()
    for k in range(4):
        assert np.array_equal(d[k].cpu().numpy(), want[k]), k
    wide = ctypes.c_uint64()
    over = ctypes.c_int()
    check(lib().pf_nucleation_status(ctx, ctypes.byref(wide), ctypes.byref(over)), "status")
    assert wide.value == nev and over.value == 0
    lib().pf_ctx_destroy(ctx)
