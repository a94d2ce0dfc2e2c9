"""The free-energy functional on the GPU (pf_energy_*, pfpa.api.summarize_energy,
Simulation.energy) against the oracle's literal restatement of
src/alloy625.cpp:752-919.

The GPU evaluates the chemical part in the closed form 1/2 mu.r of the KKS
solve; the oracle evaluates the reference's sum of the three paraboloids at the
three fictitious compositions.  Totals: relative 1e-12; per-cell maps:
1e-12 of the largest value.
"""
import ctypes

import numpy as np
import pytest
import torch

from pfpa import api
from pfpa import model as M
from pfpa._lib import check, lib
from pfpa.sim import Simulation

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _fields(Nx, Ny, bc, seed):
    rng = np.random.default_rng(seed)
    xCr = 0.3 + 0.05 * rng.standard_normal((Ny + 2, Nx + 2))
    xNb = 0.05 + 0.02 * rng.standard_normal((Ny + 2, Nx + 2))
    pd = np.clip(rng.uniform(-0.3, 1.3, (Ny + 2, Nx + 2)), 0, 1)
    pl = np.clip(rng.uniform(-0.5, 1.0, (Ny + 2, Nx + 2)), 0, 1) * (pd < 0.3)
    return [M.fill_ghosts(a, bc) for a in (xCr, xNb, pd, pl)]


@pytest.fixture(scope="module")
def ctx(gpu):
    h = ctypes.c_void_p()
    check(lib().pf_ctx_create(ctypes.byref(h)), "pf_ctx_create")
    yield h
    torch.cuda.synchronize()
    lib().pf_ctx_destroy(h)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("bc", [0, 1])
@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (37, 29), (130, 64), (300, 11)])
def test_energy_matches_oracle(ctx, oracle, shape, bc, dtype):
    Nx, Ny = shape
    nx, ny = Nx + 2, Ny + 2
    dx, dy = 2.5e-10, 3.1e-10
    kappa, omega, Omega = M.kappa(), M.omega(), M.OMEGA_3PHASE
    f0 = _fields(Nx, Ny, bc, 17 * Nx + Ny)
    dev = [torch.as_tensor(a).to(device="cuda", dtype=dtype) for a in f0]
    # the oracle sees what the device holds
    host = [a.double().cpu().numpy() for a in dev]
    want, wchem, wgrad = oracle.energy(host, bc, dx, dy, kappa, omega, Omega, maps=True)
    chem = torch.full((ny * nx,), np.nan, dtype=torch.float64, device="cuda")
    grad = torch.full_like(chem, np.nan)
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    fn = lib().pf_energy_f64 if dtype == torch.float64 else lib().pf_energy_f32
    k = (ctypes.c_double * 2)(*kappa)
    w = (ctypes.c_double * 2)(*omega)
    check(fn(ctx, nx, ny, bc, dx, dy, k, w, Omega, *map(_p, dev), _p(chem), _p(grad), _p(out),
             None), "pf_energy")
    got = out.cpu().numpy()
    for a, b in zip(got, want):
        assert a == pytest.approx(b, rel=TOL, abs=1e-30)
    gc = chem.reshape(ny, nx).cpu().numpy()[1:-1, 1:-1]
    gg = grad.reshape(ny, nx).cpu().numpy()[1:-1, 1:-1]
    sc = np.abs(wchem).max() or 1.0
    sg = np.abs(wgrad).max() or 1.0
    assert np.abs(gc - wchem[1:-1, 1:-1]).max() <= TOL * sc
    assert np.abs(gg - wgrad[1:-1, 1:-1]).max() <= TOL * sg
    # ghost ring untouched
    assert np.isnan(chem.reshape(ny, nx).cpu().numpy()[0]).all()
    # without maps the totals are the same bits
    out2 = torch.zeros(3, dtype=torch.float64, device="cuda")
    check(fn(ctx, nx, ny, bc, dx, dy, k, w, Omega, *map(_p, dev), None, None, _p(out2), None),
          "pf_energy")
    assert torch.equal(out, out2)


def test_simulation_energy_and_api(gpu, oracle):
    """Simulation.energy after some steps, and the drop-in mirror's
    summarize_energy filling the host chem_nrg / grad_nrg arrays, against the
    oracle on the same fields."""
    Nx, Ny = 150, 120
    s = M.make_setup(Nx, Ny, bc=0)
    f0 = M.init_matrix_with_particles(Nx, Ny, s.dx, 0, seed=4, radius=2.5e-9)
    sim = Simulation(s, f0)
    sim.step(30)
    tot = sim.energy()
    fields = sim.fields()
    want = oracle.energy(fields, 0, s.dx, s.dx, s.kappa, s.omega, s.alpha)
    for a, b in zip(tot, want):
        assert a == pytest.approx(b, rel=TOL)
    assert want[2] > 0 and want[1] > 0   # interfaces carry gradient and well energy
    sim.close()

    nx, ny = Nx + 2, Ny + 2
    host, dev = api.HostData(), api.CudaData()
    api.make_arrays(host, nx, ny, 3)
    api.five_point_Laplacian_stencil(s.dx, s.dx, host.mask_lap, 3)
    for k, name in enumerate(("conc_Cr_old", "conc_Nb_old", "phi_del_old", "phi_lav_old")):
        getattr(host, name)[...] = fields[k]
    D_Cr, D_Nb, Lmob = M.set_diffusion_matrix(0.30, 0.02)
    api.init_cuda(host, nx, ny, 3, M.kappa(), M.omega(), Lmob, D_Cr, D_Nb, dev)
    got = api.summarize_energy(dev, nx, ny, s.dx, s.dx, host=host)
    want, wchem, wgrad = oracle.energy(fields, 0, s.dx, s.dx, M.kappa(), M.omega(),
                                       M.OMEGA_3PHASE, maps=True)
    for a, b in zip(got, want):
        assert a == pytest.approx(b, rel=TOL)
    assert np.abs(host.chem_nrg[1:-1, 1:-1] - wchem[1:-1, 1:-1]).max() <= \
        TOL * np.abs(wchem).max()
    assert np.abs(host.grad_nrg[1:-1, 1:-1] - wgrad[1:-1, 1:-1]).max() <= \
        TOL * np.abs(wgrad).max()
    api.free_cuda(dev)


def test_energy_errors(ctx):
    f = [torch.zeros(25, dtype=torch.float64, device="cuda") for _ in range(4)]
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    k = (ctypes.c_double * 2)(1.0, 1.0)
    assert lib().pf_energy_f64(ctx, 5, 5, 0, 0.0, 1.0, k, k, 1.0, *map(_p, f), None, None,
                               _p(out), None) != 0
    assert lib().pf_energy_f64(ctx, 5, 5, 3, 1.0, 1.0, k, k, 1.0, *map(_p, f), None, None,
                               _p(out), None) != 0
    assert lib().pf_energy_f64(ctx, 5, 5, 0, 1.0, 1.0, k, k, 1.0, *map(_p, f), None, None,
                               None, None) != 0
