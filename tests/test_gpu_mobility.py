"""The variable-mobility Cahn-Hilliard divergence (chemical_convolution_*,
src/discretization.cu:605-778) on the GPU against the oracle's literal
restatement, over every cell of the padded arrays (ghost ring included).

Tolerance: |a - b| <= 1e-12 max(|a|, |b|, s) per cell with s = max |oracle|
(the fourteen terms of a cell cancel to near zero in smooth regions, so a
purely relative error is meaningless there; the GPU multiplies by 1/(2 h^2)
where the reference divides by 2 h^2, and may fuse multiply-adds)."""
import numpy as np
import pytest
import torch

from pfpa import api

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _case(Nx, Ny, seed, smooth):
    rng = np.random.default_rng(seed)
    shape = (Ny + 2, Nx + 2)
    if smooth:
        Y, X = np.mgrid[0:shape[0], 0:shape[1]]

        def f(a, b, c):
            return a + b * np.sin(2 * np.pi * X / shape[1] * c) * np.cos(2 * np.pi * Y / shape[0])
    else:
        def f(a, b, c):
            return a + b * rng.uniform(-1.0, 1.0, shape)
    conc = [f(0.2, 0.05, 1 + k % 3) for k in range(6)]
    mob_cr = [f(1e-20, 3e-21, 1 + k % 2) for k in range(8)]
    mob_nb = [f(2e-20, 5e-21, 2 - k % 2) for k in range(8)]
    phis = [np.clip(f(0.5, 0.5, 1 + k), 0.0, 1.0) for k in range(2)]
    return conc, mob_cr, mob_nb, phis


def _err(a, b):
    s = max(np.abs(b).max(), 1e-300)
    return float((np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), s)).max())


@pytest.mark.parametrize("shape", [(1, 1), (5, 3), (61, 40), (62, 63), (63, 64), (200, 131),
                                   (1000, 333)])
@pytest.mark.parametrize("smooth", [False, True])
def test_mobility_divergence_matches_oracle(gpu, oracle, shape, smooth):
    Nx, Ny = shape
    conc, mc, mn, phis = _case(Nx, Ny, 3 + Nx, smooth)
    dx, dy = 0.25e-9, 0.3e-9
    want = oracle.mobility_divergence(conc, mc, mn, phis[0], phis[1], dx, dy)
    dev = lambda a: torch.as_tensor(a).cuda()   # noqa: E731
    got = api.chemical_convolution([dev(a) for a in conc], [dev(a) for a in mc],
                                   [dev(a) for a in mn], dev(phis[0]), dev(phis[1]), dx, dy)
    torch.cuda.synchronize()
    for k in range(2):
        g = got[k].cpu().numpy()
        e = _err(g, want[k])
        print("shape", shape, "smooth", smooth, "component", k, "err %.3e" % e)
        assert e <= TOL
        assert np.abs(want[k]).max() > 0.0


def test_mobility_divergence_rejects_overlap(gpu):
    z = [torch.zeros(6, 7, dtype=torch.float64, device="cuda") for _ in range(24)]
    with pytest.raises(api._lib.PfError):
        api.chemical_convolution(z[:6], z[6:14], z[14:22], z[22], z[23], 1.0, 1.0,
                                 div_Cr=z[0])
