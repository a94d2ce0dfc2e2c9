"""CPU-side checks of libpfpa.so: it loads, exports every symbol the public
headers declare, and its host-side thermodynamics (the same source the
kernels use) match the reference's golden vectors.  No GPU calls."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import pfpa
from pfpa import model as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_c_symbols():
    text = open(os.path.join(ROOT, "include", "pfpa.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pf_\w+)\s*\(", text)))


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", pfpa.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_loads():
    L = pfpa.lib()
    assert L.pf_default_rows(4098, 4098) > 0


def test_exports_every_c_symbol():
    syms = _declared_c_symbols()
    assert len(syms) >= 20
    exp = _exported()
    missing = [s for s in syms if s not in exp]
    assert not missing, missing
    L = pfpa.lib()
    for s in syms:
        assert getattr(L, s) is not None


def test_exports_dropin_cpp_api():
    """The reference's C++ host API (src/data.cuh, src/discretization.cuh,
    src/mesh.h) is exported with C++ linkage."""
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", pfpa.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    for name in ("init_cuda(HostData*", "free_cuda(CudaData*", "read_out_result(CudaData*",
                 "device_init_prng(CudaData*", "device_boundaries(CudaData*",
                 "device_fictitious(CudaData*", "fictitious_boundaries(CudaData*",
                 "device_laplacian(CudaData*", "device_laplacian_boundaries(CudaData*",
                 "device_evolution(CudaData*", "device_nucleation(CudaData*",
                 "device_dataviz(CudaData*", "make_arrays(HostData*", "free_arrays(HostData*",
                 "swap_pointers_1D(double**", "five_point_Laplacian_stencil(double",
                 "checkCuda(hipError_t", "nTiles(int",
                 # additions: boundary mode, seed, grouped launches
                 "device_set_boundary(CudaData*", "device_seed_prng(CudaData*",
                 "device_set_steps_per_launch(CudaData*", "device_flush(CudaData*"):
        assert name in out, name


def test_no_device_calls_fail_cleanly_without_gpu():
    """Without a GPU the library reports an error instead of crashing."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    L = pfpa.lib()
    h = ctypes.c_void_p()
    rc = L.pf_ctx_create(ctypes.byref(h))
    assert rc != 0
    assert L.pf_last_error()


def test_host_thermo_matches_golden(thermo_golden):
    G = thermo_golden
    L = pfpa.lib()
    for x, p, pp in zip(G["p_in"], G["p_out"], G["pPrime_out"]):
        assert L.pf_host_p(x) == pytest.approx(p, rel=1e-14, abs=1e-17)
        assert L.pf_host_pPrime(x) == pytest.approx(pp, rel=1e-14, abs=1e-17)
    comp = G["comp_in"]
    for a, ph in enumerate(("gam", "del", "lav")):
        g = np.array([L.pf_host_g(a, c[0], c[1]) for c in comp])
        assert np.abs(g - G["g_" + ph]).max() <= 1e-14 * np.abs(G["g_" + ph]).max()
    scale = np.maximum(np.abs(G["D_CrCr"]), np.abs(G["D_NbNb"]))
    for (i, j, n) in [(0, 0, "D_CrCr"), (0, 1, "D_CrNb"), (1, 0, "D_NbCr"), (1, 1, "D_NbNb")]:
        d = np.array([L.pf_host_D(i, j, c[0], c[1]) for c in comp])
        assert np.all(np.abs(d - G[n]) <= 1e-12 * np.maximum(np.abs(G[n]), 1e-3 * scale)), n
    fin = G["fict_in"]
    names = ["fict_%s_%s" % (ph, c) for ph in ("gam", "del", "lav") for c in ("Cr", "Nb")]
    out = np.zeros((len(fin), 6))
    buf = (ctypes.c_double * 6)()
    for r, row in enumerate(fin):
        L.pf_host_fict(*[float(v) for v in row], buf)
        out[r] = list(buf)
    phys = np.abs(fin[:, 2] + fin[:, 3] + fin[:, 4] - 1.0) < 1e-12
    for k, n in enumerate(names):
        err = np.abs(out[:, k] - G[n])
        assert err[phys].max() < 1e-12, n
        assert np.all(err[~phys] <= 1e-11 * np.maximum(1.0, np.abs(G[n][~phys]))), n


def test_host_nucleation_probability_matches_oracle(oracle):
    L = pfpa.lib()
    rng = np.random.default_rng(2)
    R = ctypes.c_double()
    P = ctypes.c_double()
    for _ in range(200):
        xCr, xNb = rng.uniform(0.15, 0.35), rng.uniform(0.01, 0.25)
        sigma = rng.uniform(0.008, 0.14)
        dG = oracle.g(0, xCr, xNb)
        args = (xCr, xNb, dG, oracle.D(0, 0, xCr, xNb), oracle.D(1, 1, xCr, xNb), sigma,
                0.25 * 0.352e-9 ** 3, 1e9, 1e-20, 10 ** rng.uniform(-3, 9))
        L.pf_host_nucleation_probability(*args, ctypes.byref(R), ctypes.byref(P))
        r0, p0 = oracle.nucleation_probability(*args)
        assert R.value == pytest.approx(r0, rel=1e-14)
        assert P.value == pytest.approx(p0, rel=1e-12, abs=1e-300)


def test_model_parameters():
    """Diffusivity matrix, mobility and time-step rules at dx = 0.25 nm."""
    DCr, DNb, L = M.set_diffusion_matrix(0.30, 0.02)
    assert DCr[0] == pytest.approx(3.962e-18, rel=1e-3)
    assert DCr[1] == pytest.approx(-3.398e-17, rel=1e-3)
    assert DNb[0] == pytest.approx(3.056e-17, rel=1e-3)
    assert DNb[1] == pytest.approx(2.377e-16, rel=1e-3)
    assert L[0] == pytest.approx(4.0015e-8, rel=1e-4)
    assert M.kappa()[0] == pytest.approx(4.4318e-10, rel=1e-4)
    assert M.omega()[0] == pytest.approx(6.864e8, rel=1e-3)
    dx = 0.25e-9
    # inside delta the composition diffuses ~58.6x faster than D_NbNb
    Dmax = M.max_effective_diffusivity(DCr, DNb)
    assert Dmax == pytest.approx(1.393e-14, rel=2e-3)
    assert Dmax / DNb[1] == pytest.approx(58.6, rel=2e-3)
    assert M.dt_stable(dx, DCr, DNb) == pytest.approx(1.12e-7, rel=3e-3)
    assert M.dt_reference(dx, DCr, DNb, L) == pytest.approx(1.3e-6, rel=1e-12)
    assert M.nuc_interval(M.dt_stable(dx, DCr, DNb)) == 891


def test_max_effective_diffusivity_against_numpy(oracle):
    """Spectral radius of D H_gamma^-1 H_a over the three phases, by numpy."""
    DCr, DNb, _ = M.set_diffusion_matrix(0.30, 0.02)
    Dm = np.array([DCr, DNb])
    H = [np.array([[oracle.d2g(a, 0), oracle.d2g(a, 1)], [oracle.d2g(a, 1), oracle.d2g(a, 2)]])
         for a in range(3)]
    want = max(np.abs(np.linalg.eigvals(Dm @ np.linalg.inv(H[0]) @ H[a])).max() for a in range(3))
    assert M.max_effective_diffusivity(DCr, DNb) == pytest.approx(want, rel=1e-10)


def test_initial_conditions():
    f = M.init_matrix_with_particles(64, 64, 0.25e-9, 1, seed=3, radius=1.25e-9)
    xCr, xNb, pd, pl = f
    assert pd[1:-1, 1:-1].sum() > 0 and pl[1:-1, 1:-1].sum() > 0
    assert np.all((pd[1:-1, 1:-1] == 0) | (pd[1:-1, 1:-1] == 1))
    inside = pd[1:-1, 1:-1] == 1
    assert np.all(xCr[1:-1, 1:-1][inside] == M.lib().pf_host_xe(1, 0))
    g = M.init_enriched_band(128, 64, 1, seed=42)
    band = g[0][1:-1, 1:-1] == M.ENRICHED[0]
    assert band.sum() == 16 * 64
    mat = ~band
    assert g[0][1:-1, 1:-1][mat].min() >= M.MATRIX_CR[0]
    assert g[0][1:-1, 1:-1][mat].max() <= M.MATRIX_CR[1]
    # ghost ring is the periodic image
    assert np.array_equal(g[0][0, :], g[0][-2, :])
    assert np.array_equal(g[0][:, 0], g[0][:, -2])


def test_step_coefficients_against_oracle_paraboloids(oracle):
    """The folded coefficients the step kernels run on (pf_host_step_coefficients)
    against the oracle's paraboloids, evaluated the reference's way:
    x_a(mu) = xe_a + H_a^-1 mu at a common potential mu, then
      q_Cr = dt (D_CrCr x_g,Cr + D_CrNb x_g,Nb)      (src/discretization.cu:859-863)
      30 dt L dOmega_a = 30 dt L [g_g(x_g) - g_a(x_a) - mu.(x_g - x_a)]
                                                    (src/discretization.cu:888-942)
    A dropped term, a wrong sign or a missing factor 2 on the cross term fails."""
    s = M.make_setup(64, 64)
    from pfpa._lib import PfParams
    P = PfParams()
    P.dt = s.dt
    for k in range(2):
        P.DCr[k], P.DNb[k] = s.DCr[k], s.DNb[k]
        P.kappa[k], P.omega[k], P.Lmob[k] = s.kappa[k], s.omega[k], s.Lmob[k]
    P.alpha = s.alpha
    for k in range(9):
        P.mask[k] = s.mask[k]
    out = (ctypes.c_double * 22)()
    assert pfpa.lib().pf_host_step_coefficients(ctypes.byref(P), out) == 0
    c = np.array(out[:])
    # the five-point mask's factor 1/dy^2 rides on q and kappa dt L
    assert M.symmetric_five_point(s.mask)
    scale = s.mask[1]
    c[0:6] /= scale
    c[20:22] /= scale
    qCr, qNb, G = c[0:3], c[3:6], c[6:16].reshape(2, 5)

    def A(a):
        H = np.array([[oracle.d2g(a, 0), oracle.d2g(a, 1)], [oracle.d2g(a, 1), oracle.d2g(a, 2)]])
        return np.linalg.inv(H)

    def xa(a, mu):
        return np.array([oracle.xe(a, 0), oracle.xe(a, 1)]) + A(a) @ mu

    rng = np.random.default_rng(3)
    Hg = np.array([[oracle.d2g(0, 0), oracle.d2g(0, 1)], [oracle.d2g(0, 1), oracle.d2g(0, 2)]])
    for _ in range(50):
        # potentials of compositions across the simplex region the model visits
        x = np.array([rng.uniform(0.0, 0.6), rng.uniform(0.0, 0.3)])
        mu = Hg @ (x - np.array([oracle.xe(0, 0), oracle.xe(0, 1)]))
        xg = xa(0, mu)
        want_q = [s.dt * (s.DCr[0] * xg[0] + s.DCr[1] * xg[1]),
                  s.dt * (s.DNb[0] * xg[0] + s.DNb[1] * xg[1])]
        for q, w in zip((qCr, qNb), want_q):
            assert q[0] * mu[0] + q[1] * mu[1] + q[2] == pytest.approx(w, rel=1e-12)
        for ph in (0, 1):
            a = 1 + ph
            x_a = xa(a, mu)
            dOm = oracle.g(0, *xg) - oracle.g(a, *x_a) - mu @ (xg - x_a)
            want = 30.0 * s.dt * s.Lmob[ph] * dOm
            got = (G[ph, 0] * mu[0] ** 2 + G[ph, 1] * mu[0] * mu[1] + G[ph, 2] * mu[1] ** 2 +
                   G[ph, 3] * mu[0] + G[ph, 4] * mu[1])
            scale = 30.0 * s.dt * s.Lmob[ph] * (abs(oracle.g(0, *xg)) + abs(oracle.g(a, *x_a)) +
                                               abs(mu @ (xg - x_a)))
            assert abs(got - want) <= 1e-12 * scale
    for ph in (0, 1):
        dtL = s.dt * s.Lmob[ph]
        assert c[16 + ph] == pytest.approx(2 * s.omega[ph] * dtL, rel=1e-15)
        assert c[18 + ph] == pytest.approx(2 * s.alpha * dtL, rel=1e-15)
        assert c[20 + ph] == pytest.approx(s.kappa[ph] * dtL, rel=1e-15)
    assert pfpa.lib().pf_host_step_coefficients(None, out) != 0


def test_step_coefficients_mask_fold():
    """A symmetric five-point mask's factor mask[1] moves onto q and kappa dt L
    (the kernels divide their copy of the mask by it); a mask the folded
    Laplacian cannot take (unequal west/east weights, nine points, a zero
    mask[1]) leaves every coefficient as it is."""
    from pfpa._lib import PfParams
    s = M.make_setup(64, 64)

    def coeffs(mask):
        P = PfParams()
        P.dt = s.dt
        for k in range(2):
            P.DCr[k], P.DNb[k] = s.DCr[k], s.DNb[k]
            P.kappa[k], P.omega[k], P.Lmob[k] = s.kappa[k], s.omega[k], s.Lmob[k]
        P.alpha = s.alpha
        for k in range(9):
            P.mask[k] = mask[k]
        out = (ctypes.c_double * 22)()
        assert pfpa.lib().pf_host_step_coefficients(ctypes.byref(P), out) == 0
        return np.array(out[:])

    five = list(s.mask)
    lop = list(five)
    lop[3] *= 1.2
    nine = M.nine_point_mask(s.dx, s.dx)
    x_only = [0.0, 0.0, 0.0, five[3], -2 * five[3], five[3], 0.0, 0.0, 0.0]
    assert M.symmetric_five_point(five)
    for m in (lop, nine, x_only):
        assert not M.symmetric_five_point(m)
    base = coeffs(lop)
    for m in (nine, x_only):
        assert np.array_equal(coeffs(m), base)
    folded = coeffs(five)
    scale = five[1]
    idx = list(range(0, 6)) + [20, 21]
    assert np.array_equal(folded[idx], base[idx] * scale)
    rest = [i for i in range(22) if i not in idx]
    assert np.array_equal(folded[rest], base[rest])


@pytest.mark.parametrize("code", [53, 93, 0])
def test_cpp_set_mask_matches_python_mirror(code):
    """The C++ drop-in's set_mask (src/mesh.cpp:170-213 restated) and the
    Python mirror's build the same doubles, for square and oblong cells; the
    nine-point mask of oblong cells does not sum to zero (as the reference's)."""
    from pfpa import api
    fn = getattr(pfpa.lib(), "_Z8set_maskddiPPdi")
    fn.restype = None
    fn.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                   ctypes.POINTER(ctypes.POINTER(ctypes.c_double)), ctypes.c_int]
    for dx, dy in ((0.25e-9, 0.25e-9), (0.25e-9, 0.325e-9)):
        rows = [(ctypes.c_double * 3)() for _ in range(3)]
        ptrs = (ctypes.POINTER(ctypes.c_double) * 3)(
            *[ctypes.cast(r, ctypes.POINTER(ctypes.c_double)) for r in rows])
        fn(dx, dy, code, ptrs, 3)
        got = [rows[i][j] for i in range(3) for j in range(3)]
        m = np.zeros((3, 3))
        api.set_mask(dx, dy, code, m, 3)
        assert got == m.ravel().tolist()
        if code == 93 and dx != dy:
            assert abs(sum(got)) > 1e-3 * max(abs(v) for v in got)
        else:
            assert abs(sum(got)) <= 1e-12 * max(abs(v) for v in got)


def _rows4_model(Nx, Ny):
    """The four-step kernel's chunk-height model, restated: 2048 resident
    waves (256 CUs x 4 SIMDs x 2), a full round of them costs 2/1.3
    wave-times, a last round of at most one wave per SIMD one wave-time, and
    a wave's time is its rows plus the 3 its levels recompute."""
    strips = -(-Nx // 56)
    best, best_cost = 1, float("inf")
    for rows in range(1, min(Ny, 128) + 1):
        waves = strips * -(-Ny // rows)
        full, rem = divmod(waves, 2048)
        tail = 0.0 if rem == 0 else (1.0 if rem <= 1024 else 2 / 1.3)
        cost = (rows + 3) * (2 / 1.3 * full + tail)
        if cost < best_cost:
            best, best_cost = rows, cost
    return best


def test_default_rows4_model():
    """pf_default_rows4 follows the launch model, whose picks are the heights
    measured fastest (or within 0.7% of them) at 1024^2..8192^2; the other
    multi-step heuristics keep their documented values at 4096^2."""
    L = pfpa.lib()
    for N in (1024, 2048, 4096, 8192):
        assert L.pf_default_rows4(N + 2, N + 2, 0) == {1024: 10, 2048: 38, 4096: 75,
                                                        8192: 119}[N]
    for Nx, Ny in ((4, 4), (5, 300), (300, 5), (256, 256), (1536, 1536), (3072, 3072),
                   (4096, 512), (4096, 1024), (7, 5000), (10000, 64)):
        assert L.pf_default_rows4(Nx + 2, Ny + 2, 0) == _rows4_model(Nx, Ny), (Nx, Ny)
        assert L.pf_default_rows4(Nx + 2, Ny + 2, 1) == _rows4_model(Nx, Ny), (Nx, Ny)
    assert L.pf_default_rows3(4098, 4098, 0) == 48
    assert L.pf_default_rows2(4098, 4098, 0) == 16 and L.pf_default_rows2(4098, 4098, 1) == 32
