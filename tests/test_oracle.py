"""Pins of the CPU oracle: golden vectors from the reference's compiled code,
and what the mathematics fixes (closed forms, invariants, known answers)."""
import math

import mpmath
import numpy as np
import pytest

# ------------------------------------------------------------------ golden vectors


def test_constants_match_reference(oracle, thermo_golden):
    G = thermo_golden
    assert oracle.kT() == G["const_kT"][0]
    assert oracle.RT() == G["const_RT"][0]
    assert oracle.Vm() == G["const_Vm"][0]
    for a, ph in enumerate(("gam", "del", "lav")):
        assert oracle.xe(a, 0) == G["const_xe_%s_Cr" % ph][0]
        assert oracle.xe(a, 1) == G["const_xe_%s_Nb" % ph][0]
        assert oracle.d2g(a, 0) == G["const_d2g_%s_dxCrCr" % ph][0]
        assert oracle.d2g(a, 1) == G["const_d2g_%s_dxCrNb" % ph][0]
        assert oracle.d2g(a, 1) == G["const_d2g_%s_dxNbCr" % ph][0]
        assert oracle.d2g(a, 2) == G["const_d2g_%s_dxNbNb" % ph][0]


def test_p_golden(oracle, thermo_golden):
    G = thermo_golden
    for x, p, pp in zip(G["p_in"], G["p_out"], G["pPrime_out"]):
        # reference host form uses pow(); identical up to rounding
        assert oracle.p(x) == pytest.approx(p, rel=1e-14, abs=1e-17)
        assert oracle.pPrime(x) == pytest.approx(pp, rel=1e-14, abs=1e-17)


def test_g_dg_golden(oracle, thermo_golden):
    G = thermo_golden
    comp = G["comp_in"]
    for a, ph in enumerate(("gam", "del", "lav")):
        g = np.array([oracle.g(a, c[0], c[1]) for c in comp])
        # absolute scale: the paraboloid's typical magnitude over the set
        scale = np.abs(G["g_" + ph]).max()
        assert np.abs(g - G["g_" + ph]).max() <= 1e-14 * scale
        for k, cc in enumerate(("Cr", "Nb")):
            d = np.array([oracle.dg(a, k, c[0], c[1]) for c in comp])
            ref = G["dg_%s_dx%s" % (ph, cc)]
            assert np.abs(d - ref).max() <= 1e-14 * np.abs(ref).max()


def test_fict_golden(oracle, thermo_golden):
    """The 2x2 KKS solve reproduces the reference's sympy rational forms."""
    G = thermo_golden
    fin = G["fict_in"]
    names = ["fict_%s_%s" % (ph, c) for ph in ("gam", "del", "lav") for c in ("Cr", "Nb")]
    out = np.array([oracle.fict(*r) for r in fin])
    phys = np.abs(fin[:, 2] + fin[:, 3] + fin[:, 4] - 1.0) < 1e-12
    for k, n in enumerate(names):
        ref = G[n]
        err = np.abs(out[:, k] - ref)
        # physical phase fractions (p sum to one): 1e-12 absolute
        assert err[phys].max() < 1e-12, n
        # independent pGam: the solve is as well conditioned as det(M) allows
        assert np.all(err[~phys] <= 1e-11 * np.maximum(1.0, np.abs(ref[~phys]))), n


def test_D_golden(oracle, thermo_golden):
    G = thermo_golden
    comp = G["comp_in"]
    for (i, j, n) in [(0, 0, "D_CrCr"), (0, 1, "D_CrNb"), (1, 0, "D_NbCr"), (1, 1, "D_NbNb")]:
        d = np.array([oracle.D(i, j, c[0], c[1]) for c in comp])
        ref = G[n]
        # the reference sums three mobility terms of either sign; compare on
        # the scale of the largest entry at each composition
        scale = np.maximum(np.abs(G["D_CrCr"]), np.abs(G["D_NbNb"]))
        assert np.all(np.abs(d - ref) <= 1e-12 * np.maximum(np.abs(ref), 1e-3 * scale)), n


def test_step_golden_literal_replay(oracle, step_golden):
    """Oracle vs a literal replay of the reference's kernel sequence (ghost
    ring, boundary copies, mask convolution, CH/AC kernels) evaluated with the
    reference's compiled thermodynamics, no-flux, 4 steps."""
    G = step_golden
    F = G["fields"]
    P = oracle.make_params(G["dt"][0], G["DCr"], G["DNb"], G["kappa"], G["omega"], G["Lmob"],
                           G["alpha"][0], G["mask"])
    cur = [F[0][k].copy() for k in range(4)]
    for s in range(1, F.shape[0]):
        cur = oracle.step(cur, P, 0)
        for k in range(4):
            ref = F[s][k][1:-1, 1:-1]
            a = cur[k][1:-1, 1:-1]
            sf = np.abs(ref).max()
            rel = np.abs(a - ref) / np.maximum(np.maximum(np.abs(a), np.abs(ref)), sf)
            assert rel.max() <= 1e-12, (s, k)
            # edges of the ghost ring (corners race in the reference)
            full = np.abs(cur[k] - F[s][k])
            full[0, 0] = full[0, -1] = full[-1, 0] = full[-1, -1] = 0.0
            assert full.max() <= 1e-12 * max(sf, 1.0)


def test_step_golden_periodic_replay(oracle, step_golden):
    """Oracle in periodic mode vs the same literal replay with the ghost rings
    filled by the periodic wrap instead of the reference's copy (14 x 11,
    4 steps): pins the periodic index map, which the reference does not have,
    by something other than the oracle's own boundary code."""
    G = step_golden
    F = G["fields_periodic"]
    P = oracle.make_params(G["dt"][0], G["DCr"], G["DNb"], G["kappa"], G["omega"], G["Lmob"],
                           G["alpha"][0], G["mask"])
    cur = [F[0][k].copy() for k in range(4)]
    for s in range(1, F.shape[0]):
        cur = oracle.step(cur, P, 1)
        for k in range(4):
            ref = F[s][k][1:-1, 1:-1]
            a = cur[k][1:-1, 1:-1]
            sf = np.abs(ref).max()
            rel = np.abs(a - ref) / np.maximum(np.maximum(np.abs(a), np.abs(ref)), sf)
            assert rel.max() <= 1e-12, (s, k)
    # the periodic run differs from the no-flux one (the boundary matters here)
    assert np.abs(F[-1][2] - G["fields"][-1][2])[1:-1, 1:-1].max() > 1e-6


# ------------------------------------------------------------------ closed forms


def test_p_closed_form(oracle):
    assert oracle.p(0.0) == 0.0
    assert oracle.p(1.0) == 1.0
    assert oracle.p(0.5) == 0.5
    assert oracle.pPrime(0.0) == 0.0 and oracle.pPrime(1.0) == 0.0
    # p' is the derivative of p
    for x in np.linspace(-0.2, 1.2, 29):
        h = 1e-6
        fd = (oracle.p(x + h) - oracle.p(x - h)) / (2 * h)
        assert oracle.pPrime(x) == pytest.approx(fd, rel=1e-7, abs=1e-9)


def test_paraboloid_properties(oracle):
    for a in range(3):
        xe = (oracle.xe(a, 0), oracle.xe(a, 1))
        assert abs(oracle.g(a, *xe)) < 1e-6
        assert abs(oracle.dg(a, 0, *xe)) < 1e-3 and abs(oracle.dg(a, 1, *xe)) < 1e-3
        # Hessian by central differences of the gradient equals d2g
        h = 1e-4
        for k, (dc, dn) in enumerate(((h, 0), (0, h))):
            for comp in range(2):
                fd = (oracle.dg(a, comp, xe[0] + dc, xe[1] + dn) -
                      oracle.dg(a, comp, xe[0] - dc, xe[1] - dn)) / (2 * h)
                want = oracle.d2g(a, k + comp)
                assert fd == pytest.approx(want, rel=1e-7)
        # quadratic: g(xe + t v) = t^2 g(xe + v)
        v = (0.013, -0.007)
        g1 = oracle.g(a, xe[0] + v[0], xe[1] + v[1])
        g3 = oracle.g(a, xe[0] + 3 * v[0], xe[1] + 3 * v[1])
        assert g3 == pytest.approx(9 * g1, rel=1e-10)


def test_fict_invariants(oracle):
    rng = np.random.default_rng(3)
    for _ in range(300):
        xCr, xNb = rng.uniform(0.05, 0.5), rng.uniform(0.01, 0.25)
        pd = oracle.p(rng.uniform(0, 1))
        pl = oracle.p(rng.uniform(0, 1)) * (1 - pd)
        pg = 1 - pd - pl
        f = oracle.fict(xCr, xNb, pd, pg, pl)
        # mass balance
        assert pg * f[0] + pd * f[2] + pl * f[4] == pytest.approx(xCr, abs=1e-14)
        assert pg * f[1] + pd * f[3] + pl * f[5] == pytest.approx(xNb, abs=1e-14)
        # equal diffusion potentials across phases
        mu = [(oracle.dg(a, 0, f[2 * a], f[2 * a + 1]), oracle.dg(a, 1, f[2 * a], f[2 * a + 1]))
              for a in range(3)]
        for a in (1, 2):
            assert mu[a][0] == pytest.approx(mu[0][0], rel=1e-9, abs=1e-2)
            assert mu[a][1] == pytest.approx(mu[0][1], rel=1e-9, abs=1e-2)
    # pure gamma: x_gamma = x
    f = oracle.fict(0.3, 0.02, 0.0, 1.0, 0.0)
    assert f[0] == pytest.approx(0.3, abs=1e-15) and f[1] == pytest.approx(0.02, abs=1e-15)


def test_fict_jacobian_in_delta(oracle):
    """Inside pure delta, dx_gamma/dx = H_gamma^-1 H_delta; its entries are the
    numbers that follow from the reference's rational forms."""
    x0 = (0.05, 0.24)
    h = 1e-6
    J = np.zeros((2, 2))
    for j, (dc, dn) in enumerate(((h, 0), (0, h))):
        fp = oracle.fict(x0[0] + dc, x0[1] + dn, 1.0, 0.0, 0.0)
        fm = oracle.fict(x0[0] - dc, x0[1] - dn, 1.0, 0.0, 0.0)
        J[0, j] = (fp[0] - fm[0]) / (2 * h)
        J[1, j] = (fp[1] - fm[1]) / (2 * h)
    want = np.array([[16.349, -334.807], [-2.0826, 101.687]])
    assert np.allclose(J, want, rtol=2e-4)


def _uniform_fields(nx, ny, xCr, xNb, pd, pl):
    return [np.full((ny, nx), v) for v in (xCr, xNb, pd, pl)]


def _params(oracle, dt, dx=0.25e-9):
    DCr, DNb, L = oracle.set_diffusion_matrix(0.30, 0.02, 2.5e-9)
    kappa = [3.0 * 0.13 * 2.5e-9 / 2.2] * 2
    omega = [18.0 * 0.13 * 0.13 / kappa[0]] * 2
    return oracle.make_params(dt, DCr, DNb, kappa, omega, L, 1.07e11,
                              oracle.five_point_mask(dx, dx))


def test_fixed_points(oracle):
    P = _params(oracle, 1e-7)
    for bc in (0, 1):
        f = _uniform_fields(9, 7, 0.3, 0.02, 0.0, 0.0)
        out = oracle.step(f, P, bc)
        for a, b in zip(out, f):
            assert np.array_equal(a, b)
        f = _uniform_fields(9, 7, oracle.xe(1, 0), oracle.xe(1, 1), 1.0, 0.0)
        out = oracle.step(f, P, bc)
        for a, b in zip(out, f):
            assert np.abs(a - b).max() < 1e-14


def test_laplacian_quadratic_and_fourier(oracle):
    dx = 0.25e-9
    dt = 1e-7
    P = _params(oracle, dt, dx)
    # quadratic in x and y, pure gamma (x_gamma = x): CH increment is
    # dt * D . (2a + 2b) away from the boundary
    ny, nx = 12, 14
    yy, xx = np.mgrid[0:ny, 0:nx].astype(float)
    a, b, c, d = 2e-5, -1e-5, 3e-6, 4e-6
    xCr = 0.3 + a * xx ** 2 + b * yy ** 2
    xNb = 0.02 + c * xx ** 2 + d * yy ** 2
    z = np.zeros_like(xCr)
    out = oracle.step([xCr, xNb, z, z], P, 0)
    lapCr = 2 * (a + b) / dx ** 2
    lapNb = 2 * (c + d) / dx ** 2
    want = dt * (P.DCr[0] * lapCr + P.DCr[1] * lapNb)
    inc = (out[0] - xCr)[2:-2, 2:-2]
    assert np.allclose(inc, want, rtol=1e-6, atol=0)
    # Fourier mode, periodic: eigenvalue -(4/dx^2) sin^2(k dx / 2)
    Nx, Ny = 16, 6
    k = 2 * math.pi * 3 / (Nx * dx)
    col = np.arange(Nx)
    eps = 1e-4
    row = 0.3 + eps * np.cos(k * col * dx)
    xCr = np.zeros((Ny + 2, Nx + 2))
    xCr[1:-1, 1:-1] = row
    xNb = np.full_like(xCr, 0.02)
    z = np.zeros_like(xCr)
    out = oracle.step([xCr, xNb, z, z], P, 1)
    lam = -(4 / dx ** 2) * math.sin(k * dx / 2) ** 2
    want = dt * P.DCr[0] * lam * eps * np.cos(k * col * dx)
    inc = (out[0] - xCr)[1:-1, 1:-1]
    # the increment is recovered next to 0.3: one ulp of 0.3 is 5.6e-17
    assert np.allclose(inc, want[None, :], rtol=1e-10, atol=6e-17)


def test_mass_conservation(oracle):
    import pfpa.model as M  # noqa: F401  (imports the host package; no GPU needed)
    rng = np.random.default_rng(5)
    P = _params(oracle, 2e-7)
    for bc in (0, 1):
        ny, nx = 22, 26
        xCr = 0.3 + 1e-2 * rng.standard_normal((ny, nx))
        xNb = 0.02 + 5e-3 * rng.standard_normal((ny, nx))
        pd = np.clip(rng.uniform(-0.2, 1.2, (ny, nx)), 0, 1) * (rng.uniform(size=(ny, nx)) > 0.7)
        pl = np.zeros((ny, nx))
        f = [xCr, xNb, pd, pl]
        m0 = [f[0][1:-1, 1:-1].sum(), f[1][1:-1, 1:-1].sum()]
        out = oracle.run(f, P, bc, 20)
        m1 = [out[0][1:-1, 1:-1].sum(), out[1][1:-1, 1:-1].sum()]
        assert m1[0] == pytest.approx(m0[0], rel=1e-13)
        assert m1[1] == pytest.approx(m0[1], rel=1e-13)


def test_free_energy_decreases(oracle):
    """Allen-Cahn with no chemical driving force is a gradient flow: the
    discrete double-well + gradient energy must not increase."""
    dx = 0.25e-9
    P = _params(oracle, 1e-6, dx)
    rng = np.random.default_rng(9)
    ny, nx = 18, 18
    pd = np.clip(0.5 + 0.4 * rng.standard_normal((ny, nx)), 0, 1)
    # x at delta equilibrium where phi is 1, gamma equilibrium elsewhere makes
    # mu = 0 only for pure cells; use pure-gamma compositions and zero mobility
    # of composition by taking x = xe_gamma everywhere and phi small
    xCr = np.full((ny, nx), oracle.xe(0, 0))
    xNb = np.full((ny, nx), oracle.xe(0, 1))
    pl = np.zeros((ny, nx))

    def energy(f):
        phi = f[2][1:-1, 1:-1]
        ext = f[2]
        gx = (ext[1:-1, 2:] - ext[1:-1, 1:-1]) / dx
        gy = (ext[2:, 1:-1] - ext[1:-1, 1:-1]) / dx
        w = P.omega[0] * phi ** 2 * (1 - phi) ** 2
        return float((w + 0.5 * P.kappa[0] * (gx ** 2 + gy ** 2)).sum())

    f = [xCr, xNb, pd, pl]
    e_prev = energy(f)
    for _ in range(30):
        f = oracle.step(f, P, 1)
        e = energy(f)
        assert e <= e_prev * (1 + 1e-12)
        e_prev = e


# ------------------------------------------------------------------ nucleation


def test_nucleation_probability_textbook(oracle):
    """Classical nucleation theory in its textbook form, evaluated at 40
    digits: R* = 2 sigma/dG, G* = 4/3 pi sigma R*^2,
    Z = V_a/(2 pi R*^2) sqrt(sigma/kT), beta*_i = 4 pi R*^2 D_ii x_i / a^4,
    J_i = beta*_i Z n exp(-G*/(kT dc_i)), P = 1 - exp(-dt dV sum J)."""
    mpmath.mp.dps = 40
    kT = mpmath.mpf(oracle.kT())
    a = mpmath.mpf("0.352e-9")
    xeCr, xeNb = mpmath.mpf(oracle.xe(0, 0)), mpmath.mpf(oracle.xe(0, 1))
    cases = [
        # xCr, xNb, sigma, dt, dV scale
        (0.30, 0.20, 0.01, 1e-4, 1.0),
        (0.25, 0.22, 0.012, 1e-3, 1.0),
        (0.20, 0.23, 0.008, 1e-6, 1.0),
        (0.30, 0.02, 0.13, 1e-4, 1.0),     # matrix: underflows to P = 0
    ]
    for (xCr, xNb, sigma, dt, _) in cases:
        dG = oracle.g(0, xCr, xNb)
        DCrCr = oracle.D(0, 0, xCr, xNb)
        DNbNb = oracle.D(1, 1, xCr, xNb)
        Vatom = 0.25 * 0.352e-9 ** 3
        dV = (0.25e-9 * 64) ** 2 * 0.25e-9
        n_gam = dV / Vatom
        Rs, Pn = oracle.nucleation_probability(xCr, xNb, dG, DCrCr, DNbNb, sigma, Vatom, n_gam,
                                               dV, dt)
        s = mpmath.mpf(sigma)
        R = 2 * s / mpmath.mpf(dG)
        Gs = mpmath.mpf(4) / 3 * mpmath.pi * s * R ** 2
        Z = mpmath.mpf(Vatom) / (2 * mpmath.pi * R ** 2) * mpmath.sqrt(s / kT)
        J = 0
        for x, D, dc in ((xCr, DCrCr, xeCr - mpmath.mpf(xCr)), (xNb, DNbNb, mpmath.mpf(xNb) - xeNb)):
            beta = 4 * mpmath.pi * R ** 2 * mpmath.mpf(D) * mpmath.mpf(x) / a ** 4
            J += beta * Z * mpmath.mpf(n_gam) * mpmath.exp(-Gs / (kT * dc))
        lam = mpmath.mpf(dt) * mpmath.mpf(dV) * J
        assert Rs == pytest.approx(float(R), rel=1e-14)
        # the reference forms 1 - exp(-lambda) in FP64 (rounds to 0 below ~1e-16)
        assert Pn == pytest.approx(1.0 - math.exp(-float(lam)), rel=1e-9, abs=1e-300)
        if lam > 1e-6:
            assert Pn == pytest.approx(float(1 - mpmath.exp(-lam)), rel=1e-9)


def _rate(oracle, xCr, xNb, sigma, Vatom, n_gam, dV):
    """lambda / dt = dV (J_Cr + J_Nb), recovered from P at a small dt."""
    dG = oracle.g(0, xCr, xNb)
    args = (xCr, xNb, dG, oracle.D(0, 0, xCr, xNb), oracle.D(1, 1, xCr, xNb), sigma, Vatom,
            n_gam, dV)
    _, P = oracle.nucleation_probability(*args, 1.0)
    return -math.log1p(-P) if P < 1 else float("inf")


def test_nucleation_poisson_law(oracle):
    """P over 2 dt is the probability of at least one event in two
    independent intervals: 1 - (1 - P(dt))^2."""
    xCr, xNb, sigma = 0.27, 0.21, 0.011
    Vatom = 0.25 * 0.352e-9 ** 3
    dV = 1e-5 * 1e-5 * 0.25e-9
    n_gam = dV / Vatom
    rate = _rate(oracle, xCr, xNb, sigma, Vatom, n_gam, dV)
    assert rate > 0
    dt = 0.5 / rate
    dG = oracle.g(0, xCr, xNb)
    args = (xCr, xNb, dG, oracle.D(0, 0, xCr, xNb), oracle.D(1, 1, xCr, xNb), sigma, Vatom,
            n_gam, dV)
    _, P1 = oracle.nucleation_probability(*args, dt)
    _, P2 = oracle.nucleation_probability(*args, 2 * dt)
    assert 0.2 < P1 < 0.8
    assert P2 == pytest.approx(1 - (1 - P1) ** 2, rel=1e-12)


def test_philox_known_answers(oracle):
    """Random123 known-answer vectors for Philox4x32-10."""
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    ff = 0xffffffff
    assert oracle.philox([ff] * 4, [ff, ff]) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert oracle.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                         [0xa4093822, 0x299f31d0]) == [0xd16cfe09, 0x94fdcceb, 0x5001e420,
                                                       0x24126ea1]


def test_uniforms_in_unit_interval(oracle):
    us = [oracle.nuc_uniforms(42, call, 0, cell) for call in range(3) for cell in range(2000)]
    u = np.array(us).ravel()
    assert u.min() > 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 0.01


# Random123 known answers as (counter words, key words, output words).
_PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


def test_uniforms_bit_mapping_from_known_answers(oracle):
    """The nucleation stream's word layout and 53-bit mapping, pinned on the
    Philox known answers: counter = (cell, instance, call lo, call hi), key =
    (seed lo, seed hi); u1 = ((w1:w0) >> 11) + 1) / 2^53 from words 0-1 and u2
    from words 2-3, so u lies in (0, 1].  Computed here with exact integer and
    rational arithmetic from the published output words."""
    from fractions import Fraction
    for ctr, key, w in _PHILOX_KAT:
        cell, instance = ctr[0], ctr[1]
        call = ctr[2] | (ctr[3] << 32)
        seed = key[0] | (key[1] << 32)
        u1, u2 = oracle.nuc_uniforms(seed, call, instance, cell)
        e1 = Fraction((((w[1] << 32) | w[0]) >> 11) + 1, 1 << 53)
        e2 = Fraction((((w[3] << 32) | w[2]) >> 11) + 1, 1 << 53)
        # every value k/2^53 with 1 <= k <= 2^53 is a double: the mapping is exact
        assert Fraction(u1) == e1
        assert Fraction(u2) == e2
    # the ends of the interval: the smallest value is 2^-53, the largest 1
    assert oracle.nuc_uniforms(0, 0, 0, 0)[0] > 0.0


def test_diffusion_matrix_and_mobility_pinned(oracle, thermo_golden):
    """set_diffusion_matrix (src/alloy625.cpp:80-109) at (0.30, 0.02): D is the
    reference's compiled D_* at that composition, and the phase mobility is
    L = MobStab |(D_CrCr, D_NbNb)| / (w^2 RT / Vm) with MobStab = 1 and
    w = 2.5 nm (src/parameters.h:19-22), the same for both phases."""
    G = thermo_golden
    assert tuple(G["comp_in"][0]) == (0.30, 0.02)
    w = 2.5e-9
    DCr, DNb, L = oracle.set_diffusion_matrix(0.30, 0.02, w)
    assert DCr[0] == pytest.approx(G["D_CrCr"][0], rel=1e-14)
    assert DCr[1] == pytest.approx(G["D_CrNb"][0], rel=1e-14)
    assert DNb[0] == pytest.approx(G["D_NbCr"][0], rel=1e-14)
    assert DNb[1] == pytest.approx(G["D_NbNb"][0], rel=1e-14)
    # the same expression evaluated by hand at (0.30, 0.02): L = 4.0015e-8 m^2/(N s)
    assert L[0] == pytest.approx(4.0015e-8, rel=1e-4)
    assert L[1] == L[0]
    # the norm is over the diagonal (D_CrCr, D_NbNb), not a row or column: the
    # off-diagonal entries are ~10x D_CrCr, so mixing them in moves L by > 0.1%
    RT, Vm = G["const_RT"][0], G["const_Vm"][0]
    diag = math.hypot(G["D_CrCr"][0], G["D_NbNb"][0])
    assert L[0] == pytest.approx(diag / (w * w * RT / Vm), rel=1e-14)
    for wrong in (math.hypot(G["D_CrCr"][0], G["D_CrNb"][0]),
                  math.hypot(G["D_NbCr"][0], G["D_NbNb"][0])):
        assert abs(wrong / (w * w * RT / Vm) / L[0] - 1.0) > 1e-3
    # L scales as 1/w^2 (the interface width enters squared)
    _, _, L2 = oracle.set_diffusion_matrix(0.30, 0.02, 2.0 * w)
    assert L2[0] == pytest.approx(L[0] / 4.0, rel=1e-15)
    # the transformation-limited step of src/main.cpp:260 with kappa of
    # src/parameters.h:43: dx^2 / (4 L kappa) = 8.811e-4 s at dx = 0.25 nm
    kappa = 3.0 * 0.13 * w / 2.2
    assert (0.25e-9) ** 2 / (4.0 * L[0] * kappa) == pytest.approx(8.811e-4, rel=1e-3)


def _summary_fields(Nx, Ny, ghost):
    return [np.full((Ny + 2, Nx + 2), ghost) for _ in range(4)]


def test_summarize_uniform_fields_exact(oracle):
    """summarize_fields (src/alloy625.cpp:811-863) on uniform interiors gives
    exactly (x_Cr, x_Nb, 1 - p(phi_d) - p(phi_l), p(phi_d), p(phi_l)); p(0) = 0,
    p(1/2) = 1/2, p(1) = 1 are exact in binary.  The ghost ring holds NaN or
    1e300, so any ghost cell in the sum shows."""
    Nx, Ny = 12, 9
    cases = [(0.0, 0.0, (1.0, 0.0, 0.0)), (0.5, 0.0, (0.5, 0.5, 0.0)),
             (1.0, 0.0, (0.0, 1.0, 0.0)), (0.0, 1.0, (0.0, 0.0, 1.0)),
             (0.5, 0.5, (0.0, 0.5, 0.5))]
    for ghost in (np.nan, 1e300):
        for pd, pl, (gam, dl, lv) in cases:
            f = _summary_fields(Nx, Ny, ghost)
            for a, v in zip(f, (0.25, 0.0625, pd, pl)):
                a[1:-1, 1:-1] = v
            out = oracle.summarize(f)
            assert out == [0.25, 0.0625, gam, dl, lv], (ghost, pd, pl, out)


def test_summarize_dyadic_sum(oracle):
    """A field of dyadic values whose sums are exact in binary: the means are
    the hand-computed ones.  x_Cr = k/64 over a 4 x 3 interior in row order
    (k = 0..11) sums to 66/64, so its mean is 66/64/12 = 11/128; x_Nb is 1/8 on
    the first row and 0 elsewhere (mean 1/24 up to rounding of the division);
    phi_d = 1 on two cells, 1/2 on two (p = 1, 1/2), so delta = 3/12 = 1/4;
    phi_l = 1 on one cell, so Laves = 1/12 and gamma = 1 - 1/4 - 1/12 = 2/3."""
    Nx, Ny = 4, 3
    f = _summary_fields(Nx, Ny, np.nan)
    f[0][1:-1, 1:-1] = (np.arange(12, dtype=np.float64) / 64.0).reshape(3, 4)
    f[1][1:-1, 1:-1] = 0.0
    f[1][1, 1:-1] = 0.125
    f[2][1:-1, 1:-1] = 0.0
    f[2][1, 1] = f[2][2, 3] = 1.0
    f[2][3, 2] = f[2][3, 4] = 0.5
    f[3][1:-1, 1:-1] = 0.0
    f[3][2, 1] = 1.0
    out = oracle.summarize(f)
    assert out[0] == 11.0 / 128.0
    assert out[1] == pytest.approx(1.0 / 24.0, rel=1e-16)
    assert out[3] == 0.25
    assert out[4] == pytest.approx(1.0 / 12.0, rel=1e-16)
    assert out[2] == pytest.approx(2.0 / 3.0, rel=1e-16)
    # the phases are the interpolated fractions, not the raw order parameter:
    # phi_d = 1/4 gives p = 0.103515625 (exact in binary), not 1/4
    f[2][1:-1, 1:-1] = 0.25
    f[3][1:-1, 1:-1] = 0.0
    out = oracle.summarize(f)
    assert out[3] == 0.103515625
    assert out[2] == 1.0 - 0.103515625


def synthetic_nucleation_case(oracle, nx=40, ny=34, seed=11):
    """Fields and nucleation arguments for which P lies inside (0, 1) on most
    cells (at the benchmark's own parameters P rounds to 0).  The domain is
    given a large mesh spacing and the pass a long interval."""
    rng = np.random.default_rng(seed)
    xCr = rng.uniform(0.15, 0.3, (ny, nx))
    xNb = rng.uniform(0.19, 0.24, (ny, nx))
    pd = np.zeros((ny, nx))
    pl = np.zeros((ny, nx))
    pd[10:14, 10:14] = 0.5      # existing precipitate: no nucleation there
    pl[20:22, 5:9] = 1e-3       # p(phi) ~ 1e-8 > 1e-10: also excluded
    dx = 2.5e-7
    sig = (0.0105, 0.0108)
    dV = dx * nx * dx * ny * dx
    Vatom = 0.25 * 0.352e-9 ** 3
    rate = _rate(oracle, 0.225, 0.215, sig[0], Vatom, dV / Vatom, dV)
    dt_n = 0.7 / rate
    return [xCr, xNb, pd, pl], dict(sigma_del=sig[0], sigma_lav=sig[1], unit_a=0.352e-9,
                                    dx=dx, dz=dx, dt=dt_n)


def test_nucleation_decisions_and_embed(oracle):
    """Synthetic fields where P is in (0, 1): the event flags equal the
    decision rule applied cell by cell, and each event embeds the phase's
    equilibrium composition at its centre (delta before Laves)."""
    fields, a = synthetic_nucleation_case(oracle)
    xCr, xNb, pd, pl = fields
    ny, nx = xCr.shape
    f, nev, flags = oracle.nucleation(fields, a["sigma_del"], a["sigma_lav"], a["unit_a"],
                                      a["dx"] * nx, a["dx"] * ny, a["dz"], a["dt"], seed=7,
                                      call=3)
    assert 0 < nev < 2 * (nx - 2) * (ny - 2)
    dV = a["dx"] * nx * a["dx"] * ny * a["dz"]
    Vatom = 0.25 * a["unit_a"] ** 3
    seen = 0
    probs = []
    for y in range(1, ny - 1):
        for x in range(1, nx - 1):
            idx = nx * y + x
            want = 0
            if oracle.p(pd[y, x]) + oracle.p(pl[y, x]) < 1e-10:
                u = oracle.nuc_uniforms(7, 3, 0, idx)
                dG = oracle.g(0, xCr[y, x], xNb[y, x])
                for ph in range(2):
                    R, P = oracle.nucleation_probability(
                        xCr[y, x], xNb[y, x], dG, oracle.D(0, 0, xCr[y, x], xNb[y, x]),
                        oracle.D(1, 1, xCr[y, x], xNb[y, x]),
                        (a["sigma_del"], a["sigma_lav"])[ph], Vatom, dV / Vatom, dV, a["dt"])
                    probs.append(P)
                    if R > 0 and P - u[ph] > 0:
                        want |= 1 << ph
            assert flags[y, x] == want
            if want:
                seen += 1
                last = 2 if want & 2 else 1
                assert f[0][y, x] == oracle.xe(last, 0)
                assert f[1][y, x] == oracle.xe(last, 1)
                if want & 1:
                    assert f[2][y, x] == 1.0
                if want & 2:
                    assert f[3][y, x] == 1.0
            else:
                for k in range(4):
                    assert f[k][y, x] == fields[k][y, x]
    probs = np.array(probs)
    assert ((probs > 0.05) & (probs < 0.95)).mean() > 0.5
    assert seen > 100
    assert flags[10:14, 10:14].sum() == 0 and flags[20:22, 5:9].sum() == 0


# ------------------------------------------------------------ energy functional
# orc_energy restates src/alloy625.cpp:752-919 (gibbs, maskedgradient,
# summarize_energy).  Pinned by closed forms in which each term is isolated.

E_DX = 2.5e-10
E_KAPPA = (4.4318181818181824e-10, 4.4318181818181824e-10)
E_OMEGA = (6.8637292464878e8, 6.8637292464878e8)
E_OMEGA3 = 1.07e11


def _uniform(Nx, Ny, xCr, xNb, pd, pl):
    return [np.full((Ny + 2, Nx + 2), v, dtype=np.float64) for v in (xCr, xNb, pd, pl)]


def test_energy_pure_phases_are_the_paraboloids(oracle):
    """phi = 0: pGam = 1, x_gam = x, so the density is g_gam(x); phi_del = 1
    (phi_lav = 0): pDel = 1, x_del = x and the well vanishes, so it is
    g_del(x).  Ties the combination to the golden-pinned paraboloids."""
    Nx, Ny = 7, 5
    dV = E_DX * E_DX
    for xCr, xNb in ((0.31, 0.021), (0.45, 0.11), (0.05, 0.2)):
        f = _uniform(Nx, Ny, xCr, xNb, 0.0, 0.0)
        tot, chem, grad = oracle.energy(f, 0, E_DX, E_DX, E_KAPPA, E_OMEGA, E_OMEGA3)
        assert grad == 0.0
        want = Nx * Ny * dV * oracle.g(0, xCr, xNb)
        assert chem == pytest.approx(want, rel=1e-13)
        f = _uniform(Nx, Ny, xCr, xNb, 1.0, 0.0)
        tot, chem, grad = oracle.energy(f, 1, E_DX, E_DX, E_KAPPA, E_OMEGA, E_OMEGA3)
        assert chem == pytest.approx(Nx * Ny * dV * oracle.g(1, xCr, xNb), rel=1e-12)
        f = _uniform(Nx, Ny, xCr, xNb, 0.0, 1.0)
        tot, chem, grad = oracle.energy(f, 1, E_DX, E_DX, E_KAPPA, E_OMEGA, E_OMEGA3)
        assert chem == pytest.approx(Nx * Ny * dV * oracle.g(2, xCr, xNb), rel=1e-12)


def test_energy_double_well_and_trijunction(oracle):
    """At x = sum_a p_a xe_a the KKS potential is 0 and every g_a(x_a) = 0, so
    only the Landau terms remain: omega/16 at phi = 1/2, 2 Omega at
    phi_del = phi_lav = 1."""
    Nx, Ny = 6, 4
    dV = E_DX * E_DX
    xe = lambda a: np.array([oracle.xe(a, 0), oracle.xe(a, 1)])  # noqa: E731
    x = 0.5 * xe(0) + 0.5 * xe(1)
    f = _uniform(Nx, Ny, x[0], x[1], 0.5, 0.0)
    tot, chem, grad = oracle.energy(f, 1, E_DX, E_DX, E_KAPPA, E_OMEGA, E_OMEGA3)
    assert chem == pytest.approx(Nx * Ny * dV * E_OMEGA[0] / 16, rel=1e-9)
    x = xe(1) + xe(2) - xe(0)   # pGam = -1
    f = _uniform(Nx, Ny, x[0], x[1], 1.0, 1.0)
    tot, chem, grad = oracle.energy(f, 0, E_DX, E_DX, E_KAPPA, E_OMEGA, E_OMEGA3)
    assert chem == pytest.approx(Nx * Ny * dV * 2 * E_OMEGA3, rel=1e-9)
    assert tot == chem + grad


def test_energy_gradient_closed_forms(oracle):
    """Central differences: a linear ramp in phi_del with no-flux edges (a/dx
    inside, a/(2 dx) at the two clamped columns) and a periodic sinusoid in
    phi_lav along y (exact: A sin(theta)/dy cos(theta y))."""
    Nx, Ny = 20, 3
    a = 0.04
    dx, dy = E_DX, 1.5 * E_DX
    pd = np.zeros((Ny + 2, Nx + 2))
    pd[:, 1:-1] = a * np.arange(1, Nx + 1)
    pG = 1 - np.vectorize(oracle.p)(pd)
    pD = np.vectorize(oracle.p)(pd)
    xCr = pG * oracle.xe(0, 0) + pD * oracle.xe(1, 0)
    xNb = pG * oracle.xe(0, 1) + pD * oracle.xe(1, 1)
    f = [xCr, xNb, pd, np.zeros_like(pd)]
    tot, chem, grad = oracle.energy(f, 0, dx, dy, E_KAPPA, E_OMEGA, E_OMEGA3)
    dV = dx * dy
    want_grad = E_KAPPA[0] * dV * Ny * ((Nx - 2) * (a / dx) ** 2 + 2 * (a / (2 * dx)) ** 2)
    assert grad == pytest.approx(want_grad, rel=1e-12)
    phi = pd[1:-1, 1:-1]
    want_chem = dV * np.sum(E_OMEGA[0] * phi ** 2 * (1 - phi) ** 2)
    assert chem == pytest.approx(want_chem, rel=1e-7, abs=1e-30)

    Nx, Ny = 3, 16
    A = 0.3
    th = 2 * np.pi / Ny
    pl = np.zeros((Ny + 2, Nx + 2))
    yy = np.arange(Ny)[:, None] * np.ones((1, Nx))
    pl[1:-1, 1:-1] = 0.5 + A * np.sin(th * yy)
    f = [np.full_like(pl, 0.3), np.full_like(pl, 0.02), np.zeros_like(pl), pl]
    _, _, grad = oracle.energy(f, 1, dx, dy, E_KAPPA, E_OMEGA, E_OMEGA3)
    want = E_KAPPA[1] * dx * dy * Nx * (A * np.sin(th) / dy) ** 2 * Ny / 2
    assert grad == pytest.approx(want, rel=1e-12)


# ------------------------------------------------ variable-mobility divergence


def _mob_case(Nx, Ny, seed):
    rng = np.random.default_rng(seed)
    shape = (Ny + 2, Nx + 2)
    conc = [rng.uniform(0.0, 0.5, shape) for _ in range(6)]
    mob_cr = [rng.uniform(0.5, 2.0, shape) for _ in range(8)]
    mob_nb = [rng.uniform(0.5, 2.0, shape) for _ in range(8)]
    phis = [rng.uniform(0.0, 1.0, shape) for _ in range(2)]
    return conc, mob_cr, mob_nb, phis


def _fill(a, bc):
    """Ghost ring: no-flux copy of the adjacent interior cell (corners from the
    diagonal one) or periodic wrap."""
    a = a.copy()
    if bc == 0:
        a[0, :], a[-1, :] = a[1, :], a[-2, :]
        a[:, 0], a[:, -1] = a[:, 1], a[:, -2]
    else:
        a[0, :], a[-1, :] = a[-2, :], a[1, :]
        a[:, 0], a[:, -1] = a[:, -2], a[:, 1]
    return a


def test_mobility_divergence_constant_mobility_is_laplacian(oracle):
    """Constant mobility M on one composition term (all others zero): the
    face average (2M)(c_hi - c) - (2M)(c - c_lo) over 2 h^2 is M times the
    five-point Laplacian, at every interior cell (dx != dy)."""
    Nx, Ny, dx, dy = 13, 9, 0.5, 0.25
    conc, _, _, phis = _mob_case(Nx, Ny, 1)
    z = np.zeros_like(conc[0])
    for t in range(6):
        for comp in range(2):
            mc = [z] * 8
            mn = [z] * 8
            M = 1.75
            (mc if comp == 0 else mn)[t] = np.full_like(z, M)
            out = oracle.mobility_divergence(conc, mc, mn, phis[0], phis[1], dx, dy)
            c = conc[t]
            lap = ((c[1:-1, 2:] - 2 * c[1:-1, 1:-1] + c[1:-1, :-2]) / dx ** 2 +
                   (c[2:, 1:-1] - 2 * c[1:-1, 1:-1] + c[:-2, 1:-1]) / dy ** 2)
            got = out[comp][1:-1, 1:-1]
            assert np.allclose(got, M * lap, rtol=1e-13, atol=1e-13 * np.abs(M * lap).max())
            assert not np.any(out[1 - comp])           # the other component sees nothing


def test_mobility_divergence_linear_mobility_quadratic_field(oracle):
    """M = a + b x (linear) and c = x^2: the face-averaged difference equals
    the continuous d/dx (M dc/dx) = 2a + 4 b x exactly (up to rounding); the
    same in y.  A transposed or mis-paired face average fails this."""
    Nx, Ny, h = 20, 16, 0.125
    X = (np.arange(Nx + 2) * h)[None, :] * np.ones((Ny + 2, 1))
    Y = (np.arange(Ny + 2) * h)[:, None] * np.ones((1, Nx + 2))
    a, b = 1.5, 0.75
    z = np.zeros_like(X)
    for axis, coord in ((0, X), (1, Y)):
        conc = [coord ** 2] + [z] * 5
        mc = [a + b * coord] + [z] * 7
        out = oracle.mobility_divergence(conc, mc, [z] * 8, z, z, h, h)[0]
        want = 2 * a + 4 * b * coord
        assert np.allclose(out[1:-1, 1:-1], want[1:-1, 1:-1], rtol=1e-12, atol=0)


def test_mobility_divergence_phase_terms_are_x_only(oracle):
    """Terms 7 and 8 (the phase-field terms) difference in x only, as the
    reference's kernels do: a phase field varying in y alone contributes
    nothing; one varying in x alone gives the 1-D face average."""
    Nx, Ny, dx, dy = 11, 12, 0.5, 0.5
    z = np.zeros((Ny + 2, Nx + 2))
    Yv = (np.arange(Ny + 2, dtype=float)[:, None] ** 2) * np.ones((1, Nx + 2))
    Xv = (np.arange(Nx + 2, dtype=float)[None, :] ** 2) * np.ones((Ny + 2, 1))
    M = np.full_like(z, 2.0)
    for slot, which in ((6, 0), (7, 1)):
        mc = [z] * 8
        mc[slot] = M
        phis = [z, z]
        phis[which] = Yv
        out = oracle.mobility_divergence([z] * 6, mc, [z] * 8, phis[0], phis[1], dx, dy)[0]
        assert not np.any(out[1:-1, 1:-1])
        phis[which] = Xv
        out = oracle.mobility_divergence([z] * 6, mc, [z] * 8, phis[0], phis[1], dx, dy)[0]
        # (4 (c_hi - c) - 4 (c - c_lo)) / (2 dx^2) with c = x^2: 2 * 2 / dx^2 * 2 / 2
        assert np.allclose(out[1:-1, 1:-1], 2.0 * 2.0 / dx ** 2)


@pytest.mark.parametrize("bc", [0, 1])
def test_mobility_divergence_conserves_mass(oracle, bc):
    """With the ghost ring filled by the no-flux copy or the periodic wrap
    (fields and mobilities alike), the fluxes through the domain edge vanish
    or cancel: the interior sum of the divergence is zero to rounding."""
    Nx, Ny = 17, 10
    conc, mc, mn, phis = _mob_case(Nx, Ny, 5)
    conc = [_fill(a, bc) for a in conc]
    mc = [_fill(a, bc) for a in mc]
    mn = [_fill(a, bc) for a in mn]
    phis = [_fill(a, bc) for a in phis]
    dx, dy = 0.5, 0.4
    out = oracle.mobility_divergence(conc, mc, mn, phis[0], phis[1], dx, dy)
    for o in out:
        interior = o[1:-1, 1:-1]
        assert abs(interior.sum()) <= 1e-12 * np.abs(interior).sum()
        assert np.abs(interior).max() > 1.0        # not trivially zero


def test_five_point_mask_anisotropic(oracle):
    """five_point_Laplacian_stencil (src/mesh.cpp:187-196) with dx != dy,
    applied as the reference's mask convolution: quadratics in x and in y give
    the exact Laplacian 2 each, constants give 0 (a swapped dx / dy or a wrong
    centre fails this)."""
    dx, dy = 0.5, 0.25
    m = np.array(oracle.five_point_mask(dx, dy)).reshape(3, 3)
    X = np.arange(9)[None, :] * dx * np.ones((7, 1))
    Y = np.arange(7)[:, None] * dy * np.ones((1, 9))

    def conv(f):
        out = np.zeros((5, 7))
        for j in range(3):
            for i in range(3):
                out += m[j, i] * f[j:j + 5, i:i + 7]
        return out

    assert np.allclose(conv(X ** 2), 2.0, rtol=1e-12)
    assert np.allclose(conv(Y ** 2), 2.0, rtol=1e-12)
    assert np.allclose(conv(3.0 * X ** 2 - Y ** 2 + 0.7), 4.0, rtol=1e-12)
    assert np.abs(conv(np.full((7, 9), 0.3))).max() <= 1e-12 * np.abs(m).max()


@pytest.mark.parametrize("n", [1, 2, 3, 6])
def test_run_equals_repeated_steps(oracle, n):
    """orc_run's buffer ping-pong leaves the result of n steps in the first
    buffer set, odd or even n: the same bits as n single steps."""
    P = _params(oracle, 1e-7)
    f = [np.ascontiguousarray(a) for a in
         (np.full((9, 11), 0.3), np.full((9, 11), 0.02), np.zeros((9, 11)), np.zeros((9, 11)))]
    rng = np.random.default_rng(n)
    f[0] = f[0] + 1e-3 * rng.standard_normal(f[0].shape)
    f[2][3:6, 4:8] = 0.8
    cur = [a.copy() for a in f]
    for _ in range(n):
        cur = oracle.step(cur, P, 1)
    got = oracle.run(f, P, 1, n)
    for a, b in zip(got, cur):
        assert np.array_equal(a, b)
