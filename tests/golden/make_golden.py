"""Generate golden vectors from the reference's own compiled code.

Run in a container where the reference is mounted (it reads nothing else):

    make -C oracle            # compiles /root/reference/thermo/parabola625.c
    python tests/golden/make_golden.py

Writes tests/golden/thermo_golden.npz (pointwise thermodynamics) and
tests/golden/step_golden.npz (a literal replay of the reference's per-step
kernel sequence, src/main.cpp:289-305 -> src/discretization.cu, evaluated with
the reference's compiled pointwise functions).  Both are data fixtures: inputs
and the reference's outputs.  The reference itself never travels.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

SEED = 20261202


def thermo_vectors(R):
    rng = np.random.default_rng(SEED)
    out = {}
    # constants
    for name in O.REF_FUNCS_0:
        out["const_" + name] = np.array([getattr(R, name)()])

    # p, pPrime
    xs = np.concatenate([np.linspace(-0.2, 1.2, 57), rng.uniform(-0.1, 1.1, 200),
                         [0.0, 0.5, 1.0]])
    out["p_in"] = xs
    out["p_out"] = np.array([R.p(x) for x in xs])
    out["pPrime_out"] = np.array([R.pPrime(x) for x in xs])

    # compositions: random over the region the simulations visit plus named points
    special = [
        (0.30, 0.02),                      # set_diffusion_matrix point, src/main.cpp:254
        (0.3041, 0.0235),                  # matrix composition of the 256^2 case
        (0.272, 0.169),                    # enriched band
        (0.2794, 0.0202), (0.3288, 0.0269),  # matrix range corners
        (0.5175, 0.0220),                  # thermo/check-nucleation.cpp:162-165
    ]
    for a in range(3):
        ph = ("gam", "del", "lav")[a]
        special.append((getattr(R, "xe_%s_Cr" % ph)(), getattr(R, "xe_%s_Nb" % ph)()))
    comp = np.array(special + list(zip(rng.uniform(0.005, 0.6, 400),
                                       rng.uniform(0.005, 0.3, 400))))
    out["comp_in"] = comp
    for name in O.REF_FUNCS_2:
        out[name] = np.array([getattr(R, name)(c[0], c[1]) for c in comp])

    # fictitious compositions: physical phase fractions (pDel + pLav <= 1 with
    # pGam = 1 - pDel - pLav) from phi in [0, 1], pure phases, and a few
    # independent-pGam points
    rows = []
    for c in comp[:9]:
        for (pd, pg, pl) in [(0, 1, 0), (1, 0, 0), (0, 0, 1), (0.5, 0.5, 0), (0, 0.5, 0.5),
                             (0.5, 0, 0.5), (0.25, 0.5, 0.25)]:
            rows.append((c[0], c[1], pd, pg, pl))
    for _ in range(600):
        xCr = rng.uniform(0.005, 0.6)
        xNb = rng.uniform(0.005, 0.3)
        pd = R.p(rng.uniform(0, 1))
        pl = R.p(rng.uniform(0, 1)) if rng.uniform() < 0.5 else 0.0
        if pd + pl > 1.0:
            pl = 1.0 - pd
        rows.append((xCr, xNb, pd, 1.0 - pd - pl, pl))
    for _ in range(50):
        rows.append((rng.uniform(0.05, 0.5), rng.uniform(0.01, 0.25), rng.uniform(0, 1),
                     rng.uniform(0, 1), rng.uniform(0, 1)))
    fin = np.array(rows)
    out["fict_in"] = fin
    for name in O.REF_FUNCS_5:
        out[name] = np.array([getattr(R, name)(*r) for r in fin])
    return out


# --------------------------------------------------------------------------
# Literal replay of the reference's per-step kernel sequence
# --------------------------------------------------------------------------

def boundary(f):
    """boundary_kernel, nm = 3 (src/discretization.cu:360-394): left, right,
    bottom, top ghost := adjacent interior (corners: last writer wins)."""
    ny, nx = f.shape
    f[:, 0] = f[:, 1]
    f[:, nx - 1] = f[:, nx - 2]
    f[0, :] = f[1, :]
    f[ny - 1, :] = f[ny - 2, :]


def boundary_periodic(f):
    """The same ghost ring filled by the periodic wrap that the 256^2 plain case
    asks for (the reference itself has only the no-flux copy): ghost columns
    and rows := the interior column / row at the opposite side, corners from
    the diagonally opposite interior corner."""
    ny, nx = f.shape
    f[:, 0] = f[:, nx - 2]
    f[:, nx - 1] = f[:, 1]
    f[0, :] = f[ny - 2, :]
    f[ny - 1, :] = f[1, :]


def convolution(src, mask):
    """convolution_kernel interior outputs (src/discretization.cu:550-603),
    accumulated in mask order; ghost outputs are left as garbage (NaN)."""
    ny, nx = src.shape
    out = np.full_like(src, np.nan)
    for y in range(1, ny - 1):
        for x in range(1, nx - 1):
            v = 0.0
            for j in range(3):
                for i in range(3):
                    v += mask[j * 3 + i] * src[y + j - 1, x + i - 1]
            out[y, x] = v
    return out


def replay_step(R, old, P, boundary=boundary):
    """One pass of src/main.cpp:289-305 with the reference's kernels;
    `boundary` fills the ghost rings (the reference's copy, or the periodic
    wrap)."""
    xCr, xNb, pd, pl = [a.copy() for a in old]
    ny, nx = xCr.shape
    for a in (xCr, xNb, pd, pl):
        boundary(a)                                   # device_boundaries
    fict = np.zeros((6, ny, nx))
    names = O.REF_FUNCS_5
    for y in range(ny):                                # device_fictitious
        for x in range(nx):
            pD = R.p(pd[y, x])
            pL = R.p(pl[y, x])
            pG = 1.0 - pD - pL
            for k, n in enumerate(names):
                fict[k, y, x] = getattr(R, n)(xCr[y, x], xNb[y, x], pD, pG, pL)
    for k in range(6):
        boundary(fict[k])                              # fictitious_boundaries
    lap = [convolution(pd, P["mask"]), convolution(pl, P["mask"]),
           convolution(fict[0], P["mask"]), convolution(fict[1], P["mask"])]
    for a in lap:
        boundary(a)                                    # device_laplacian_boundaries
    dt = P["dt"]
    nCr = xCr + dt * (P["DCr"][0] * lap[2] + P["DCr"][1] * lap[3])   # composition_kernel
    nNb = xNb + dt * (P["DNb"][0] * lap[2] + P["DNb"][1] * lap[3])
    npd = np.empty_like(pd)
    npl = np.empty_like(pl)
    for y in range(ny):                                # allen_cahn_kernel
        for x in range(nx):
            pD = R.p(pd[y, x])
            pL = R.p(pl[y, x])
            pG = 1.0 - pD - pL
            gCr = R.fict_gam_Cr(xCr[y, x], xNb[y, x], pD, pG, pL)
            gNb = R.fict_gam_Nb(xCr[y, x], xNb[y, x], pD, pG, pL)
            gam_nrg = R.g_gam(gCr, gNb)
            muCr = R.dg_gam_dxCr(gCr, gNb)
            muNb = R.dg_gam_dxNb(gCr, gNb)
            for ph, (fa, fb, gfun, phi, other, lp, outarr) in enumerate([
                    (R.fict_del_Cr, R.fict_del_Nb, R.g_del, pd, pl, lap[0], npd),
                    (R.fict_lav_Cr, R.fict_lav_Nb, R.g_lav, pl, pd, lap[1], npl)]):
                aCr = fa(xCr[y, x], xNb[y, x], pD, pG, pL)
                aNb = fb(xCr[y, x], xNb[y, x], pD, pG, pL)
                a_nrg = gfun(aCr, aNb)
                gp = gam_nrg - a_nrg - muCr * (gCr - aCr) - muNb * (gNb - aNb)
                f = phi[y, x]
                o = other[y, x]
                dF = (-R.pPrime(f) * gp
                      + 2.0 * P["omega"][ph] * f * (f - 1.0) * (2.0 * f - 1.0)
                      + 2.0 * P["alpha"] * f * o * o
                      - P["kappa"][ph] * lp[y, x])
                outarr[y, x] = f - dt * P["Lmob"][ph] * dF
    return [nCr, nNb, npd, npl]


def step_vectors(R):
    """Small seeded grid with one delta and one Laves particle, no-flux, the
    reference's own parameter values (src/parameters.h, meshres = 0.25 nm)."""
    rng = np.random.default_rng(SEED + 1)
    Nx, Ny = 14, 11
    nx, ny = Nx + 2, Ny + 2
    dx = 0.25e-9
    xCr = 0.3041 + 1e-3 * rng.standard_normal((ny, nx))
    xNb = 0.0235 + 1e-3 * rng.standard_normal((ny, nx))
    pd = np.zeros((ny, nx))
    pl = np.zeros((ny, nx))
    yy, xx = np.mgrid[0:ny, 0:nx]
    # soft-edged particles so that interfaces carry fractional phi
    rd = np.hypot(xx - 4.3, yy - 4.1)
    rl = np.hypot(xx - 10.6, yy - 8.2)
    pd[:] = 0.5 - 0.5 * np.tanh((rd - 2.5) / 1.2)
    pl[:] = 0.5 - 0.5 * np.tanh((rl - 2.2) / 1.2)
    xCr = np.where(pd > 0.5, R.xe_del_Cr(), np.where(pl > 0.5, R.xe_lav_Cr(), xCr))
    xNb = np.where(pd > 0.5, R.xe_del_Nb(), np.where(pl > 0.5, R.xe_lav_Nb(), xNb))

    sigma = R.s_delta()
    ifce_width = 2.5e-9
    kappa = [3.0 * sigma * ifce_width / 2.2] * 2
    omega = [18.0 * sigma * sigma / kappa[0]] * 2
    DCr = [R.D_CrCr(0.30, 0.02), R.D_CrNb(0.30, 0.02)]
    DNb = [R.D_NbCr(0.30, 0.02), R.D_NbNb(0.30, 0.02)]
    Dnorm = np.sqrt(DCr[0] ** 2 + DNb[1] ** 2)
    L = Dnorm / (ifce_width ** 2 * R.RT() / R.Vm())
    mask = [0.0] * 9
    mask[1] = 1.0 / dx ** 2
    mask[3] = 1.0 / dx ** 2
    mask[4] = -2.0 * (dx * dx + dx * dx) / (dx * dx * dx * dx)
    mask[5] = 1.0 / dx ** 2
    mask[7] = 1.0 / dx ** 2
    # a time step large enough that a few steps move the fields visibly but
    # inside the delta-phase stability limit
    dt = 5.0e-7
    P = dict(dt=dt, DCr=DCr, DNb=DNb, kappa=kappa, omega=omega, Lmob=[L, L],
             alpha=1.07e11, mask=mask)
    fields = [xCr, xNb, pd, pl]
    hist = [np.array(fields)]
    nsteps = 4
    for _ in range(nsteps):
        fields = replay_step(R, fields, P)
        hist.append(np.array(fields))
    # the same start and kernels with periodic ghost rings
    fields = [a.copy() for a in hist[0]]
    hist_p = [np.array(fields)]
    for _ in range(nsteps):
        fields = replay_step(R, fields, P, boundary_periodic)
        hist_p.append(np.array(fields))
    out = {"fields": np.array(hist), "fields_periodic": np.array(hist_p),
           "dt": np.array([dt]), "DCr": np.array(DCr),
           "DNb": np.array(DNb), "kappa": np.array(kappa), "omega": np.array(omega),
           "Lmob": np.array([L, L]), "alpha": np.array([1.07e11]), "mask": np.array(mask),
           "dx": np.array([dx])}
    return out


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libparabola625_ref.so missing: run `make -C oracle` where "
                 "the reference is mounted")
    R = O.ref()
    tv = thermo_vectors(R)
    np.savez_compressed(os.path.join(HERE, "thermo_golden.npz"), **tv)
    sv = step_vectors(R)
    np.savez_compressed(os.path.join(HERE, "step_golden.npz"), **sv)
    print("wrote", len(tv), "thermo arrays and", sv["fields"].shape, "step history")


if __name__ == "__main__":
    main()
