"""Model parameters, time-step rules and seeded initial conditions.

Parameter values follow src/parameters.h:14-48 (interface width, LinStab,
lattice constant, sigma, Omega, kappa/omega formulas) and
src/alloy625.cpp:80-121 (diffusivity matrix at (0.30, 0.02), Allen-Cahn
mobility, diffusion time-step limit).  Thermodynamic evaluations go through
the library's host-side functions (libpfpa.so, pf_host_*).
"""
import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import lib

# src/parameters.h
IFCE_WIDTH = 2.5e-9          # interface thickness 2*lambda (m), :19
LINSTAB = 0.02               # von Neumann threshold, :21
LATTICE_CONST = 0.352e-9     # FCC Ni (m), :27
SIGMA = (0.13, 0.13)         # s_delta(), s_laves() (J/m^2), :37 via thermo/parabola625.c:299-316
OMEGA_3PHASE = 1.07e11       # three-phase coefficient (J/m^3), :38
WIDTH_FACTOR = 2.2           # :43
MOBSTAB = 1.0                # :22
RT = 9504.6840941999999      # thermo/parabola625.c:47-54
VM = 1.0000000000000001e-5   # thermo/parabola625.c:56-63

# composition ranges of the benchmark initial conditions
MATRIX_CR = (0.2794, 0.3288)
MATRIX_NB = (0.0202, 0.0269)
ENRICHED = (0.272, 0.169)    # (x_Cr, x_Nb) of the enriched band


def kappa():
    """Gradient energy coefficients (src/parameters.h:44-46)."""
    return [3.0 * s * IFCE_WIDTH / WIDTH_FACTOR for s in SIGMA]


def omega():
    """Double-well heights (src/parameters.h:47-49)."""
    k = kappa()
    return [18.0 * SIGMA[i] * SIGMA[i] / k[i] for i in range(2)]


def D(i, j, xCr, xNb):
    return lib().pf_host_D(i, j, xCr, xNb)


def set_diffusion_matrix(xCr0=0.30, xNb0=0.02):
    """(D_Cr, D_Nb, Lmob) as src/alloy625.cpp:80-109 computes them."""
    DCr = [D(0, 0, xCr0, xNb0), D(0, 1, xCr0, xNb0)]
    DNb = [D(1, 0, xCr0, xNb0), D(1, 1, xCr0, xNb0)]
    Dnorm = math.sqrt(DCr[0] * DCr[0] + DNb[1] * DNb[1])
    L = MOBSTAB * Dnorm / (IFCE_WIDTH * IFCE_WIDTH * RT / VM)
    return DCr, DNb, [L, L]


def five_point_mask(dx, dy):
    """src/mesh.cpp:187-196 as a row-major 3x3 list."""
    m = [0.0] * 9
    m[1] = 1.0 / (dy * dy)
    m[3] = 1.0 / (dx * dx)
    m[4] = -2.0 * (dx * dx + dy * dy) / (dx * dx * dy * dy)
    m[5] = 1.0 / (dx * dx)
    m[7] = 1.0 / (dy * dy)
    return m


def nine_point_mask(dx, dy):
    """src/mesh.cpp:198-213 as a row-major 3x3 list (the same expressions, so
    the same doubles as the C++ set_mask(dx, dy, 93, ...))."""
    c = 1.0 / (6.0 * dx * dy)
    return [c, 4.0 / (6.0 * dy * dy), c,
            4.0 / (6.0 * dx * dx), -10.0 * (dx * dx + dy * dy) / (6.0 * dx * dx * dy * dy),
            4.0 / (6.0 * dx * dx),
            c, 4.0 / (6.0 * dy * dy), c]


def symmetric_five_point(mask):
    """Whether a row-major 3x3 mask is a symmetric five-point one (zero corners,
    m[0][1] = m[2][1] != 0 and finite, m[1][0] = m[1][2]): the masks the two- and
    three-step kernels and the batched form take (every five-point mask the
    reference builds, src/mesh.cpp:187-196, is one)."""
    m = [float(v) for v in mask]
    return (m[0] == 0.0 and m[2] == 0.0 and m[6] == 0.0 and m[8] == 0.0 and m[1] == m[7]
            and m[3] == m[5] and m[1] != 0.0 and math.isfinite(m[1]))


def max_effective_diffusivity(DCr, DNb):
    a = (ctypes.c_double * 2)(*DCr)
    b = (ctypes.c_double * 2)(*DNb)
    return lib().pf_host_max_effective_diffusivity(a, b)


def dt_reference(dx, DCr, DNb, Lmob, dim=2):
    """The reference's rule (src/main.cpp:259-268 with src/alloy625.cpp:111-121):
    LinStab * min(transformation, diffusion) limit, floored to a round number."""
    k = kappa()
    dt_T = (dx * dx) / (2.0 * dim * Lmob[0] * k[0])
    dt_D = (dx * dx) / (4.0 * max(DCr[0], DNb[1]))
    round_dt = 0.0
    roundoff = 4e6
    while round_dt < 1e-20:
        roundoff *= 10
        round_dt = math.floor(roundoff * LINSTAB * min(dt_T, dt_D)) / roundoff
    return round_dt


def dt_stable(dx, DCr, DNb, factor=0.1):
    """factor * dx^2 / (4 D_max) with D_max the largest effective diffusivity
    over the three phases (inside a precipitate the composition diffuses with
    D * H_gamma^-1 * H_phase, up to ~59x D_NbNb for delta)."""
    return factor * dx * dx / (4.0 * max_effective_diffusivity(DCr, DNb))


def nuc_interval(dt):
    """src/main.cpp:273."""
    return int(0.0001 / dt)


@dataclass
class Setup:
    """Everything one simulation needs besides its fields."""
    N: tuple            # interior (Nx, Ny)
    dx: float
    dt: float
    bc: int
    DCr: list = field(default_factory=list)
    DNb: list = field(default_factory=list)
    Lmob: list = field(default_factory=list)
    kappa: list = field(default_factory=list)
    omega: list = field(default_factory=list)
    alpha: float = OMEGA_3PHASE
    mask: list = field(default_factory=list)

    @property
    def nx(self):
        return self.N[0] + 2

    @property
    def ny(self):
        return self.N[1] + 2


def make_setup(Nx, Ny, dx=0.25e-9, bc=1, dt_rule="stable"):
    DCr, DNb, L = set_diffusion_matrix(0.30, 0.02)
    if dt_rule == "stable":
        dt = dt_stable(dx, DCr, DNb)
    elif dt_rule == "reference":
        dt = dt_reference(dx, DCr, DNb, L)
    else:
        dt = float(dt_rule)
    return Setup(N=(Nx, Ny), dx=dx, dt=dt, bc=bc, DCr=DCr, DNb=DNb, Lmob=L, kappa=kappa(),
                 omega=omega(), alpha=OMEGA_3PHASE, mask=five_point_mask(dx, dx))


# ---------------------------------------------------------------- initial conditions

def fill_ghosts(a, bc):
    """Ghost ring of a padded (ny, nx) array: no-flux copy or periodic wrap."""
    if bc == 0:
        a[:, 0] = a[:, 1]
        a[:, -1] = a[:, -2]
        a[0, :] = a[1, :]
        a[-1, :] = a[-2, :]
    else:
        a[:, 0] = a[:, -2]
        a[:, -1] = a[:, 1]
        a[0, :] = a[-2, :]
        a[-1, :] = a[1, :]
    return a


def _xe(phase, comp):
    return lib().pf_host_xe(phase, comp)


def _place_discs(rng, Nx, Ny, R, count, min_gap, existing=()):
    centres = list(existing)
    placed = []
    tries = 0
    while len(placed) < count:
        tries += 1
        if tries > 100000:
            raise RuntimeError("could not place precipitates")
        c = (rng.uniform(0, Nx), rng.uniform(0, Ny))
        ok = True
        for e in centres:
            ddx = abs(c[0] - e[0])
            ddy = abs(c[1] - e[1])
            ddx = min(ddx, Nx - ddx)
            ddy = min(ddy, Ny - ddy)
            if math.hypot(ddx, ddy) < min_gap:
                ok = False
                break
        if ok:
            centres.append(c)
            placed.append(c)
    return placed, centres


def init_matrix_with_particles(Nx, Ny, dx, bc, seed=1, xCr0=0.3041, xNb0=0.0235, noise=1e-3,
                               n_del=2, n_lav=2, radius=5e-9):
    """256^2 plain case: uniform matrix plus seeded Gaussian noise, and circular
    delta and Laves particles at seeded positions with phi = 1 and the
    equilibrium paraboloid compositions inside."""
    rng = np.random.default_rng(seed)
    nx, ny = Nx + 2, Ny + 2
    xCr = np.zeros((ny, nx))
    xNb = np.zeros((ny, nx))
    pd = np.zeros((ny, nx))
    pl = np.zeros((ny, nx))
    xCr[1:-1, 1:-1] = xCr0 + noise * rng.standard_normal((Ny, Nx))
    xNb[1:-1, 1:-1] = xNb0 + noise * rng.standard_normal((Ny, Nx))
    _stamp_particles(rng, [xCr, xNb, pd, pl], dx, bc, n_del, n_lav, radius)
    for a in (xCr, xNb, pd, pl):
        fill_ghosts(a, bc)
    return [xCr, xNb, pd, pl]


def add_particles(fields, dx, bc, seed=1, n_del=2, n_lav=2, radius=5e-9):
    """Stamp seeded delta and Laves discs (phi = 1 and the equilibrium
    compositions inside) onto padded fields (x_Cr, x_Nb, phi_del, phi_lav)
    in place, and refresh their ghost rings."""
    rng = np.random.default_rng(seed)
    _stamp_particles(rng, fields, dx, bc, n_del, n_lav, radius)
    for a in fields:
        fill_ghosts(a, bc)
    return fields


def _stamp_particles(rng, fields, dx, bc, n_del, n_lav, radius):
    xCr, xNb, pd, pl = fields
    Ny, Nx = xCr.shape[0] - 2, xCr.shape[1] - 2
    R = radius / dx
    dels, cs = _place_discs(rng, Nx, Ny, R, n_del, 2 * R + 8)
    lavs, cs = _place_discs(rng, Nx, Ny, R, n_lav, 2 * R + 8, cs)

    def box(c, n):
        # the interior indices within R + 1 of centre c (wrapped when periodic)
        i = np.arange(int(math.floor(c - R)) - 1, int(math.ceil(c + R)) + 2)
        return np.unique(i % n) if bc == 1 else i[(i >= 0) & (i < n)]

    for centres, phi, ph in ((dels, pd, 1), (lavs, pl, 2)):
        xe = (_xe(ph, 0), _xe(ph, 1))
        for (cx, cy) in centres:
            xs, ys = box(cx, Nx), box(cy, Ny)
            # per cell the same arithmetic as over the whole grid, on the box only
            ddx = np.abs(xs[None, :] + 0.5 - cx)
            ddy = np.abs(ys[:, None] + 0.5 - cy)
            if bc == 1:
                ddx = np.minimum(ddx, Nx - ddx)
                ddy = np.minimum(ddy, Ny - ddy)
            inside = np.hypot(ddx, ddy) <= R
            rows, cols = np.broadcast_arrays(ys[:, None] + 1, xs[None, :] + 1)
            rows, cols = rows[inside], cols[inside]
            phi[rows, cols] = 1.0
            xCr[rows, cols] = xe[0]
            xNb[rows, cols] = xe[1]


def init_enriched_band(Nx, Ny, bc, seed=42, dtype=np.float64):
    """4096^2 / 8192^2 cases: per-cell seeded uniform matrix composition within
    the matrix ranges, a central enriched strip 1/8 of the domain width, no
    precipitates."""
    rng = np.random.default_rng(seed)
    nx, ny = Nx + 2, Ny + 2
    xCr = np.empty((ny, nx), dtype=dtype)
    xNb = np.empty((ny, nx), dtype=dtype)
    xCr[1:-1, 1:-1] = rng.uniform(MATRIX_CR[0], MATRIX_CR[1], (Ny, Nx))
    xNb[1:-1, 1:-1] = rng.uniform(MATRIX_NB[0], MATRIX_NB[1], (Ny, Nx))
    w = max(1, Nx // 8)
    x0 = (Nx - w) // 2
    xCr[1:-1, 1 + x0:1 + x0 + w] = ENRICHED[0]
    xNb[1:-1, 1 + x0:1 + x0 + w] = ENRICHED[1]
    pd = np.zeros((ny, nx), dtype=dtype)
    pl = np.zeros((ny, nx), dtype=dtype)
    for a in (xCr, xNb):
        fill_ghosts(a, bc)
    return [xCr, xNb, pd, pl]
