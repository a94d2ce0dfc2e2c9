"""ctypes wrapper around the CPU oracle (oracle/liborc.so) and, where it was
built, the reference's compiled thermodynamics (oracle/_ref/libparabola625_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's
cpu_baseline leg, never by the product.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PF_ORACLE_LIB: another build of pf_oracle.c (the CPU baseline's -O3
# -march=native build, oracle/cpu_bench.py); the checker is liborc.so
LIB = os.environ.get("PF_ORACLE_LIB") or os.path.join(HERE, "liborc.so")
REF_LIB = os.path.join(HERE, "_ref", "libparabola625_ref.so")

_d = ctypes.c_double
_i = ctypes.c_int
_dp = ctypes.POINTER(ctypes.c_double)
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("dt", _d),
        ("DCr", _d * 2),
        ("DNb", _d * 2),
        ("kappa", _d * 2),
        ("omega", _d * 2),
        ("Lmob", _d * 2),
        ("alpha", _d),
        ("mask", _d * 9),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        for name in ("orc_p", "orc_pPrime"):
            getattr(L, name).restype = _d
            getattr(L, name).argtypes = [_d]
        for name in ("orc_kT", "orc_RT", "orc_Vm"):
            getattr(L, name).restype = _d
            getattr(L, name).argtypes = []
        L.orc_xe.restype = _d
        L.orc_xe.argtypes = [_i, _i]
        L.orc_d2g.restype = _d
        L.orc_d2g.argtypes = [_i, _i]
        L.orc_sigma.restype = _d
        L.orc_sigma.argtypes = [_i]
        L.orc_g.restype = _d
        L.orc_g.argtypes = [_i, _d, _d]
        L.orc_dg.restype = _d
        L.orc_dg.argtypes = [_i, _i, _d, _d]
        L.orc_fict.restype = None
        L.orc_fict.argtypes = [_d, _d, _d, _d, _d, _dp]
        L.orc_D.restype = _d
        L.orc_D.argtypes = [_i, _i, _d, _d]
        L.orc_nucleation_probability.restype = None
        L.orc_nucleation_probability.argtypes = [_d] * 10 + [_dp, _dp]
        L.orc_set_diffusion_matrix.restype = None
        L.orc_set_diffusion_matrix.argtypes = [_d, _d, _d, _dp, _dp, _dp]
        L.orc_five_point_mask.restype = None
        L.orc_five_point_mask.argtypes = [_d, _d, _dp]
        L.orc_step.restype = None
        L.orc_step.argtypes = [_i, _i, _i, ctypes.POINTER(OrcParams)] + [_dp] * 8
        L.orc_run.restype = None
        L.orc_run.argtypes = [_i, _i, _i, ctypes.POINTER(OrcParams), _i] + [_dp] * 8
        L.orc_philox4x32_10.restype = None
        L.orc_philox4x32_10.argtypes = [ctypes.POINTER(_u32), ctypes.POINTER(_u32),
                                        ctypes.POINTER(_u32)]
        L.orc_nuc_uniforms.restype = None
        L.orc_nuc_uniforms.argtypes = [_u64, _u64, _u32, _u32, _dp, _dp]
        L.orc_nucleation.restype = _i
        L.orc_nucleation.argtypes = ([_i, _i] + [_dp] * 4 + [_d] * 7 + [_u64, _u64, _u32,
                                     ctypes.POINTER(ctypes.c_ubyte)])
        L.orc_summarize.restype = None
        L.orc_summarize.argtypes = [_i, _i] + [_dp] * 5
        L.orc_mobility_divergence.restype = None
        L.orc_mobility_divergence.argtypes = [_i, _i, _d, _d, ctypes.POINTER(_dp) ,
                                              ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                                              _dp, _dp, _dp, _dp]
        L.orc_energy.restype = None
        L.orc_energy.argtypes = [_i, _i, _i, _d, _d, _dp, _dp, _d] + [_dp] * 7
        L.orc_num_threads.restype = _i
        L.orc_num_threads.argtypes = []
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [_i]
        _lib = L
    return _lib


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


# ---------------------------------------------------------------- pointwise

def p(x):
    return lib().orc_p(float(x))


def pPrime(x):
    return lib().orc_pPrime(float(x))


def g(phase, xCr, xNb):
    return lib().orc_g(phase, xCr, xNb)


def dg(phase, comp, xCr, xNb):
    return lib().orc_dg(phase, comp, xCr, xNb)


def fict(xCr, xNb, pDel, pGam, pLav):
    out = (ctypes.c_double * 6)()
    lib().orc_fict(xCr, xNb, pDel, pGam, pLav, out)
    return list(out)


def D(i, j, xCr, xNb):
    return lib().orc_D(i, j, xCr, xNb)


def nucleation_probability(xCr, xNb, dG, DCrCr, DNbNb, sigma, Vatom, n_gam, dV, dt):
    R = ctypes.c_double()
    P = ctypes.c_double()
    lib().orc_nucleation_probability(xCr, xNb, dG, DCrCr, DNbNb, sigma, Vatom, n_gam, dV, dt,
                                     ctypes.byref(R), ctypes.byref(P))
    return R.value, P.value


def xe(phase, comp):
    return lib().orc_xe(phase, comp)


def d2g(phase, which):
    return lib().orc_d2g(phase, which)


def kT():
    return lib().orc_kT()


def RT():
    return lib().orc_RT()


def Vm():
    return lib().orc_Vm()


def set_diffusion_matrix(xCr0, xNb0, ifce_width):
    DCr = (ctypes.c_double * 2)()
    DNb = (ctypes.c_double * 2)()
    L = (ctypes.c_double * 2)()
    lib().orc_set_diffusion_matrix(xCr0, xNb0, ifce_width, DCr, DNb, L)
    return list(DCr), list(DNb), list(L)


def five_point_mask(dx, dy):
    m = (ctypes.c_double * 9)()
    lib().orc_five_point_mask(dx, dy, m)
    return list(m)


# ---------------------------------------------------------------- stepping

def make_params(dt, DCr, DNb, kappa, omega, Lmob, alpha, mask):
    P = OrcParams()
    P.dt = dt
    for k in range(2):
        P.DCr[k] = DCr[k]
        P.DNb[k] = DNb[k]
        P.kappa[k] = kappa[k]
        P.omega[k] = omega[k]
        P.Lmob[k] = Lmob[k]
    P.alpha = alpha
    for k in range(9):
        P.mask[k] = mask[k]
    return P


def step(fields, params, bc):
    """One step; fields = (xCr, xNb, phi_del, phi_lav) padded (ny, nx) arrays.
    Returns new arrays."""
    ny, nx = fields[0].shape
    ins = [np.ascontiguousarray(f, dtype=np.float64) for f in fields]
    outs = [np.empty_like(ins[0]) for _ in range(4)]
    lib().orc_step(nx, ny, bc, ctypes.byref(params), *[_ptr(a) for a in ins],
                   *[_ptr(a) for a in outs])
    return outs


def run(fields, params, bc, nsteps):
    ny, nx = fields[0].shape
    a = [np.array(f, dtype=np.float64, order="C", copy=True) for f in fields]
    b = [np.empty_like(a[0]) for _ in range(4)]
    lib().orc_run(nx, ny, bc, ctypes.byref(params), int(nsteps), *[_ptr(x) for x in a],
                  *[_ptr(x) for x in b])
    return a


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def nuc_uniforms(seed, call, instance, cell):
    u1 = ctypes.c_double()
    u2 = ctypes.c_double()
    lib().orc_nuc_uniforms(seed, call, instance, cell, ctypes.byref(u1), ctypes.byref(u2))
    return u1.value, u2.value


def nucleation(fields, sigma_del, sigma_lav, unit_a, dx_arg, dy_arg, dz, dt_arg,
               seed, call, instance=0):
    """In-place nucleation on copies of the fields; returns (fields, nevents, flags)."""
    ny, nx = fields[0].shape
    f = [np.array(x, dtype=np.float64, order="C", copy=True) for x in fields]
    flags = np.zeros((ny, nx), dtype=np.uint8)
    nev = lib().orc_nucleation(nx, ny, *[_ptr(x) for x in f], sigma_del, sigma_lav, unit_a,
                               dx_arg, dy_arg, dz, dt_arg, seed, call, instance,
                               flags.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
    return f, nev, flags


def summarize(fields):
    ny, nx = fields[0].shape
    out = (ctypes.c_double * 5)()
    f = [np.ascontiguousarray(x, dtype=np.float64) for x in fields]
    lib().orc_summarize(nx, ny, *[_ptr(x) for x in f], out)
    return list(out)


def num_threads():
    return lib().orc_num_threads()


def set_threads(n):
    """Threads of the oracle's parallel loops from now on."""
    lib().orc_set_threads(int(n))


# ------------------------------------------------ reference thermo (compiled)

_ref = None

REF_FUNCS_1 = ["p", "pPrime"]
REF_FUNCS_0 = ["kT", "RT", "Vm", "xe_gam_Cr", "xe_gam_Nb", "xe_del_Cr", "xe_del_Nb",
               "xe_lav_Cr", "xe_lav_Nb", "s_delta", "s_laves"] + [
    "d2g_%s_dx%s" % (ph, c) for ph in ("gam", "del", "lav")
    for c in ("CrCr", "CrNb", "NbCr", "NbNb")]
REF_FUNCS_2 = ["g_gam", "g_del", "g_lav"] + [
    "dg_%s_dx%s" % (ph, c) for ph in ("gam", "del", "lav") for c in ("Cr", "Nb")] + [
    "D_CrCr", "D_CrNb", "D_NbCr", "D_NbNb"]
REF_FUNCS_5 = ["fict_%s_%s" % (ph, c) for ph in ("gam", "del", "lav") for c in ("Cr", "Nb")]


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    """The reference's thermo/parabola625.c compiled from its own source
    (oracle/Makefile).  Only present where the reference was mounted."""
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_LIB)
        for name in REF_FUNCS_0:
            getattr(L, name).restype = _d
            getattr(L, name).argtypes = []
        for name in REF_FUNCS_1:
            getattr(L, name).restype = _d
            getattr(L, name).argtypes = [_d]
        for name in REF_FUNCS_2:
            getattr(L, name).restype = _d
            getattr(L, name).argtypes = [_d, _d]
        for name in REF_FUNCS_5:
            getattr(L, name).restype = _d
            getattr(L, name).argtypes = [_d] * 5
        _ref = L
    return _ref


def energy(fields, bc, dx, dy, kappa, omega, Omega, maps=False):
    """orc_energy: (total, sum chem, sum grad) of the free energy, and the
    per-cell chem / grad arrays when `maps`."""
    f = [np.ascontiguousarray(a, dtype=np.float64) for a in fields]
    ny, nx = f[0].shape
    chem = np.zeros((ny, nx)) if maps else None
    grad = np.zeros((ny, nx)) if maps else None
    k = np.ascontiguousarray(kappa, dtype=np.float64)
    w = np.ascontiguousarray(omega, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_energy(nx, ny, bc, dx, dy, _ptr(k), _ptr(w), Omega, *[_ptr(a) for a in f],
                     _ptr(chem) if maps else None, _ptr(grad) if maps else None, _ptr(out))
    if maps:
        return tuple(out), chem, grad
    return tuple(out)


def mobility_divergence(conc, mob_cr, mob_nb, phi_del, phi_lav, dx, dy):
    """orc_mobility_divergence: (div_Cr, div_Nb) over every cell of the padded
    arrays.  conc: 6 arrays (gam/del/lav x Cr/Nb), mob_cr / mob_nb: 8 each
    (terms 1-6 then the two phase terms)."""
    c = [np.ascontiguousarray(a, dtype=np.float64) for a in conc]
    mc = [np.ascontiguousarray(a, dtype=np.float64) for a in mob_cr]
    mn = [np.ascontiguousarray(a, dtype=np.float64) for a in mob_nb]
    pd = np.ascontiguousarray(phi_del, dtype=np.float64)
    pl = np.ascontiguousarray(phi_lav, dtype=np.float64)
    ny, nx = pd.shape
    for a in c + mc + mn:
        assert a.shape == (ny, nx)
    cp = (_dp * 6)(*[_ptr(a) for a in c])
    mcp = (_dp * 8)(*[_ptr(a) for a in mc])
    mnp = (_dp * 8)(*[_ptr(a) for a in mn])
    oc = np.empty((ny, nx))
    on = np.empty((ny, nx))
    lib().orc_mobility_divergence(nx, ny, dx, dy, cp, mcp, mnp, _ptr(pd), _ptr(pl), _ptr(oc),
                                  _ptr(on))
    return oc, on
