"""Python mirror of the reference's host API for the step path.

Same names, argument order and meaning as the C++ functions the driver calls
(src/main.cpp:236-321):

    make_arrays, five_point_Laplacian_stencil,
    nine_point_Laplacian_stencil, set_mask               src/mesh.h:20-86
    init_cuda, read_out_result, free_cuda               src/data.cuh:16-35
    device_init_prng, device_boundaries, device_fictitious, fictitious_boundaries,
    device_laplacian, device_laplacian_boundaries, device_evolution,
    device_nucleation, device_dataviz                   src/discretization.cuh:39-118
    swap_pointers_1D                                    src/mesh.h:44
    summarize_fields, summarize_energy                  src/alloy625.cpp:811-919 (on the GPU)

Device memory and streams come from torch (plumbing only); every computation
runs in libpfpa.so.  Errors raise PfError (the reference asserts in checkCuda,
src/discretization.cu:23-30).

Differences from the reference, by design:
  * device_evolution runs one fused kernel that performs the whole step
    (boundaries, fictitious compositions, Laplacians, Cahn-Hilliard and
    Allen-Cahn); the five calls before it only record that they were made.
  * Calls to device_evolution are grouped (as in the C++ drop-in): eight
    consecutive steps, the caller swapping the old/new fields in between, run
    as two launches of the four-step kernel (an even number of launches, so
    the result lands in the set the caller's swaps put it in).  Every
    function here that reads the fields first runs the pending steps, so what
    the caller observes is unchanged; device_flush runs them explicitly,
    device_set_steps_per_launch(dev, 3) groups three steps per three-step
    launch instead and device_set_steps_per_launch(dev, 1) turns grouping off.
  * The per-cell cuRAND state (seeded from clock()) is replaced by a
    counter-based Philox4x32-10 stream: device_init_prng(seed=...) sets the
    key, and each device_nucleation call advances the stream position.
  * swap_pointers_1D returns the swapped pair (Python has no pointer-to-pointer).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import model as M
from ._lib import PfNucParams, PfParams, check, lib

DEFAULT_SEED = 42


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class HostData:
    """Host arrays of src/type.h:28-53 (2-D numpy views of 1-D storage)."""

    FIELDS = ("mask_lap", "conc_Ni", "phi", "chem_nrg", "grad_nrg",
              "conc_Cr_old", "conc_Cr_new", "conc_Cr_lap", "conc_Cr_gam",
              "conc_Nb_old", "conc_Nb_new", "conc_Nb_lap", "conc_Nb_gam",
              "phi_del_old", "phi_del_new", "phi_del_lap",
              "phi_lav_old", "phi_lav_new", "phi_lav_lap")

    def __init__(self):
        for f in self.FIELDS:
            setattr(self, f, None)


class CudaData:
    """Device arrays of src/type.h:58-107 that the step path uses.

    conc_*_old/new and phi_*_old/new are torch CUDA tensors (nx*ny FP64, the
    reference's layout); the driver swaps them between steps."""

    def __init__(self):
        self.conc_Cr_old = self.conc_Cr_new = None
        self.conc_Nb_old = self.conc_Nb_new = None
        self.phi_del_old = self.phi_del_new = None
        self.phi_lav_old = self.phi_lav_new = None
        self.conc_Ni = self.phi = None
        self.ctx = None
        self.params = None
        self.bc = _lib.PF_BC_NOFLUX
        self.seed = DEFAULT_SEED
        self.nuc_calls = 0
        self.pending = []
        self.group = 4        # steps per launch of grouped device_evolution calls
        self.queued = 0       # steps recorded and not yet run (0..2)
        self.first = None     # (old, new, alpha, dt, nx, ny) of the first of them


def make_arrays(host, nx, ny, nm):
    """src/mesh.cpp:12-97."""
    for f in HostData.FIELDS:
        if f == "mask_lap":
            host.mask_lap = np.zeros((nm, nm))
        else:
            setattr(host, f, np.zeros((ny, nx)))


def free_arrays(host):
    for f in HostData.FIELDS:
        setattr(host, f, None)


def five_point_Laplacian_stencil(dx, dy, mask_lap, nm):
    """src/mesh.cpp:187-196 (fills mask_lap in place)."""
    if nm != 3:
        raise ValueError("five_point_Laplacian_stencil requires nm == 3")
    mask_lap[0, 1] = 1.0 / (dy * dy)
    mask_lap[1, 0] = 1.0 / (dx * dx)
    mask_lap[1, 1] = -2.0 * (dx * dx + dy * dy) / (dx * dx * dy * dy)
    mask_lap[1, 2] = 1.0 / (dx * dx)
    mask_lap[2, 1] = 1.0 / (dy * dy)


def nine_point_Laplacian_stencil(dx, dy, mask_lap, nm):
    """src/mesh.cpp:198-213 (fills mask_lap in place)."""
    if nm != 3:
        raise ValueError("nine_point_Laplacian_stencil requires nm == 3")
    mask_lap[:, :] = np.array(M.nine_point_mask(dx, dy)).reshape(3, 3)


def set_mask(dx, dy, code, mask_lap, nm):
    """src/mesh.cpp:170-185: 93 selects the nine-point mask, anything else the
    five-point one."""
    if code == 93:
        nine_point_Laplacian_stencil(dx, dy, mask_lap, nm)
    else:
        five_point_Laplacian_stencil(dx, dy, mask_lap, nm)


def swap_pointers_1D(a, b):
    """src/mesh.cpp:163-168."""
    return b, a


def init_cuda(host, nx, ny, nm, kappa, omega, Lmob, D_Cr, D_Nb, dev, bc=_lib.PF_BC_NOFLUX):
    """src/data.cu:10-92: allocate device fields, upload parameters and the
    host's *_old fields.  bc selects no-flux (the reference's) or periodic."""
    if nm != 3:
        raise ValueError("only 3x3 stencils (nm == 3) are supported")
    if not torch.cuda.is_available():
        raise _lib.PfError("init_cuda: no GPU visible")
    h = ctypes.c_void_p()
    check(lib().pf_ctx_create(ctypes.byref(h)), "pf_ctx_create")
    dev.ctx = h
    dev.bc = int(bc)
    P = PfParams()
    for k in range(2):
        P.DCr[k] = D_Cr[k]
        P.DNb[k] = D_Nb[k]
        P.kappa[k] = kappa[k]
        P.omega[k] = omega[k]
        P.Lmob[k] = Lmob[k]
    mask = np.asarray(host.mask_lap, dtype=np.float64).reshape(-1)
    for k in range(9):
        P.mask[k] = mask[k]
    dev.params = P
    opts = dict(dtype=torch.float64, device="cuda")
    for name in ("conc_Cr", "conc_Nb", "phi_del", "phi_lav"):
        src = np.ascontiguousarray(getattr(host, name + "_old"), dtype=np.float64)
        if src.shape != (ny, nx):
            raise ValueError("host.%s_old has shape %s, expected %s" % (name, src.shape, (ny, nx)))
        setattr(dev, name + "_old", torch.from_numpy(src).to("cuda").reshape(-1))
        setattr(dev, name + "_new", torch.zeros(nx * ny, **opts))
    dev.conc_Ni = torch.zeros(nx * ny, **opts)
    dev.phi = torch.zeros(nx * ny, **opts)
    dev.nx, dev.ny = nx, ny


def free_cuda(dev):
    """src/data.cu:94-154."""
    if dev.ctx is not None:
        device_flush(dev)
        torch.cuda.synchronize()
        lib().pf_ctx_destroy(dev.ctx)
        dev.ctx = None
    for name in ("conc_Cr", "conc_Nb", "phi_del", "phi_lav"):
        setattr(dev, name + "_old", None)
        setattr(dev, name + "_new", None)
    dev.conc_Ni = dev.phi = None


def device_init_prng(dev, nx, ny, nm, bx, by, seed=DEFAULT_SEED):
    """src/discretization.cu:1033-1044: key the counter-based stream."""
    dev.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    dev.nuc_calls = 0


def _record(dev, name):
    dev.pending.append(name)


def device_boundaries(dev, nx, ny, nm, bx, by):
    """src/discretization.cu:396-451 -- folded into device_evolution."""
    _record(dev, "boundaries")


def device_fictitious(dev, nx, ny, nm, bx, by):
    """src/discretization.cu:128-173 -- folded into device_evolution."""
    _record(dev, "fictitious")


def fictitious_boundaries(dev, nx, ny, nm, bx, by):
    """src/discretization.cu:453-480 -- folded into device_evolution."""
    _record(dev, "fictitious_boundaries")


def device_laplacian(dev, nx, ny, nm, bx, by, dx, dy):
    """src/discretization.cu:780-841 -- folded into device_evolution (the mask
    uploaded by init_cuda carries 1/dx^2; dx, dy are unused, as there)."""
    _record(dev, "laplacian")


def device_laplacian_boundaries(dev, nx, ny, nm, bx, by):
    """src/discretization.cu:524-548 -- folded into device_evolution."""
    _record(dev, "laplacian_boundaries")


_FIELDS = ("conc_Cr", "conc_Nb", "phi_del", "phi_lav")


def _fields(dev, which):
    return tuple(getattr(dev, f + which) for f in _FIELDS)


def _launch(dev, fn, nx, ny, alpha, dt, old, new, *extra):
    P = dev.params
    P.alpha = alpha
    P.dt = dt
    check(fn(dev.ctx, nx, ny, dev.bc, ctypes.byref(P), *map(_ptr, old), *map(_ptr, new),
             *extra, _stream()), "device_evolution")


def pending_launches(p, four):
    """Steps per launch of the launches that run p pending steps.  After p
    steps the newest fields are in the first step's new set if p is odd and
    in its old set if p is even; launches alternate between the two sets, so
    their number has the parity of p: pairs of four-step launches (`four`),
    then three-step launches, then one-step launches."""
    out = []
    while four and p >= 8:
        out += [4, 4]
        p -= 8
    while p >= 3:
        out.append(3)
        p -= 3
    return out + [1] * p


def group_size(group, nx, ny):
    """Pending steps that trigger the launches: eight (two four-step
    launches) with four steps per launch on a grid of at least 4x4 interior
    cells, else three."""
    return 8 if group == 4 and nx >= 6 and ny >= 6 else 3


def device_flush(dev):
    """Run the steps device_evolution has recorded and not yet launched
    (pending_launches: first old -> new, then back)."""
    if dev.queued:
        old, new, alpha, dt, nx, ny = dev.first
        fns = {1: lib().pf_step_f64, 3: lib().pf_step3_f64, 4: lib().pf_step4_f64}
        four = group_size(dev.group, nx, ny) == 8
        for i, k in enumerate(pending_launches(dev.queued, four)):
            a, b = (old, new) if i % 2 == 0 else (new, old)
            _launch(dev, fns[k], nx, ny, alpha, dt, a, b, *([None] if k == 1 else []))
    dev.queued = 0
    dev.first = None


def device_set_steps_per_launch(dev, steps):
    """4 (default): group device_evolution calls eight to two four-step
    launches; 3: three per three-step launch; 1: one per launch."""
    device_flush(dev)
    dev.group = steps if steps in (3, 4) else 1


def device_evolution(dev, nx, ny, nm, bx, by, alpha, dt):
    """src/discretization.cu:982-1020: one fused step *_old -> *_new, grouped
    with the calls around it (see the module docstring)."""
    dev.pending.clear()
    old, new = _fields(dev, "_old"), _fields(dev, "_new")
    five_point = M.symmetric_five_point(dev.params.mask)
    if dev.group == 1 or not five_point or nx < 5 or ny < 5:
        device_flush(dev)
        _launch(dev, lib().pf_step_f64, nx, ny, alpha, dt, old, new, None)
        return
    if dev.queued:
        f_old, f_new, f_alpha, f_dt, f_nx, f_ny = dev.first
        want_old, want_new = (f_new, f_old) if dev.queued % 2 else (f_old, f_new)
        same = (lambda a, b: all(x.data_ptr() == y.data_ptr() for x, y in zip(a, b)))
        if not (same(old, want_old) and same(new, want_new) and (alpha, dt, nx, ny) ==
                (f_alpha, f_dt, f_nx, f_ny)):
            device_flush(dev)
    if not dev.queued:
        dev.first = (old, new, alpha, dt, nx, ny)
    dev.queued += 1
    if dev.queued == group_size(dev.group, nx, ny):
        device_flush(dev)


def nuc_params(dev, sigma_del, sigma_lav, unit_a, ifce_width, dx, dy, dz, dt):
    N = PfNucParams()
    N.sigma_del, N.sigma_lav, N.unit_a, N.ifce_width = sigma_del, sigma_lav, unit_a, ifce_width
    N.dx, N.dy, N.dz, N.dt = dx, dy, dz, dt
    N.seed = dev.seed
    N.call = dev.nuc_calls
    N.instance = 0
    return N


def device_nucleation(dev, nx, ny, nm, bx, by, sigma_del, sigma_lav, unit_a, ifce_width,
                      dx, dy, dz, dt):
    """src/discretization.cu:1179-1209: nucleation pass on the *_new fields."""
    device_flush(dev)
    N = nuc_params(dev, sigma_del, sigma_lav, unit_a, ifce_width, dx, dy, dz, dt)
    check(lib().pf_nucleation_f64(dev.ctx, nx, ny, _ptr(dev.conc_Cr_new), _ptr(dev.conc_Nb_new),
                                  _ptr(dev.phi_del_new), _ptr(dev.phi_lav_new),
                                  ctypes.byref(N), _stream()), "device_nucleation")
    dev.nuc_calls += 1


def read_out_result(dev, host, nx, ny):
    """src/data.cu:156-172: device *_old -> host *_new (synchronous)."""
    device_flush(dev)
    torch.cuda.synchronize()
    for name in ("conc_Cr", "conc_Nb", "phi_del", "phi_lav"):
        getattr(host, name + "_new")[...] = (
            getattr(dev, name + "_old").reshape(ny, nx).cpu().numpy())


def device_dataviz(dev, host, nx, ny, nm, bx, by):
    """src/discretization.cu:1231-1263."""
    device_flush(dev)
    check(lib().pf_dataviz_f64(nx, ny, _ptr(dev.conc_Cr_old), _ptr(dev.conc_Nb_old),
                               _ptr(dev.phi_del_old), _ptr(dev.phi_lav_old),
                               _ptr(dev.conc_Ni), _ptr(dev.phi), _stream()), "device_dataviz")
    read_out_result(dev, host, nx, ny)
    host.conc_Ni[...] = dev.conc_Ni.reshape(ny, nx).cpu().numpy()
    host.phi[...] = dev.phi.reshape(ny, nx).cpu().numpy()


def summarize_fields(dev, nx, ny):
    """Means of {x_Cr, x_Nb, gamma, delta, Laves} over the interior of the
    latest fields (src/alloy625.cpp:811-863), on the GPU."""
    device_flush(dev)
    out = torch.zeros(5, dtype=torch.float64, device="cuda")
    check(lib().pf_summarize_f64(dev.ctx, nx, ny, _ptr(dev.conc_Cr_old), _ptr(dev.conc_Nb_old),
                                 _ptr(dev.phi_del_old), _ptr(dev.phi_lav_old), _ptr(out),
                                 _stream()), "summarize_fields")
    return out.cpu().tolist()


def summarize_energy(dev, nx, ny, dx, dy, host=None, Omega=None):
    """Total free energy of the latest fields, src/alloy625.cpp:865-919
    (gibbs + kappa |grad phi|^2, times dV = dx dy, summed over the interior),
    on the GPU with compensated sums.  kappa and omega are the values init_cuda
    received; Omega defaults to src/parameters.h:37.  With `host`, the per-cell
    energies go to host.chem_nrg / host.grad_nrg as the reference writes
    them.  Returns (total, chemical, gradient)."""
    from .model import OMEGA_3PHASE
    device_flush(dev)
    P = dev.params
    kappa = (ctypes.c_double * 2)(P.kappa[0], P.kappa[1])
    omega = (ctypes.c_double * 2)(P.omega[0], P.omega[1])
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    chem = grad = None
    if host is not None:
        chem = torch.zeros(ny * nx, dtype=torch.float64, device="cuda")
        grad = torch.zeros(ny * nx, dtype=torch.float64, device="cuda")
    check(lib().pf_energy_f64(dev.ctx, nx, ny, dev.bc, dx, dy, kappa, omega,
                              OMEGA_3PHASE if Omega is None else Omega,
                              _ptr(dev.conc_Cr_old), _ptr(dev.conc_Nb_old),
                              _ptr(dev.phi_del_old), _ptr(dev.phi_lav_old),
                              _ptr(chem) if chem is not None else None,
                              _ptr(grad) if grad is not None else None, _ptr(out), _stream()),
          "summarize_energy")
    if host is not None:
        c = chem.reshape(ny, nx).cpu().numpy()
        g = grad.reshape(ny, nx).cpu().numpy()
        host.chem_nrg[1:-1, 1:-1] = c[1:-1, 1:-1]
        host.grad_nrg[1:-1, 1:-1] = g[1:-1, 1:-1]
    return tuple(out.cpu().tolist())


def nucleation_status(dev):
    wide = ctypes.c_uint64()
    over = ctypes.c_int()
    rc = lib().pf_nucleation_status(dev.ctx, ctypes.byref(wide), ctypes.byref(over))
    check(rc, "pf_nucleation_status")
    return wide.value, bool(over.value)


def chemical_convolution(conc, mob_Cr, mob_Nb, phi_del, phi_lav, dx, dy, div_Cr=None,
                         div_Nb=None, stream=None):
    """The variable-mobility Cahn-Hilliard divergence of
    chemical_convolution_Cr_kernel / chemical_convolution_Nb_kernel
    (src/discretization.cu:605-778) on device tensors: conc = the six
    fictitious compositions (gam, del, lav x Cr, Nb), mob_Cr / mob_Nb = the
    eight mobility arrays of each component (terms 1-6, then the delta and
    Laves phase terms, which are differenced in x only), every array a CUDA
    FP64 tensor of the padded (ny, nx) shape with its ghost ring filled.
    Returns (div_Cr, div_Nb) over every cell.  The reference never launches
    these kernels (src/discretization.cu:809-835) and leaves the mobility
    functions d_M_* undefined, so the mobilities are the caller's."""
    arrs = list(conc) + list(mob_Cr) + list(mob_Nb) + [phi_del, phi_lav]
    if len(conc) != 6 or len(mob_Cr) != 8 or len(mob_Nb) != 8:
        raise ValueError("need 6 compositions and 8 + 8 mobility arrays")
    shape = phi_del.shape
    for a in arrs:
        if a.dtype != torch.float64 or not a.is_cuda or a.shape != shape or \
                not a.is_contiguous():
            raise ValueError("every array must be a contiguous CUDA float64 tensor of one shape")
    ny, nx = shape
    if div_Cr is None:
        div_Cr = torch.empty_like(phi_del)
    if div_Nb is None:
        div_Nb = torch.empty_like(phi_del)
    f = _lib.PfMobilityFields()
    for k in range(6):
        f.conc[k] = conc[k].data_ptr()
    for k in range(8):
        f.mob_Cr[k] = mob_Cr[k].data_ptr()
        f.mob_Nb[k] = mob_Nb[k].data_ptr()
    f.phi_del = phi_del.data_ptr()
    f.phi_lav = phi_lav.data_ptr()
    st = ctypes.c_void_p(stream.cuda_stream) if stream is not None else _stream()
    check(lib().pf_mobility_divergence_f64(nx, ny, dx, dy, ctypes.byref(f), _ptr(div_Cr),
                                           _ptr(div_Nb), st), "pf_mobility_divergence_f64")
    return div_Cr, div_Nb
