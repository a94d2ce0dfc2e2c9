"""Native stepping loop over the C-ABI.  Steps run up to four per launch
(the temporally blocked kernels) wherever no step of the group is a
nucleation step; nucleation steps, the ones the reference's schedule selects
(kernel_time % nuc_interval == 0, src/main.cpp:307-316), run one per launch
with the nucleation pass fused."""
import ctypes

import numpy as np
import torch

from ._lib import PfNucParams, PfParams, check, lib
from . import model as M
from .model import LATTICE_CONST, IFCE_WIDTH, SIGMA


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def iter_launches(t0, n, max_steps, nuc_every=0):
    """Launches that advance steps t0 .. t0+n-1, one (k, nucleation) at a
    time with k steps per launch.  A launch groups up to `max_steps` steps
    none of which is a nucleation step (t % nuc_every == 0,
    src/main.cpp:307-316); a nucleation step runs alone with the nucleation
    pass (nucleation True).  nuc_every 0 = nucleation off.  O(1) per launch,
    so a stepping loop can start launching at once."""
    t, end = t0, t0 + n
    while t < end:
        if nuc_every > 0:
            nxt = -(-t // nuc_every) * nuc_every  # first nucleation step >= t
            if nxt == t:
                yield (1, True)
                t += 1
                continue
            k = min(max_steps, end - t, nxt - t)
        else:
            k = min(max_steps, end - t)
        yield (k, False)
        t += k


# FP32 storage: four steps per launch below this many interior cells, three
# from it on (the four-step kernel is 5-15% faster at 1024^2-2048^2 and 3-4%
# slower at 3072^2-8192^2, where the FP32 kernels are bound by arithmetic)
FP32_FOUR_STEP_CELLS = 2560 * 2560


def default_steps_per_launch(setup, dtype, requested=None):
    """Steps per launch of a Simulation: `requested` (1 to 4), or by default
    the faster kernel on the grids measured: four for FP64 storage, and for
    FP32 four below FP32_FOUR_STEP_CELLS interior cells and three above;
    reduced to what the grid and mask allow: the multi-step kernels take
    symmetric five-point masks and an interior of at least 3x3 (four steps:
    4x4)."""
    if requested is None:
        cells = (setup.nx - 2) * (setup.ny - 2)
        requested = 4 if dtype == torch.float64 or cells < FP32_FOUR_STEP_CELLS else 3
    if requested not in (1, 2, 3, 4):
        raise ValueError("steps_per_launch must be 1, 2, 3 or 4")
    if not (M.symmetric_five_point(setup.mask) and setup.nx >= 5 and setup.ny >= 5):
        return 1
    if requested == 4 and (setup.nx < 6 or setup.ny < 6):
        return 3
    return requested


def plan_launches(t0, n, max_steps, nuc_every=0):
    """iter_launches as a list."""
    return list(iter_launches(t0, n, max_steps, nuc_every))


class Simulation:
    """Double-buffered fields resident on the GPU.

    dtype torch.float64 is the plain form; torch.float32 the precision form
    (FP32 storage, FP64 arithmetic).

    storage_offset (FP32 storage only): None keeps x in the arrays; "median",
    "mean" or (x_Cr, x_Nb) stores x - offset ("median" / "mean" of the initial
    interior compositions; the median is the matrix composition wherever the
    matrix is most of the domain, which is where the offset pays: an FP32 ulp
    of x - offset there is 8-100x below that of x, so each step's small
    composition increment keeps that many more bits).  Simulations sharing a
    context share its offset (pf_ctx_set_storage_offset).
    fields(), summarize() and energy() report absolute compositions."""

    def __init__(self, setup, fields, dtype=torch.float64, nucleation=False, seed=42,
                 nuc_every=None, instance=0, stream=None, ctx=None, nuc_dx=None,
                 steps_per_launch=None, storage_offset=None):
        if not torch.cuda.is_available():
            raise RuntimeError("Simulation needs a GPU")
        self.s = setup
        self.dtype = dtype
        self.nx, self.ny = setup.nx, setup.ny
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        self.xoff = (0.0, 0.0)
        if storage_offset is not None:
            if dtype != torch.float32:
                raise ValueError("storage_offset applies to FP32 storage only")
            if isinstance(storage_offset, str):
                stat = {"median": np.median, "mean": np.mean}.get(storage_offset)
                if stat is None:
                    raise ValueError("storage_offset: None, 'median', 'mean' or (x_Cr, x_Nb)")
                self.xoff = tuple(float(stat(np.asarray(f, dtype=np.float64)[1:-1, 1:-1]))
                                  for f in fields[:2])
            else:
                self.xoff = (float(storage_offset[0]), float(storage_offset[1]))
        self.owns_ctx = ctx is None
        if ctx is None:
            h = ctypes.c_void_p()
            check(lib().pf_ctx_create(ctypes.byref(h)), "pf_ctx_create")
            ctx = h
        self.ctx = ctx
        if dtype == torch.float32:
            check(lib().pf_ctx_set_storage_offset(ctx, self.xoff[0], self.xoff[1]),
                  "pf_ctx_set_storage_offset")

        def stored(k, f):
            off = self.xoff[k] if k < 2 else 0.0
            return np.ascontiguousarray(np.asarray(f, dtype=np.float64) - off if off else f)

        with torch.cuda.stream(self.stream):
            self.a = [torch.as_tensor(stored(k, f)).to(device="cuda", dtype=dtype)
                      for k, f in enumerate(fields)]
            self.b = [torch.empty_like(x) for x in self.a]
        P = PfParams()
        P.dt = setup.dt
        for k in range(2):
            P.DCr[k] = setup.DCr[k]
            P.DNb[k] = setup.DNb[k]
            P.kappa[k] = setup.kappa[k]
            P.omega[k] = setup.omega[k]
            P.Lmob[k] = setup.Lmob[k]
        P.alpha = setup.alpha
        for k in range(9):
            P.mask[k] = setup.mask[k]
        self.P = P
        self.nucleation = nucleation
        self.nuc_every = nuc_every if nuc_every else max(1, int(0.0001 / setup.dt))
        N = PfNucParams()
        N.sigma_del, N.sigma_lav = SIGMA
        N.unit_a = LATTICE_CONST
        N.ifce_width = IFCE_WIDTH
        # mesh spacing the nucleation pass uses (the stencil's by default)
        N.dx = N.dy = N.dz = nuc_dx if nuc_dx else setup.dx
        N.dt = self.nuc_every * setup.dt
        N.seed = seed
        N.call = 0
        N.instance = instance
        self.N = N
        self.steps_done = 0
        self._step = lib().pf_step_f64 if dtype == torch.float64 else lib().pf_step_f32
        f64 = dtype == torch.float64
        self._stepk = {2: lib().pf_step2_f64 if f64 else lib().pf_step2_f32,
                       3: lib().pf_step3_f64 if f64 else lib().pf_step3_f32,
                       4: lib().pf_step4_f64 if f64 else lib().pf_step4_f32}
        self.steps_per_launch = default_steps_per_launch(setup, dtype, steps_per_launch)
        self._summ = lib().pf_summarize_f64 if dtype == torch.float64 else lib().pf_summarize_f32

    def set_tile(self, rows, remap=1, depth=0):
        check(lib().pf_ctx_set_tile(self.ctx, rows, remap, depth), "pf_ctx_set_tile")

    def set_tile2(self, rows, depth=0):
        check(lib().pf_ctx_set_tile2(self.ctx, rows, depth), "pf_ctx_set_tile2")

    def set_tile3(self, rows, depth=0):
        check(lib().pf_ctx_set_tile3(self.ctx, rows, depth), "pf_ctx_set_tile3")

    def set_tile4(self, rows, depth=0):
        check(lib().pf_ctx_set_tile4(self.ctx, rows, depth), "pf_ctx_set_tile4")

    def tune_tile3(self, rows_set=(6, 8, 12, 16, 24, 32, 48, 64, 72), launches=20):
        """tune_tile for the three-step kernel (the simulation must run three
        steps per launch)."""
        if self.steps_per_launch != 3:
            raise ValueError("tune_tile3 needs three steps per launch")
        return self.tune_tile(rows_set, launches)

    def tune_tile(self, rows_set=None, launches=20):
        """Time the simulation's multi-step kernel (steps_per_launch 2, 3 or
        4) at each chunk height in `rows_set` on scratch copies of the
        current fields (the simulation does not advance) and keep the
        fastest.  The default heights (a launch model for four steps, a
        heuristic for two and three) are within a few per cent of the best on
        the grids measured; this measures the grid at hand.  rows_set None:
        the kernel's default height and 6, 8, 12, 16, 24, 32, 48, 60, 64, 72,
        96 (those up to the interior height).  Results do not depend on the
        tile.  Returns (rows, {rows: us per launch})."""
        k = self.steps_per_launch
        if k < 2:
            raise ValueError("tune_tile needs two or more steps per launch")
        if rows_set is None:
            storage = 0 if self.dtype == torch.float64 else 1
            default = getattr(lib(), "pf_default_rows%d" % k)(self.nx, self.ny, storage)
            rows_set = sorted({default} | {r for r in (6, 8, 12, 16, 24, 32, 48, 60, 64, 72, 96)
                                          if r <= self.ny - 2})
        fn = self._stepk[k]
        set_tile = {2: self.set_tile2, 3: self.set_tile3, 4: self.set_tile4}[k]
        st = ctypes.c_void_p(self.stream.cuda_stream)
        # the copies are made on the simulation's stream, after the steps
        # queued there, and the kernels that read them run there too
        with torch.cuda.stream(self.stream):
            a = [x.clone() for x in self.a]
            b = [torch.empty_like(x) for x in a]

        def run(n):
            for i in range(n):
                x, y = (a, b) if i % 2 == 0 else (b, a)
                check(fn(self.ctx, self.nx, self.ny, self.s.bc, ctypes.byref(self.P),
                         *[_ptr(t) for t in x], *[_ptr(t) for t in y], st), "pf_step%d" % k)

        times = {}
        for rows in rows_set:
            set_tile(rows, 3)
            run(2)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(self.stream)
            run(launches)
            e1.record(self.stream)
            e1.synchronize()
            times[rows] = e0.elapsed_time(e1) * 1e3 / launches
        best = min(times, key=times.get)
        set_tile(best, 3)
        return best, times

    def step(self, n=1):
        st = ctypes.c_void_p(self.stream.cuda_stream)
        plan = iter_launches(self.steps_done, n, self.steps_per_launch,
                             self.nuc_every if self.nucleation else 0)
        for k, nucleate in plan:
            a, b = self.a, self.b
            t = self.steps_done
            if k > 1:
                check(self._stepk[k](self.ctx, self.nx, self.ny, self.s.bc, ctypes.byref(self.P),
                                     _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), _ptr(a[3]),
                                     _ptr(b[0]), _ptr(b[1]), _ptr(b[2]), _ptr(b[3]), st),
                      "pf_step%d" % k)
                done = k
            else:
                nuc = None
                if nucleate:
                    self.N.call = t // self.nuc_every
                    nuc = ctypes.byref(self.N)
                check(self._step(self.ctx, self.nx, self.ny, self.s.bc, ctypes.byref(self.P),
                                 _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), _ptr(a[3]),
                                 _ptr(b[0]), _ptr(b[1]), _ptr(b[2]), _ptr(b[3]), nuc, st),
                      "pf_step")
                done = 1
            self.a, self.b = self.b, self.a
            self.steps_done += done

    def fields(self):
        """The current fields as FP64 arrays, compositions in absolute units."""
        self.stream.synchronize()
        out = [x.reshape(self.ny, self.nx).double().cpu().numpy() for x in self.a]
        for k in range(2):
            if self.xoff[k] != 0.0:
                out[k] = out[k] + self.xoff[k]
        return out

    def summarize(self):
        # allocated and cleared on the simulation's stream, in order with the
        # reduction that writes it (a stream of its own need not wait for the
        # current one)
        with torch.cuda.stream(self.stream):
            out = torch.zeros(5, dtype=torch.float64, device="cuda")
        a = self.a
        check(self._summ(self.ctx, self.nx, self.ny, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]),
                         _ptr(a[3]), _ptr(out), ctypes.c_void_p(self.stream.cuda_stream)),
              "pf_summarize")
        self.stream.synchronize()
        return out.cpu().tolist()

    def energy(self, maps=False):
        """Free energy of the current fields (src/alloy625.cpp:865-919):
        (total, chemical, gradient), and the per-cell chem / grad arrays when
        `maps`."""
        f64 = self.dtype == torch.float64
        fn = lib().pf_energy_f64 if f64 else lib().pf_energy_f32
        s = self.s
        kappa = (ctypes.c_double * 2)(*s.kappa)
        omega = (ctypes.c_double * 2)(*s.omega)
        chem = grad = None
        with torch.cuda.stream(self.stream):
            out = torch.zeros(3, dtype=torch.float64, device="cuda")
            if maps:
                chem = torch.zeros(self.nx * self.ny, dtype=torch.float64, device="cuda")
                grad = torch.zeros_like(chem)
        a = self.a
        check(fn(self.ctx, self.nx, self.ny, s.bc, s.dx, s.dx, kappa, omega, s.alpha,
                 _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), _ptr(a[3]),
                 _ptr(chem) if maps else None, _ptr(grad) if maps else None, _ptr(out),
                 ctypes.c_void_p(self.stream.cuda_stream)), "pf_energy")
        self.stream.synchronize()
        tot = tuple(out.cpu().tolist())
        if maps:
            return tot, chem.reshape(self.ny, self.nx).cpu().numpy(), \
                grad.reshape(self.ny, self.nx).cpu().numpy()
        return tot

    def close(self):
        if self.ctx is not None and self.owns_ctx:
            torch.cuda.synchronize()
            lib().pf_ctx_destroy(self.ctx)
        self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
