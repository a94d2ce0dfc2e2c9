"""Batched form over the C-ABI (pf_batch_*): many independent simulations of
uneven size in one device pool, advanced by one persistent launch per one,
two or three steps, optionally replayed from a captured hipGraph."""
import ctypes

import numpy as np

from ._lib import PfError, PfParams, check, lib
from .model import LATTICE_CONST, SIGMA

PF_STORAGE_F64 = 0
PF_STORAGE_F32 = 1


class PfJobDesc(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("bc", ctypes.c_int),
                ("nucleation", ctypes.c_int), ("dx", ctypes.c_double), ("prm", PfParams),
                ("steps", ctypes.c_uint64), ("seed", ctypes.c_uint64),
                ("instance", ctypes.c_uint32), ("nuc_every", ctypes.c_int),
                ("sigma_del", ctypes.c_double), ("sigma_lav", ctypes.c_double),
                ("unit_a", ctypes.c_double), ("x_offset", ctypes.c_double * 2)]


_vp = ctypes.c_void_p
_i = ctypes.c_int
PROTOS = {
    "pf_batch_job_bytes": (ctypes.c_size_t, [_i, _i, _i]),
    "pf_batch_create": (_i, [ctypes.POINTER(_vp), _i, ctypes.c_size_t, _i]),
    "pf_batch_destroy": (None, [_vp]),
    "pf_batch_free_bytes": (ctypes.c_size_t, [_vp]),
    "pf_batch_set_tile": (_i, [_vp, _i, _i]),
    "pf_batch_set_launch_steps": (_i, [_vp, _i]),
    "pf_batch_add": (_i, [_vp, ctypes.POINTER(PfJobDesc), _vp, _vp, _vp, _vp,
                          ctypes.POINTER(_i)]),
    "pf_batch_remove": (_i, [_vp, _i]),
    "pf_batch_commit": (_i, [_vp]),
    "pf_batch_run": (_i, [_vp, _i, _i]),
    "pf_batch_sync": (_i, [_vp]),
    "pf_batch_poll": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(_i)]),
    "pf_batch_poll_all": (_i, [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_uint64),
                               ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "pf_batch_read": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "pf_batch_summarize": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_double)]),
    "pf_batch_status": (_i, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "pf_batch_waves": (_i, [_vp]),
    "pf_batch_hold": (_i, [_vp, _i, _i]),
    "pf_batch_stream": (_vp, [_vp]),
}


def blib():
    L = lib()
    if not getattr(L, "_batch_protos", False):
        for n, (r, a) in PROTOS.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        L._batch_protos = True
    return L


def job_bytes(storage, Nx, Ny):
    return blib().pf_batch_job_bytes(storage, Nx, Ny)


def make_desc(setup, steps, nucleation=False, seed=42, instance=0, nuc_every=None, nuc_dx=None,
              x_offset=(0.0, 0.0)):
    """A job descriptor.  x_offset (FP32 storage): the job's composition arrays
    hold x - x_offset (Batch.add / read convert; see Simulation's
    storage_offset)."""
    d = PfJobDesc()
    d.Nx, d.Ny = setup.N
    d.bc = setup.bc
    d.nucleation = 1 if nucleation else 0
    d.dx = nuc_dx if nuc_dx else setup.dx
    P = d.prm
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
    d.steps = int(steps)
    d.seed = int(seed)
    d.instance = int(instance)
    d.nuc_every = int(nuc_every if nuc_every else max(1, int(0.0001 / setup.dt)))
    d.sigma_del, d.sigma_lav = SIGMA
    d.unit_a = LATTICE_CONST
    d.x_offset[0], d.x_offset[1] = float(x_offset[0]), float(x_offset[1])
    return d


class Batch:
    def __init__(self, storage=PF_STORAGE_F32, pool_bytes=1 << 30, max_jobs=4096):
        self.storage = storage
        self.dtype = np.float32 if storage == PF_STORAGE_F32 else np.float64
        h = _vp()
        check(blib().pf_batch_create(ctypes.byref(h), storage, int(pool_bytes), max_jobs),
              "pf_batch_create")
        self.h = h
        self.shapes = {}
        self.offsets = {}
        self.launch_steps = 1

    def set_tile(self, rows, depth=3):
        check(blib().pf_batch_set_tile(self.h, rows, depth), "pf_batch_set_tile")

    def set_launch_steps(self, steps):
        """Steps per launch: 1, 2 or 3 (2 and 3: the temporally blocked marches)."""
        check(blib().pf_batch_set_launch_steps(self.h, steps), "pf_batch_set_launch_steps")
        self.launch_steps = steps

    def add(self, desc, fields):
        """Admit a job (fields in absolute compositions); returns its id, or
        None if the pool is full."""
        off = (desc.x_offset[0], desc.x_offset[1], 0.0, 0.0)
        arrs = [np.ascontiguousarray(np.asarray(f, dtype=np.float64) - off[k] if off[k] else f,
                                     dtype=self.dtype) for k, f in enumerate(fields)]
        ny, nx = arrs[0].shape
        if (nx, ny) != (desc.Nx + 2, desc.Ny + 2):
            raise ValueError("field shape %s does not match the job size" % (arrs[0].shape,))
        jid = _i()
        rc = blib().pf_batch_add(self.h, ctypes.byref(desc), *[a.ctypes.data_as(_vp)
                                                               for a in arrs], ctypes.byref(jid))
        if rc == 3:  # PF_ERR_OVERFLOW: no room now
            return None
        check(rc, "pf_batch_add")
        self.shapes[jid.value] = (ny, nx)
        self.offsets[jid.value] = off[:2]
        return jid.value

    def remove(self, jid):
        check(blib().pf_batch_remove(self.h, jid), "pf_batch_remove")
        self.shapes.pop(jid, None)
        self.offsets.pop(jid, None)

    def commit(self):
        check(blib().pf_batch_commit(self.h), "pf_batch_commit")

    def run(self, launches, launches_per_graph=0):
        """Enqueue `launches` step launches (each advances every active job by
        up to set_launch_steps() steps), replayed from a captured graph of
        `launches_per_graph` launches when that is > 1."""
        check(blib().pf_batch_run(self.h, int(launches), int(launches_per_graph)),
              "pf_batch_run")

    def sync(self):
        check(blib().pf_batch_sync(self.h), "pf_batch_sync")

    def poll(self, jid):
        s = ctypes.c_uint64()
        a = _i()
        check(blib().pf_batch_poll(self.h, jid, ctypes.byref(s), ctypes.byref(a)), "poll")
        return s.value, bool(a.value)

    def poll_all(self, capacity=65536):
        ids = (_i * capacity)()
        st = (ctypes.c_uint64 * capacity)()
        ac = (_i * capacity)()
        n = _i()
        check(blib().pf_batch_poll_all(self.h, capacity, ids, st, ac, ctypes.byref(n)),
              "poll_all")
        return [(ids[k], st[k], bool(ac[k])) for k in range(min(n.value, capacity))]

    def read(self, jid, stored=False):
        """A job's current fields; compositions absolute (as FP64 when the job
        has an offset) unless `stored`."""
        ny, nx = self.shapes[jid]
        out = [np.empty((ny, nx), dtype=self.dtype) for _ in range(4)]
        check(blib().pf_batch_read(self.h, jid, *[a.ctypes.data_as(_vp) for a in out]),
              "pf_batch_read")
        off = self.offsets.get(jid, (0.0, 0.0))
        if not stored:
            for k in range(2):
                if off[k]:
                    out[k] = out[k].astype(np.float64) + off[k]
        return out

    def summarize(self, jid):
        o = (ctypes.c_double * 5)()
        check(blib().pf_batch_summarize(self.h, jid, o), "pf_batch_summarize")
        return list(o)

    def status(self, raise_on_wide=True):
        """Nucleation events that needed an embed wider than one cell (batch
        mode counts them instead of applying them).  Raises PfError when there
        were any, unless raise_on_wide is False."""
        w = ctypes.c_uint64()
        rc = blib().pf_batch_status(self.h, ctypes.byref(w))
        if rc == 1 and w.value and not raise_on_wide:   # PF_ERR_ARG: wide events counted
            return w.value
        check(rc, "pf_batch_status")
        return w.value

    def hold(self, jid, on=True):
        """Hold (or release) a job: it keeps its state but takes no part in
        the launches until released."""
        check(blib().pf_batch_hold(self.h, jid, 1 if on else 0), "pf_batch_hold")

    def stream(self):
        """The batch's hipStream_t handle (an int), e.g. for
        torch.cuda.ExternalStream."""
        return blib().pf_batch_stream(self.h)

    def waves(self):
        return blib().pf_batch_waves(self.h)

    def free_bytes(self):
        return blib().pf_batch_free_bytes(self.h)

    def close(self):
        if self.h is not None:
            blib().pf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
