"""pf_summarize_f64_noalloc (no context, no allocation: partials in a static
device array) against pf_summarize_f64 through the C-ABI: the same bits on
two calls in a row, the second on a smaller grid (the array keeps the first
call's partials, which the second must not see)."""
import ctypes

import numpy as np
import pytest
import torch

from pfpa import model as M
from pfpa._lib import check, lib

pytestmark = pytest.mark.gpu


def _noalloc():
    f = lib().pf_summarize_f64_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 6
    return f


def _fields(Nx, Ny, seed):
    f = M.init_matrix_with_particles(Nx, Ny, 0.25e-9, 1, seed=seed, radius=min(5e-9, Nx * 2e-11))
    return [torch.as_tensor(a).cuda() for a in f]


def test_noalloc_summary_matches_context_summary(gpu):
    ctx = ctypes.c_void_p()
    check(lib().pf_ctx_create(ctypes.byref(ctx)), "pf_ctx_create")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    fn = _noalloc()
    try:
        for Nx, Ny, seed in ((1500, 1300, 3), (70, 40, 4)):
            f = _fields(Nx, Ny, seed)
            ptrs = [ctypes.c_void_p(a.data_ptr()) for a in f]
            want = torch.zeros(5, dtype=torch.float64, device="cuda")
            got = torch.full((5,), float("nan"), dtype=torch.float64, device="cuda")
            check(lib().pf_summarize_f64(ctx, Nx + 2, Ny + 2, *ptrs,
                                         ctypes.c_void_p(want.data_ptr()), st), "summarize")
            check(fn(Nx + 2, Ny + 2, *ptrs, ctypes.c_void_p(got.data_ptr()), st),
                  "pf_summarize_f64_noalloc")
            torch.cuda.synchronize()
            w, g = want.cpu().numpy(), got.cpu().numpy()
            print("grid", (Nx, Ny), "summary", g.tolist(), flush=True)
            assert np.array_equal(w, g), (Nx, Ny, w, g)
            assert 0.0 < g[3] < 1.0 and 0.0 < g[4] < 1.0   # particles present
    finally:
        lib().pf_ctx_destroy(ctx)


def test_noalloc_summary_rejects_bad_sizes(gpu):
    fn = _noalloc()
    z = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = ctypes.c_void_p(z.data_ptr())
    assert fn(2, 8, p, p, p, p, p, None) != 0
    assert fn(1 << 16, 1 << 16, p, p, p, p, p, None) != 0
    assert fn(4, 4, p, p, p, p, None, None) != 0
