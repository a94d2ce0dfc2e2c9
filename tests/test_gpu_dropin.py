"""The drop-in host API driven exactly as the reference's driver drives it
(src/main.cpp:225-327), from C++ (tests/cpp/replay_main) and from Python
(pfpa.api), checked against the oracle."""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

from pfpa import api
from pfpa import model as M

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "cpp", "replay_main")
TOL = 1e-12


def mixed_err(a, b):
    a = a[1:-1, 1:-1]
    b = b[1:-1, 1:-1]
    sf = max(np.abs(b).max(), 1e-300)
    return float((np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), sf)).max())


def oracle_reference(oracle, f0, setup, steps):
    P = oracle.make_params(setup.dt, setup.DCr, setup.DNb, setup.kappa, setup.omega,
                           setup.Lmob, setup.alpha, setup.mask)
    return oracle.run(f0, P, 0, steps)


def run_harness(f0, s, steps, nuc, io, spl=None, bc=None, mask_code=None):
    Nx, Ny = s.N
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(fin, "wb") as fh:
            fh.write(struct.pack("ii", Nx + 2, Ny + 2))
            for a in f0:
                fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        args = [HARNESS, fin, fout, str(steps), repr(s.dt), repr(s.dx), str(nuc), "42", str(io)]
        if spl is not None or bc is not None:
            args.append(str(spl or 0))
        if bc is not None:
            args.append(str(bc))
            if mask_code is not None:
                args.append(str(mask_code))
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        raw = np.fromfile(fout, dtype=np.uint8)
    return raw[8:].view(np.float64).reshape(6, Ny + 2, Nx + 2)


def test_cpp_driver_replay(gpu, oracle):
    """The reference's call sequence with nucleation every 100 steps and I/O
    every 150: grouped (four steps per launch, the default; and three) and one
    launch per call give the same bits, and all match the oracle."""
    assert os.path.exists(HARNESS), "tests/cpp/replay_main not built: run __graft_entry__.build()"
    Nx, Ny = 120, 90
    s = M.make_setup(Nx, Ny, bc=0)
    f0 = M.init_matrix_with_particles(Nx, Ny, s.dx, 0, seed=5, radius=2.5e-9)
    steps = 300
    arr = run_harness(f0, s, steps, 100, 150)
    one = run_harness(f0, s, steps, 100, 150, spl=1)
    three = run_harness(f0, s, steps, 100, 150, spl=3)
    want = oracle_reference(oracle, f0, s, steps)
    assert np.array_equal(three, one)
    for k in range(4):
        assert np.array_equal(arr[k], one[k]), k
        assert mixed_err(arr[k], want[k]) <= TOL, k
    # visualisation fields were produced at the last I/O step (step 300)
    assert np.array_equal(arr[4:], one[4:])
    assert np.allclose(arr[4], 1.0 - arr[0] - arr[1], atol=1e-15)
    p = lambda x: x ** 3 * (6 * x * x - 15 * x + 10)  # noqa: E731
    assert np.allclose(arr[5], p(arr[2]) + p(arr[3]), rtol=1e-13, atol=1e-15)
    assert (Nx + 2) * (Ny + 2) == arr[0].size


@pytest.mark.parametrize("bc,mask_code", [(1, 53), (1, 93), (0, 93)])
def test_cpp_driver_replay_boundary_and_mask(gpu, oracle, bc, mask_code):
    """device_set_boundary (periodic) and set_mask(.., 93, ..) (the nine-point
    mask, one launch per call) through the C++ API: grouped == one launch per
    call, and both match the oracle with the same mask."""
    Nx, Ny = 96, 72
    s = M.make_setup(Nx, Ny, bc=bc, dt_rule=5e-8)
    if mask_code == 93:
        s.mask = M.nine_point_mask(s.dx, s.dx)
    f0 = M.init_matrix_with_particles(Nx, Ny, s.dx, bc, seed=13, radius=2e-9)
    steps = 40
    a = run_harness(f0, s, steps, 0, 0, spl=3, bc=bc, mask_code=mask_code)
    b = run_harness(f0, s, steps, 0, 0, spl=1, bc=bc, mask_code=mask_code)
    c = run_harness(f0, s, steps, 0, 0, spl=4, bc=bc, mask_code=mask_code)
    P = oracle.make_params(s.dt, s.DCr, s.DNb, s.kappa, s.omega, s.Lmob, s.alpha, s.mask)
    want = oracle.run(f0, P, bc, steps)
    for k in range(4):
        assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(c[k], b[k]), k
        assert mixed_err(a[k], want[k]) <= TOL, k


def test_cpp_driver_replay_4096(gpu, oracle):
    """The drop-in benchmark's grid: the reference's call sequence on 4096^2
    (no-flux, the enriched-band workload, nucleation every 891 steps, I/O at
    the last step), 11 steps grouped eight to two four-step launches and a
    three-step launch, every cell against the oracle."""
    assert os.path.exists(HARNESS), "tests/cpp/replay_main not built: run __graft_entry__.build()"
    s = M.make_setup(4096, 4096, bc=0)
    f0 = M.init_enriched_band(4096, 4096, 0, seed=42)
    steps = 11
    arr = run_harness(f0, s, steps, 891, steps)
    want = oracle_reference(oracle, f0, s, steps)
    for k in range(4):
        assert mixed_err(arr[k], want[k]) <= TOL, k


@pytest.mark.parametrize("io", [1, 2, 7])
def test_cpp_driver_replay_frequent_io(gpu, io):
    """I/O (device_dataviz + read_out_result) every 1, 2 or 7 steps breaks the
    groups at every phase; results equal one launch per call."""
    Nx, Ny = 80, 48
    s = M.make_setup(Nx, Ny, bc=1)
    f0 = M.init_matrix_with_particles(Nx, Ny, s.dx, 1, seed=9, radius=1.2e-9)
    a = run_harness(f0, s, 23, 5, io)      # 23 steps: a pending tail flushed by read-out
    b = run_harness(f0, s, 23, 5, io, spl=1)
    c = run_harness(f0, s, 23, 5, io, spl=3)
    assert np.array_equal(a, b)
    assert np.array_equal(c, b)


@pytest.mark.parametrize("spl", [4, 3, 1])
def test_python_api_replay(gpu, oracle, spl):
    """pfpa.api in the reference's call sequence, I/O steps included, with
    grouped launches (4, the default, and 3) and one launch per call."""
    Nx, Ny = 100, 70
    s = M.make_setup(Nx, Ny, bc=0)
    f0 = M.init_matrix_with_particles(Nx, Ny, s.dx, 0, seed=8, radius=2e-9)
    nx, ny, nm, bx, by = Nx + 2, Ny + 2, 3, 13, 13
    host = api.HostData()
    dev = api.CudaData()
    api.make_arrays(host, nx, ny, nm)
    api.five_point_Laplacian_stencil(s.dx, s.dx, host.mask_lap, nm)
    host.conc_Cr_old[...] = f0[0]
    host.conc_Nb_old[...] = f0[1]
    host.phi_del_old[...] = f0[2]
    host.phi_lav_old[...] = f0[3]
    D_Cr, D_Nb, Lmob = M.set_diffusion_matrix(0.30, 0.02)
    api.init_cuda(host, nx, ny, nm, M.kappa(), M.omega(), Lmob, D_Cr, D_Nb, dev)
    api.device_init_prng(dev, nx, ny, nm, bx, by, seed=42)
    api.device_set_steps_per_launch(dev, spl)
    steps, nuc_interval, io_interval = 200, 50, 64
    for t in range(steps):
        api.device_boundaries(dev, nx, ny, nm, bx, by)
        api.device_fictitious(dev, nx, ny, nm, bx, by)
        api.fictitious_boundaries(dev, nx, ny, nm, bx, by)
        api.device_laplacian(dev, nx, ny, nm, bx, by, s.dx, s.dx)
        api.device_laplacian_boundaries(dev, nx, ny, nm, bx, by)
        api.device_evolution(dev, nx, ny, nm, bx, by, M.OMEGA_3PHASE, s.dt)
        if t % nuc_interval == 0:
            api.device_nucleation(dev, nx, ny, nm, bx, by, M.SIGMA[0], M.SIGMA[1],
                                  M.LATTICE_CONST, M.IFCE_WIDTH, s.dx, s.dx, s.dx,
                                  nuc_interval * s.dt)
        dev.conc_Cr_old, dev.conc_Cr_new = api.swap_pointers_1D(dev.conc_Cr_old, dev.conc_Cr_new)
        dev.conc_Nb_old, dev.conc_Nb_new = api.swap_pointers_1D(dev.conc_Nb_old, dev.conc_Nb_new)
        dev.phi_del_old, dev.phi_del_new = api.swap_pointers_1D(dev.phi_del_old, dev.phi_del_new)
        dev.phi_lav_old, dev.phi_lav_new = api.swap_pointers_1D(dev.phi_lav_old, dev.phi_lav_new)
        if (t + 1) % io_interval == 0:
            api.device_dataviz(dev, host, nx, ny, nm, bx, by)
            api.read_out_result(dev, host, nx, ny)
    api.read_out_result(dev, host, nx, ny)
    summary = api.summarize_fields(dev, nx, ny)
    assert api.nucleation_status(dev) == (0, False)
    api.free_cuda(dev)
    want = oracle_reference(oracle, f0, s, steps)
    got = [host.conc_Cr_new, host.conc_Nb_new, host.phi_del_new, host.phi_lav_new]
    for k in range(4):
        assert mixed_err(got[k], want[k]) <= TOL, k
    ws = oracle.summarize(want)
    for a, b in zip(summary, ws):
        assert a == pytest.approx(b, rel=1e-12, abs=1e-15)


def _grouping_input(Nx, Ny, bc, seed):
    s = M.make_setup(Nx, Ny, bc=bc)
    f0 = M.init_matrix_with_particles(Nx, Ny, s.dx, bc, seed=seed, radius=2e-9)
    return s, f0


def test_cpp_grouping_edge_cases(gpu):
    """tests/cpp/grouping_main: a dt change inside a group, a step whose
    pointers were not swapped, a nucleation pass with events, a direct device
    read after device_flush; grouped == one launch per call, bit for bit."""
    exe = os.path.join(ROOT, "tests", "cpp", "grouping_main")
    assert os.path.exists(exe), "tests/cpp/grouping_main not built: run __graft_entry__.build()"
    Nx, Ny = 90, 64
    s, f0 = _grouping_input(Nx, Ny, 0, 12)
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        with open(fin, "wb") as fh:
            fh.write(struct.pack("ii", Nx + 2, Ny + 2))
            for a in f0:
                fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        r = subprocess.run([exe, fin, repr(s.dt), repr(s.dx)], capture_output=True, text=True,
                           timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "grouping ok" in r.stdout


def _python_grouping_sequence(spl, s, f0):
    import torch
    Nx, Ny = s.N
    nx, ny, nm, bx, by = Nx + 2, Ny + 2, 3, 13, 13
    host, dev = api.HostData(), api.CudaData()
    api.make_arrays(host, nx, ny, nm)
    api.five_point_Laplacian_stencil(s.dx, s.dx, host.mask_lap, nm)
    for name, a in zip(("conc_Cr_old", "conc_Nb_old", "phi_del_old", "phi_lav_old"), f0):
        getattr(host, name)[...] = a
    D_Cr, D_Nb, Lmob = M.set_diffusion_matrix(0.30, 0.02)
    api.init_cuda(host, nx, ny, nm, M.kappa(), M.omega(), Lmob, D_Cr, D_Nb, dev)
    api.device_init_prng(dev, nx, ny, nm, bx, by, seed=7)
    api.device_set_steps_per_launch(dev, spl)

    def swap():
        dev.conc_Cr_old, dev.conc_Cr_new = api.swap_pointers_1D(dev.conc_Cr_old, dev.conc_Cr_new)
        dev.conc_Nb_old, dev.conc_Nb_new = api.swap_pointers_1D(dev.conc_Nb_old, dev.conc_Nb_new)
        dev.phi_del_old, dev.phi_del_new = api.swap_pointers_1D(dev.phi_del_old, dev.phi_del_new)
        dev.phi_lav_old, dev.phi_lav_new = api.swap_pointers_1D(dev.phi_lav_old, dev.phi_lav_new)

    def step(h, do_swap=True):
        api.device_evolution(dev, nx, ny, nm, bx, by, M.OMEGA_3PHASE, h)
        if do_swap:
            swap()

    dt = s.dt
    step(dt)
    step(dt)
    step(0.5 * dt)              # dt changes inside the group
    step(0.5 * dt)
    step(0.5 * dt)
    step(dt, do_swap=False)     # the chain breaks: old -> new twice
    step(dt)
    step(dt)
    step(dt)
    api.device_evolution(dev, nx, ny, nm, bx, by, M.OMEGA_3PHASE, dt)
    api.device_nucleation(dev, nx, ny, nm, bx, by, 0.0105, 0.0108, M.LATTICE_CONST,
                          M.IFCE_WIDTH, s.dx, s.dx, s.dx, 3e38)
    swap()
    step(dt)
    step(dt)
    api.device_flush(dev)
    torch.cuda.synchronize()
    mid = dev.conc_Cr_old.cpu().numpy().copy()
    for _ in range(23):         # two groups of eight and a tail of seven at 4
        step(dt)
    api.read_out_result(dev, host, nx, ny)
    out = [host.conc_Cr_new.copy(), host.conc_Nb_new.copy(), host.phi_del_new.copy(),
           host.phi_lav_new.copy()]
    api.free_cuda(dev)
    return mid, out


def test_python_grouping_edge_cases(gpu):
    """The same sequence through the Python mirror (pfpa.api)."""
    s, f0 = _grouping_input(90, 64, 0, 12)
    m1, o1 = _python_grouping_sequence(1, s, f0)
    for spl in (4, 3):
        m, o = _python_grouping_sequence(spl, s, f0)
        assert np.array_equal(m, m1), spl
        for a, b in zip(o, o1):
            assert np.array_equal(a, b), spl
