"""Parity at the full benchmark size: 256 x 256 x 64 grid, K = 20, cubic anisotropy 0.15, moving
Gaussian source, 500 grains (the throughput workload of bench.py).

The state after 30 FP64 steps from the one-hot start (interfaces diffused, a melt pool under the
source) is stepped by the fused HIP kernels and by the C/OpenMP FP64 oracle (oracle/pf_oracle.c),
and every tile shape and z-chunk of the FP32 step is compared bitwise with the generic kernel.
"""
import os

import numpy as np
import pytest
import torch

from oracle import c_oracle as C
from pfkit import PFSolver, _lib, workloads

pytestmark = pytest.mark.gpu

WARM = 30  # FP64 steps before the compared steps


def _threads():
    return min(16, len(os.sched_getaffinity(0)))


@pytest.fixture(scope="module")
def c1_state(cuda):
    wl = workloads.c1()
    s64 = PFSolver(wl.pf_args, wl.polycrystal, dtype=torch.float64, clean=False)
    ts = wl.ts(WARM + 3)
    eta = s64.run(s64.ini_cond(), ts[:WARM + 1], wl.laser)
    return wl, ts, eta


def _oracle_steps(wl, ts, eta_host, n0, n):
    params = C.make_params(wl.pf_args, wl.polycrystal.mesh_h_xyz, wl.laser.as_dict())
    a = np.ascontiguousarray(eta_host)
    b = np.empty_like(a)
    for i in range(n0, n0 + n):
        C.step(a, b, params, wl.polycrystal.unique_grain_directions, ts[i], ts[i + 1], threads=_threads())
        a, b = b, a
    return a


def test_c1_f64_steps_match_c_oracle(c1_state):
    if not C.available():
        pytest.skip("oracle/_build/libpforacle.so not built (make -C oracle)")
    wl, ts, eta = c1_state
    host = eta.cpu().numpy()
    # the developed state is not trivial: diffuse interfaces (~1.2e5 values strictly between 0 and
    # 1) and a melt pool of all-zero cells (~8e4) under the source
    assert np.count_nonzero((host > 1e-3) & (host < 1 - 1e-3)) > 5 * 10 ** 4
    assert np.count_nonzero(np.all(host == 0.0, axis=1)) > 10 ** 4
    s64 = PFSolver(wl.pf_args, wl.polycrystal, dtype=torch.float64, clean=False)
    got = s64.run(eta, ts[WARM:WARM + 3], wl.laser).cpu().numpy()
    ref = _oracle_steps(wl, ts, host, WARM, 2)
    assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_c1_f32_step_matches_c_oracle(c1_state):
    if not C.available():
        pytest.skip("oracle/_build/libpforacle.so not built (make -C oracle)")
    wl, ts, eta = c1_state
    e32 = eta.to(torch.float32)
    s32 = PFSolver(wl.pf_args, wl.polycrystal, dtype=torch.float32, clean=False)
    out = torch.empty_like(e32)
    s32.state_rhs.step(e32, out, wl.laser, ts[WARM], ts[WARM + 1], s32.tile, s32.zchunk)
    got = out.to(torch.float64).cpu().numpy()
    ref = _oracle_steps(wl, ts, e32.to(torch.float64).cpu().numpy(), WARM, 1)
    assert np.max(np.abs(got - ref)) <= 2e-6 * np.max(np.abs(ref)) + 1e-7


def test_c1_tile_shapes_and_zchunks_agree_bitwise(c1_state):
    wl, ts, eta = c1_state
    lib = _lib.load()
    for dtype in (torch.float32, torch.bfloat16):
        e = eta.to(dtype)
        s = PFSolver(wl.pf_args, wl.polycrystal, dtype=dtype, clean=False)

        def step(tile, zc):
            out = torch.empty_like(e)
            s.state_rhs.step(e, out, wl.laser, ts[WARM], ts[WARM + 1], tile, zc)
            return out

        base = step(-2, 0)  # generic kernel
        for tile in range(lib.pf_tile_count(_lib.TORCH_TO_PF[dtype], wl.K)):
            for zc in (0, 8, 12, 64, 10 | _lib.ZCHUNK_TAIL, 16 | _lib.ZCHUNK_TAIL):
                assert torch.equal(step(tile, zc), base), (dtype, tile, zc)


def test_c0_full_run_f64_matches_c_oracle(cuda):
    """The CPU-runnable configuration (64 x 64 x 32 grid, K = 10, 50 grains, isotropic, moving
    Gaussian source, 200 steps from the one-hot start): the whole FP64 trajectory on the GPU
    against the C/OpenMP FP64 oracle (relative L-inf <= 1e-10 after 200 steps), and the FP32
    run against it (max |d eta| <= 1e-5)."""
    if not C.available():
        pytest.skip("oracle/_build/libpforacle.so not built (make -C oracle)")
    wl = workloads.c0()
    ts = wl.ts()
    n = len(ts) - 1
    s64 = PFSolver(wl.pf_args, wl.polycrystal, dtype=torch.float64, clean=False)
    eta0 = s64.ini_cond()
    got = s64.run(eta0, ts, wl.laser).cpu().numpy()
    ref = _oracle_steps(wl, ts, eta0.cpu().numpy(), 0, n)
    assert n == 200
    assert np.count_nonzero((ref > 1e-3) & (ref < 1 - 1e-3)) > 1000   # interfaces have moved
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))
    s32 = PFSolver(wl.pf_args, wl.polycrystal, dtype=torch.float32, clean=False)
    got32 = s32.run(s32.ini_cond(), ts, wl.laser).to(torch.float64).cpu().numpy()
    assert np.max(np.abs(got32 - ref)) <= 1e-5
