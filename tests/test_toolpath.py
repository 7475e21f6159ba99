"""Multi-segment toolpath (pfkit.laser.Toolpath, multi_scan/example.py:52-105)."""
import numpy as np
import pytest

from oracle import pf_oracle as O
from pfkit.laser import Quench, Rosenthal, Toolpath
from problems import args, poly, random_eta


def test_schedule_by_hand():
    # corners (0,0,0) -> (0.9,0,0) -> (0.9,0.6,0); vel 1, dt 0.25
    tp = Toolpath((0.0, 0.9, 0.9), (0.0, 0.0, 0.6), (0.0, 0.0, 0.0), (1, 0, 0), vel=1.0, dt=0.25)
    ts, xs, ys, zs, ps, d = tp.schedule()
    # segment 1: arange(0, 0.9, 0.25) = 0, .25, .5, .75 -> x = linspace(0, .9, 4) = 0, .3, .6, .9
    # segment 2: t from 0.9: arange(0.9, 1.5, 0.25) = .9, 1.15, 1.4 -> y = linspace(0, .6, 3)
    np.testing.assert_allclose(ts, [0, .25, .5, .75, .9, 1.15, 1.4], atol=1e-15)
    np.testing.assert_allclose(xs, [0, .3, .6, .9, .9, .9, .9], atol=1e-15)
    np.testing.assert_allclose(ys, [0, 0, 0, 0, 0, .3, .6], atol=1e-15)
    np.testing.assert_array_equal(zs, np.zeros(7))
    np.testing.assert_array_equal(ps, [1, 1, 1, 1, 0, 0, 0])
    np.testing.assert_array_equal(d[:4], np.tile([1.0, 0, 0], (4, 1)))
    np.testing.assert_array_equal(d[4:], np.tile([0, 1.0, 0], (3, 1)))
    assert len(tp) == 7


def test_descriptor_places_the_source_at_the_scheduled_point():
    tp = Toolpath((1e-4, 9e-4, 9e-4), (1.7e-4, 1.7e-4, 0.7e-4), (1e-4, 1e-4, 1e-4), (1, 1, 0), vel=0.25, dt=2e-7)
    ts, xs, ys, zs, ps, d = tp.schedule()
    for i in (0, 17, len(ts) - 1):
        L = tp.laser(i)
        assert isinstance(L, Rosenthal) and L.T_cap == 2000.0
        pos = np.array(L.pos0) + L.vel * ts[i] * np.array(L.dir)
        np.testing.assert_allclose(pos, [xs[i], ys[i], zs[i]], rtol=0, atol=1e-18)
    off = Toolpath((0.0, 1.0, 1.0), (0.0, 0.0, 1.0), (0.0, 0.0, 0.0), (1, 0, 0), vel=1.0, dt=0.3, apply_switch=True)
    assert isinstance(off.laser(len(off) - 1), Quench) and isinstance(off.laser(0), Rosenthal)


def _oracle_toolpath(eta0, tp, a, pc, n):
    ts, xs, ys, zs, ps, d = tp.schedule()
    cent = O.centroids(a["Nz"], a["Ny"], a["Nx"], pc.mesh_h_xyz)
    dirs = pc.unique_grain_directions if a["anisotropy"] > 0 else None
    eta = eta0
    for i in range(n):
        # the reference's get_T_laser with the explicit source position of step i
        L = dict(model="rosenthal", pos0=(xs[i], ys[i], zs[i]), dir=tuple(d[i]), vel=tp.vel, Q=tp.Q,
                 alpha=tp.alpha, kappa_T=tp.kappa_T, T_ambient=tp.T_ambient, T_cap=tp.T_cap)
        T = O.laser_T(L, cent, 0.0)
        eta = O.explicit_euler(eta, ts[i], ts[i + 1], T, a, pc.mesh_h_xyz, dirs)
    return eta


def _problem():
    Nx, Ny, Nz, K = 18, 12, 8, 6
    a = args(Nx, Ny, Nz, K, anisotropy=0.15, force_liquid=True)
    pc = poly(Nx, Ny, Nz, K)
    # an L-shaped scan over the top surface (micrometres, microseconds)
    tp = Toolpath((3.0, 14.0, 14.0), (3.0, 3.0, 9.0), (8.0, 8.0, 8.0), (1, 1, 0), vel=2.0, dt=0.4, Q=1.6e6,
                  alpha=5.2, kappa_T=27.0)
    eta0 = random_eta(Nx * Ny * Nz, K, seed=11)
    return a, pc, tp, eta0


@pytest.mark.gpu
def test_run_toolpath_matches_oracle_and_stepper(cuda):
    import torch
    from pfkit import PFSolver
    a, pc, tp, eta0 = _problem()
    n = len(tp) - 1
    ref = _oracle_toolpath(eta0, tp, a, pc, n)
    s = PFSolver(a, pc, dtype=torch.float64, device=cuda, clean=False)
    got = s.run_toolpath(eta0, tp).cpu().numpy()
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) <= 1e-12
    # the reference's own loop: stepper with the descriptor of step i as ode_params
    ts = tp.schedule()[0]
    state = (s.state_rhs.check_state(eta0), ts[0])
    for i in range(n):
        state, sol = s.stepper(state, ts[i + 1], [tp.laser(i)])
    assert torch.equal(sol, torch.as_tensor(got, device=sol.device))


@pytest.mark.gpu
def test_batched_toolpath_equals_run_toolpath(cuda):
    """A toolpath instance in the batch engine (per-step source table) next to an ordinary
    instance: bitwise the result of PFSolver.run_toolpath, for FP64, FP32 and BF16 storage."""
    import torch
    from pfkit import PFSolver, workloads
    from pfkit.batch import BatchEngine, Instance
    a, pc, tp, eta0 = _problem()
    n = len(tp) - 1
    ts = tp.schedule()[0]
    other = workloads.make("other", 20, 16, 12, 10, 12, 9, "float32", 0.15, seed=5, vel=2.0)
    for dtype in ("float64", "float32", "bfloat16"):
        s = PFSolver(a, pc, dtype=dtype, device=cuda, clean=False)
        solo = s.run_toolpath(eta0, tp)
        eng = BatchEngine(max_instances=4)
        slot = eng.add(Instance(pf_args=a, h_xyz=pc.mesh_h_xyz, grain_dirs=pc.unique_grain_directions, laser=None,
                                dtype=dtype, steps=n, dt=tp.dt, eta0=eta0, toolpath=tp))
        slot2 = eng.add(other.instance())
        eng.run_to_completion(chunk=4)
        steps_left, t, par = eng.status()
        assert steps_left[slot] == 0 and steps_left[slot2] == 0
        assert t[slot] == ts[n]
        assert torch.equal(eng.state(slot, par[slot]), solo), dtype
        eng.close()
