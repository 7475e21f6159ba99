"""Writes the golden fixtures golden_*.npz of this directory (run from the repository root:
``python tests/golden/make_golden.py``).

The reference ships no golden vectors, known-answer tests or fixtures for its phase-field
path (its only fixtures are FEM files unrelated to it), and its JAX code cannot run here.
These fixtures therefore hold the FP64 oracle's output (oracle/pf_oracle.py) on small
seeded problems, after the oracle itself is pinned by the closed-form, invariant and
brute-force checks of tests/test_oracle.py.  Each file stores every input the step needs
(parameters, laser, grain directions, time grid, initial state) next to the expected
states, so the GPU tests compare against them without rebuilding the problem, and the CPU
tests detect any drift of the oracle.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "jax-am_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

from oracle import pf_oracle as O  # noqa: E402
from problems import args, gaussian, poly, quench, random_eta, rosenthal  # noqa: E402

# name: (Nx, Ny, Nz, K, anisotropy, laser, force_liquid, start, steps, dt)
CASES = {
    "iso_gauss_k6": (12, 10, 8, 6, 0.0, "gauss", True, "random", 5, 0.02),
    "aniso_rosen_k5": (9, 7, 6, 5, 0.15, "rosen", False, "random", 3, 0.02),
    "aniso_quench_k10_onehot": (16, 12, 10, 10, 0.15, "quench", True, "onehot", 10, 0.05),
    "aniso_gauss_k17": (11, 9, 7, 17, 0.3, "gauss", True, "random", 4, 0.02),
}


def build(name):
    Nx, Ny, Nz, K, eps, lk, fl, start, steps, dt = CASES[name]
    a = args(Nx, Ny, Nz, K, anisotropy=eps, force_liquid=fl)
    pc = poly(Nx, Ny, Nz, K)
    laser = {"gauss": gaussian(Nx, Ny, Nz), "rosen": rosenthal(Nx, Ny, Nz), "quench": quench(Nz)}[lk]
    N = Nx * Ny * Nz
    eta0 = O.ini_cond(pc.cell_ori_inds, K) if start == "onehot" else random_eta(N, K, seed=Nx * 100 + K)
    ts = np.arange(steps + 1) * dt + 0.1
    return a, pc, laser, eta0, ts


def expected(a, pc, laser, eta0, ts):
    cent = O.centroids(a["Nz"], a["Ny"], a["Nx"], pc.mesh_h_xyz)
    dirs = pc.unique_grain_directions if a["anisotropy"] > 0 else None
    eta1 = O.run(eta0, ts[:2], a, pc.mesh_h_xyz, dirs, laser=laser.as_dict(), cent=cent)
    etaN = O.run(eta0, ts, a, pc.mesh_h_xyz, dirs, laser=laser.as_dict(), cent=cent)
    T_end = O.laser_T(laser.as_dict(), cent, ts[-1])
    f_loc, f_grad = O.free_energy(etaN, T_end, a, pc.mesh_h_xyz)
    return dict(eta_step1=eta1, eta_final=etaN, T_final=T_end.reshape(-1),
                free_energy=np.array([f_loc, f_grad]), volumes=O.volumes(etaN, pc.mesh_h_xyz))


def main():
    for name in CASES:
        a, pc, laser, eta0, ts = build(name)
        exp = expected(a, pc, laser, eta0, ts)
        np.savez_compressed(os.path.join(HERE, f"golden_{name}.npz"),
                            pf_args=json.dumps({k: (float(v) if isinstance(v, (np.floating, float)) else v)
                                                for k, v in a.items()}, sort_keys=True),
                            laser=json.dumps(laser.as_dict()), h_xyz=np.array(pc.mesh_h_xyz),
                            grain_dirs=pc.unique_grain_directions, cell_ori_inds=pc.cell_ori_inds,
                            eta0=eta0, ts=ts, **exp)
        print(name, exp["eta_final"].shape, float(np.abs(exp["eta_final"] - eta0).max()))


if __name__ == "__main__":
    main()
