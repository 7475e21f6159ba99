"""rk4, CFD-coupling temperature, quaternion output, checkpoints, layer stacking, VTU output."""
import os

import numpy as np
import pytest

from oracle import pf_oracle as O
from problems import args, gaussian, poly, random_eta


def _smooth_problem():
    Nx, Ny, Nz, K = 6, 5, 4, 3
    a = args(Nx, Ny, Nz, K, anisotropy=0.1)
    pc = poly(Nx, Ny, Nz, K)
    z, y, x = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    base = 0.5 + 0.3 * np.sin(0.7 * x + 0.4 * y + 0.2 * z).reshape(-1)
    eta = np.stack([base, 0.8 - base, 0.3 + 0.1 * np.cos(0.5 * x).reshape(-1)], axis=1)
    T = np.full((Nx * Ny * Nz, 1), 1500.0)
    return a, pc, eta, T


def test_oracle_rk4_is_fourth_order():
    # isotropic: the right-hand side is a polynomial in eta (the anisotropy factor has kinks)
    a, pc, eta, T = _smooth_problem()
    a = dict(a, anisotropy=0.0)
    T = np.full_like(T, 1900.0)
    dirs = None
    H = 0.8

    def integrate(n):
        y = eta
        for i in range(n):
            y = O.rk4(y, i * H / n, (i + 1) * H / n, T, a, pc.mesh_h_xyz, dirs)
        return y
    ref = integrate(256)
    e1 = np.max(np.abs(integrate(8) - ref))
    e2 = np.max(np.abs(integrate(16) - ref))
    assert 13.0 < e1 / e2 < 24.0          # 2^4 = 16 (a third-order scheme would give ~8)
    # and the Euler step it generalises is first order
    y1 = O.explicit_euler(eta, 0.0, H / 8, T, a, pc.mesh_h_xyz, dirs)
    assert np.max(np.abs(y1 - O.rk4(eta, 0.0, H / 8, T, a, pc.mesh_h_xyz, dirs))) > 1e-8


def test_stack_layer_geometry():
    from pfkit.polycrystal import stack_layer
    lo = poly(8, 6, 4, 5, seed=1)
    hi = poly(8, 6, 3, 5, seed=2)
    hi.quat, hi.unique_grain_directions = lo.quat, lo.unique_grain_directions
    st = stack_layer(lo, hi)
    assert st.Nz == 7 and st.n_cells == 8 * 6 * 7
    np.testing.assert_array_equal(st.cell_ori_inds[:lo.n_cells], lo.cell_ori_inds)
    np.testing.assert_array_equal(st.cell_ori_inds[lo.n_cells:], hi.cell_ori_inds)
    # the stacked centroids are those of the 8 x 6 x 7 grid
    np.testing.assert_allclose(st.centroids, O.centroids(7, 6, 8, st.mesh_h_xyz))


def test_field_writes_quaternions(tmp_path):
    from pfkit.polycrystal import Field
    pf_args = dict(Nx=6, Ny=5, Nz=4, num_grains=7, num_oris=3, data_dir=str(tmp_path))
    f = Field(pf_args)
    q = np.load(os.path.join(tmp_path, "numpy/pf/quat.npy"))
    np.testing.assert_array_equal(q, f.quat)
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, rtol=1e-12)


@pytest.mark.gpu
def test_rk4_matches_oracle(cuda):
    import torch
    from pfkit import PFSolver, rk4
    a, pc, eta, T = _smooth_problem()
    s = PFSolver(a, pc, dtype=torch.float64, device=cuda, clean=False)
    (y, t), _ = rk4((eta, 0.1), 0.15, s.state_rhs, [T])
    ref = O.rk4(eta, 0.1, 0.15, T, a, pc.mesh_h_xyz, pc.unique_grain_directions)
    assert np.max(np.abs(y.cpu().numpy() - ref)) / np.max(np.abs(ref)) <= 1e-12 and t == 0.15


@pytest.mark.gpu
def test_cfd_temperature_interpolates_and_reorders(cuda):
    import torch
    from pfkit.coupling import cfd_temperature
    rng = np.random.default_rng(5)
    Nx, Ny, Nz = 7, 5, 3
    Tp, Tf = rng.uniform(300, 2000, (Nx, Ny, Nz)), rng.uniform(300, 2000, (Nx, Ny, Nz))
    t_pf, t_cfd, dt = 3.3e-6, 4.0e-6, 1.0e-6
    got = cfd_temperature(torch.as_tensor(Tp, device=cuda), torch.as_tensor(Tf, device=cuda), t_pf, t_cfd, dt)
    r = (t_cfd - t_pf) / dt
    ref = np.transpose(r * Tp + (1 - r) * Tf, axes=(2, 1, 0)).reshape(-1, 1)   # coupled/example.py:31-42
    np.testing.assert_allclose(got.cpu().numpy(), ref, rtol=1e-15)
    got32 = cfd_temperature(torch.as_tensor(Tp, dtype=torch.float32, device=cuda),
                            torch.as_tensor(Tf, dtype=torch.float32, device=cuda), t_pf, t_cfd, dt)
    np.testing.assert_allclose(got32.double().cpu().numpy(), ref, rtol=1e-6)


@pytest.mark.gpu
def test_checkpoint_roundtrip(cuda, tmp_path):
    import torch
    from pfkit import PFSolver
    a = args(9, 7, 5, 4, anisotropy=0.15)
    pc = poly(9, 7, 5, 4)
    for dt in (torch.float64, torch.float32, torch.bfloat16):
        s = PFSolver(a, pc, dtype=dt, device=cuda, clean=False)
        eta = s.state_rhs.check_state(random_eta(9 * 7 * 5, 4, seed=3))
        path = s.save_checkpoint(os.path.join(tmp_path, f"ck_{dt}.npz"), eta, 1.25, 17)
        e2, t, step = s.load_checkpoint(path)
        assert torch.equal(e2, eta) and e2.dtype == dt and t == 1.25 and step == 17


@pytest.mark.gpu
def test_stacked_layer_step_matches_oracle(cuda):
    import torch
    from pfkit import PFSolver
    from pfkit.polycrystal import stack_layer
    lo, hi = poly(10, 8, 4, 5, seed=1), poly(10, 8, 3, 5, seed=2)
    hi.quat, hi.unique_grain_directions = lo.quat, lo.unique_grain_directions
    st = stack_layer(lo, hi)
    a = args(10, 8, 7, 5, anisotropy=0.15, force_liquid=True)
    L = gaussian(10, 8, 7)
    s = PFSolver(a, st, dtype=torch.float64, device=cuda, clean=False)
    eta0 = s.ini_cond()
    got = s.run(eta0, np.array([0.2, 0.25, 0.3]), L).cpu().numpy()
    ref = O.run(O.ini_cond(st.cell_ori_inds, 5), np.array([0.2, 0.25, 0.3]), a, st.mesh_h_xyz,
                st.unique_grain_directions, laser=L.as_dict(), cent=st.centroids)
    assert np.max(np.abs(got - ref)) <= 1e-12


@pytest.mark.gpu
def test_write_sols_ipf_on_device(cuda, tmp_path):
    import torch
    from pfkit import PFSolver
    a = args(8, 6, 5, 4, anisotropy=0.15)
    a.update(data_dir=str(tmp_path), write_sol_interval=1)
    pc = poly(8, 6, 5, 4)
    s = PFSolver(a, pc, dtype=torch.float32, device=cuda, clean=False)
    L = gaussian(8, 6, 5)
    eta = s.ini_cond()
    path = s.write_sols(eta, L, 3, t=0.0)
    assert os.path.exists(path) and os.path.getsize(path) > 0
    from pfkit.vtu import read_cell_data
    cd = read_cell_data(path)
    inds = np.asarray(pc.cell_ori_inds)
    np.testing.assert_array_equal(cd["ori_inds"], inds)
    T = O.laser_T(L.as_dict(), O.centroids(5, 6, 8, pc.mesh_h_xyz), 0.0).reshape(-1)
    exp = np.take(np.asarray(pc.unique_oris_rgb)[0], inds, axis=0)
    exp[T > a["T_liquidus"]] = 0.0
    np.testing.assert_allclose(cd["ipf_x"], exp, rtol=1e-6)
