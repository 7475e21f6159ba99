"""Pins for the FP64 oracle (oracle/pf_oracle.py) that do not retype its formulas.

Each check uses an independent route to the same number: autograd or finite
differences of the literal reference energy, closed forms, invariants of the
discretisation, brute-force scalar recursions, or the discrete energy functional.
"""
import itertools

import numpy as np
import pytest
import torch

from oracle import pf_oracle as O
from problems import args, gaussian, poly, quench, random_eta, rosenthal

H = (1.0, 1.0, 1.0)


def _energy_torch(eta, zeta, m_g, gamma):
    """allen_cahn.py:42-49 written with torch ops (vmap(np.outer) -> batched outer product)."""
    outer = eta[:, :, None] * eta[:, None, :]
    e1 = torch.sum(eta**4 / 4. - eta**2 / 2.)
    e2 = gamma * (torch.sum(outer**2) - torch.sum(eta**4))
    e3 = torch.sum((1 - zeta.reshape(-1))**2 * torch.sum(eta**2, dim=1).reshape(-1))
    return m_g * (e1 + e2 + e3)


@pytest.mark.parametrize("gamma", [1.0, 0.75])
def test_local_grad_equals_autograd_of_reference_energy(gamma):
    rng = np.random.default_rng(1)
    eta = rng.uniform(-0.2, 1.2, size=(50, 7))
    zeta = rng.choice([0.0, 0.5, 1.0], size=(50, 1))
    P = dict(m_g=2.4, gamma=gamma)
    e = torch.tensor(eta, requires_grad=True)
    (g,) = torch.autograd.grad(_energy_torch(e, torch.tensor(zeta), 2.4, gamma), e)
    got = O.local_energy_grad(eta, zeta, P)
    assert np.max(np.abs(got - g.numpy())) <= 1e-13 * np.max(np.abs(g.numpy()))
    # the oracle's literal energy is the same function
    assert O.local_energy(eta, zeta, P) == pytest.approx(float(_energy_torch(e, torch.tensor(zeta), 2.4, gamma)),
                                                         rel=1e-14)


def test_local_grad_equals_finite_differences():
    rng = np.random.default_rng(2)
    eta = rng.uniform(0, 1, size=(6, 5))
    zeta = rng.uniform(0, 1, size=(6, 1))
    P = dict(m_g=1.3, gamma=0.8)
    g = O.local_energy_grad(eta, zeta, P)
    d = 1e-6
    for c, i in itertools.product(range(6), range(5)):
        ep, em = eta.copy(), eta.copy()
        ep[c, i] += d
        em[c, i] -= d
        fd = (O.local_energy(ep, zeta, P) - O.local_energy(em, zeta, P)) / (2 * d)
        assert fd == pytest.approx(g[c, i], rel=1e-7, abs=1e-8)


def test_local_grad_fixed_points():
    P = dict(m_g=3.0, gamma=0.75)
    onehot = np.eye(4)
    assert np.all(O.local_energy_grad(onehot, np.ones((4, 1)), P) == 0.0)      # solid grain: minimum
    assert np.all(O.local_energy_grad(np.zeros((3, 4)), np.zeros((3, 1)), P) == 0.0)
    g = O.local_energy_grad(onehot, np.zeros((4, 1)), P)                       # melted grain: 2 m_g
    assert np.allclose(np.diag(g), 2 * 3.0) and np.all(g[~np.eye(4, dtype=bool)] == 0)


def test_zeta_is_a_step_at_liquidus_and_T_is_clipped():
    P = dict(T_liquidus=1623.0)
    T = np.array([[300.0], [1622.99], [1623.0], [1623.01], [1999.0], [2500.0]])
    Tc, z = O.zeta_of_T(T, P)
    assert list(z.reshape(-1)) == [1.0, 1.0, 0.5, 0.0, 0.0, 0.0]
    assert Tc[-1, 0] == 2000.0 and Tc[-2, 0] == 1999.0


def _rhs_parts(eta, a, pc, T):
    return O.state_rhs(eta, T, a, pc.mesh_h_xyz, pc.unique_grain_directions)


def test_laplacian_closed_forms_and_neumann_sum():
    Nx, Ny, Nz, K = 9, 6, 5, 2
    h = (0.5, 1.0, 2.0)
    a = args(Nx, Ny, Nz, K, m_g=0.0, kappa_g=1.7, ad_hoc=1.3)
    pc = poly(Nx, Ny, Nz, K, h=h)
    T = np.full((Nx * Ny * Nz, 1), 1500.0)
    L = a["L0"] * np.exp(-a["Qg"] / (1500.0 * a["gas_const"]))
    # eta = x^2 (x = cell index * h_x): interior Laplacian = 2
    x = (np.arange(Nx) * h[0])
    eta = np.broadcast_to((x**2)[None, None, :, None], (Nz, Ny, Nx, K)).reshape(-1, K).copy()
    rhs = _rhs_parts(eta, a, pc, T).reshape(Nz, Ny, Nx, K)
    lap = rhs / (L * a["kappa_g"] * a["ad_hoc"])
    assert np.allclose(lap[:, :, 1:-1, :], 2.0, rtol=1e-12)
    # edge replication at x = 0: (eta_1 - eta_0) / h^2
    assert np.allclose(lap[:, :, 0, :], (x[1]**2 - x[0]**2) / h[0]**2, rtol=1e-12)
    # constant field: zero Laplacian exactly
    assert np.all(_rhs_parts(np.full_like(eta, 0.37), a, pc, T) == 0.0)
    # zero-flux boundaries: the Laplacian sums to zero over the domain
    rnd = random_eta(Nx * Ny * Nz, K, seed=3)
    lap = _rhs_parts(rnd, a, pc, T) / (L * a["kappa_g"] * a["ad_hoc"])
    assert abs(np.sum(lap * h[0] * h[1] * h[2], axis=0)).max() <= 1e-12 * np.abs(lap).max() * lap.shape[0]


def _aniso_closed_form(edges, dirs, eps):
    e = edges.reshape(-1, edges.shape[-2], 3)
    dots = np.einsum('kid,nid->nki', dirs, e)
    n2 = np.sum(e * e, axis=-1)
    with np.errstate(invalid='ignore', divide='ignore'):
        q = np.max(dots**2, axis=1) / n2
    q = np.where(np.isfinite(q), np.minimum(q, 1.0), 1.0)
    return (1 + eps * (q**2 + (1 - q)**2)).reshape(edges.shape[:-1])


def test_anisotropy_special_directions_and_closed_form():
    pc = poly(4, 4, 4, 6, seed=11)
    dirs = pc.unique_grain_directions          # (3, K, 3)
    eps = 0.15
    K = dirs.shape[1]
    along = np.stack([dirs[1, i] * 0.3 for i in range(K)])[None]                       # parallel to a cube axis
    diag = np.stack([(dirs[0, i] + dirs[2, i]) * 0.2 for i in range(K)])[None]        # 45 deg in a cube face
    zero = np.zeros((1, K, 3))
    assert np.allclose(O.anisotropy_factor(along, dirs, eps), 1 + eps, rtol=1e-12)
    assert np.allclose(O.anisotropy_factor(diag, dirs, eps), 1 + eps / 2, rtol=1e-12)
    assert np.all(O.anisotropy_factor(zero, dirs, eps) == 1 + eps)                    # NaN -> angle 0
    rnd = np.random.default_rng(4).normal(size=(200, K, 3))
    assert np.allclose(O.anisotropy_factor(rnd, dirs, eps), _aniso_closed_form(rnd, dirs, eps), rtol=1e-12, atol=0)


def test_anisotropy_invariant_under_cube_symmetry():
    pc = poly(4, 4, 4, 3, seed=12)
    dirs = pc.unique_grain_directions
    rnd = np.random.default_rng(5).normal(size=(100, 3, 3))
    base = O.anisotropy_factor(rnd, dirs, 0.2)
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            d2 = np.stack([signs[j] * dirs[perm[j]] for j in range(3)])   # the same cube, relabelled axes
            assert np.allclose(O.anisotropy_factor(rnd, d2, 0.2), base, rtol=1e-12)


def test_uniform_single_grain_is_a_fixed_point():
    Nx, Ny, Nz, K = 6, 5, 4, 3
    for eps in (0.0, 0.2):
        a = args(Nx, Ny, Nz, K, anisotropy=eps)
        pc = poly(Nx, Ny, Nz, K)
        eta = np.zeros((Nx * Ny * Nz, K))
        eta[:, 1] = 1.0
        T = np.full((eta.shape[0], 1), 1500.0)
        assert np.all(O.explicit_euler(eta, 0.0, 0.05, T, a, pc.mesh_h_xyz, pc.unique_grain_directions) == eta)


def test_uniform_scalar_ode_matches_brute_force_recursion():
    Nx, Ny, Nz, K = 3, 3, 3, 1
    a = args(Nx, Ny, Nz, K, m_g=1.7)
    pc = poly(Nx, Ny, Nz, K, grains=2)
    for Tval, zeta in ((1500.0, 1.0), (1800.0, 0.0)):
        T = np.full((27, 1), Tval)
        L = a["L0"] * np.exp(-a["Qg"] / (Tval * a["gas_const"]))
        eta = np.full((27, 1), 0.3)
        y = 0.3
        ts = np.arange(0, 40) * 0.05
        for i in range(len(ts) - 1):
            eta = O.explicit_euler(eta, ts[i], ts[i + 1], T, a, pc.mesh_h_xyz)
            h = ts[i + 1] - ts[i]
            y = y - h * L * 1.7 * (y**3 - y + 2 * (1 - zeta)**2 * y)
        assert np.allclose(eta, y, rtol=1e-13)


def test_free_energy_gradient_is_the_isotropic_rhs():
    """d F / d eta_ci (central differences of the discrete functional) = -rhs_ci / L / dV."""
    Nx, Ny, Nz, K = 5, 4, 3, 3
    h = (1.0, 0.7, 1.3)
    a = args(Nx, Ny, Nz, K, m_g=0.9, kappa_g=1.4, ad_hoc=1.2)
    pc = poly(Nx, Ny, Nz, K, h=h)
    eta = random_eta(Nx * Ny * Nz, K, seed=6)
    T = np.full((eta.shape[0], 1), 1550.0)
    T[::3] = 1700.0
    L = a["L0"] * np.exp(-a["Qg"] / (np.minimum(T, 2000.0) * a["gas_const"]))
    rhs = O.state_rhs(eta, T, a, h)
    dV = h[0] * h[1] * h[2]
    d = 1e-6
    rng = np.random.default_rng(0)
    for _ in range(25):
        c, i = rng.integers(eta.shape[0]), rng.integers(K)
        ep, em = eta.copy(), eta.copy()
        ep[c, i] += d
        em[c, i] -= d
        fd = (sum(O.free_energy(ep, T, a, h)) - sum(O.free_energy(em, T, a, h))) / (2 * d)
        assert fd == pytest.approx(-rhs[c, i] / L[c, 0] * dV, rel=1e-6, abs=1e-9)


def test_euler_descends_the_free_energy_at_stable_dt():
    from pfkit.config import stable_dt
    Nx, Ny, Nz, K = 10, 8, 6, 4
    a = args(Nx, Ny, Nz, K)
    pc = poly(Nx, Ny, Nz, K)
    dt, _, _ = stable_dt(a, pc.mesh_h_xyz)
    eta = O.ini_cond(pc.cell_ori_inds, K) * 0.9 + 0.05 * random_eta(Nx * Ny * Nz, K, seed=1)
    T = np.full((eta.shape[0], 1), 1600.0)
    F = [sum(O.free_energy(eta, T, a, pc.mesh_h_xyz))]
    for i in range(30):
        eta = O.explicit_euler(eta, i * dt, (i + 1) * dt, T, a, pc.mesh_h_xyz)
        F.append(sum(O.free_energy(eta, T, a, pc.mesh_h_xyz)))
    assert all(F[i + 1] <= F[i] for i in range(30))


def test_orientation_permutation_equivariance():
    Nx, Ny, Nz, K = 6, 5, 4, 5
    a = args(Nx, Ny, Nz, K)
    pc = poly(Nx, Ny, Nz, K)
    eta = random_eta(Nx * Ny * Nz, K, seed=7)
    T = O.laser_T(gaussian(Nx, Ny, Nz).as_dict(), O.centroids(Nz, Ny, Nx, pc.mesh_h_xyz), 0.1)
    perm = np.array([3, 0, 4, 1, 2])
    r1 = O.state_rhs(eta, T, a, pc.mesh_h_xyz)[:, perm]
    r2 = O.state_rhs(eta[:, perm], T, a, pc.mesh_h_xyz)
    assert np.allclose(r1, r2, rtol=1e-13, atol=1e-300)


def test_mirror_symmetry_in_x():
    Nx, Ny, Nz, K = 7, 4, 3, 3
    a = args(Nx, Ny, Nz, K)
    pc = poly(Nx, Ny, Nz, K)
    eta = random_eta(Nx * Ny * Nz, K, seed=8)
    T = np.random.default_rng(1).uniform(1400, 1900, size=(Nx * Ny * Nz, 1))
    flip = lambda v: v.reshape(Nz, Ny, Nx, -1)[:, :, ::-1].reshape(-1, v.shape[1])
    r1 = flip(O.state_rhs(eta, T, a, pc.mesh_h_xyz))
    r2 = O.state_rhs(flip(eta), flip(T), a, pc.mesh_h_xyz)
    assert np.allclose(r1, r2, rtol=1e-13, atol=1e-300)


def test_laser_closed_forms():
    cent = np.array([[10.0, 5.0, 4.0], [4.0, 5.0, 4.0], [16.0, 5.0, 4.0], [10.0, 5.0, 1.0]])
    L = dict(model="rosenthal", pos0=(10.0, 5.0, 4.0), vel=0.5, Q=100.0, alpha=5.2, kappa_T=27.0, T_ambient=300.0)
    T = O.laser_T(L, cent[1:3], 0.0).reshape(-1)
    q = 100.0 / (2 * np.pi * 27.0)
    assert T[0] == pytest.approx(300.0 + q / 6.0, rel=1e-14)                                 # behind: R + X = 0
    assert T[1] == pytest.approx(300.0 + q / 6.0 * np.exp(-0.5 / 10.4 * 12.0), rel=1e-14)    # ahead: R + X = 2R
    # the source moves: at t = 12 it sits 6 um further along x, so (10, 5, 4) is 6 um behind it
    assert O.laser_T(L, cent[:1], 12.0)[0, 0] == pytest.approx(300.0 + q / 6.0, rel=1e-13)
    G = dict(model="gaussian", pos0=(10.0, 5.0, 4.0), vel=1.0, dT_peak=1900.0, sigma_xy=3.0, sigma_z=2.0,
             T_ambient=300.0)
    assert O.laser_T(G, cent[:1], 0.0)[0, 0] == 2200.0
    assert O.laser_T(G, cent[3:4], 0.0)[0, 0] == pytest.approx(300 + 1900 * np.exp(-9 / 8), rel=1e-14)
    Qd = dict(model="quench", T_ref=1623.0, grad_T=5e5, cool_rate=1e5)
    assert O.laser_T(Qd, cent[:1], 1e-5)[0, 0] == pytest.approx(1623.0 + 5e5 * 4.0 - 1.0, rel=1e-14)


def test_force_eta_resets_liquid_cells_only():
    a = args(2, 1, 1, 3, force_liquid=True)
    eta = np.array([[0.2, 0.9, 0.1], [0.3, 0.3, 0.4]])
    T = np.array([[1625.0], [1623.5]])   # T_l + 1 = 1624
    out = O.force_eta(eta, T, a)
    assert np.all(out[0] == 0.0) and np.array_equal(out[1], eta[1])


def test_ini_cond_and_inspect():
    eta0 = O.ini_cond(np.array([2, 0, 1, 2]), 3)
    assert np.array_equal(eta0, np.eye(3)[[2, 0, 1, 2]])
    eta = eta0.copy()
    eta[1] = [0.1, 0.8, 0.1]
    r = O.inspect(eta, eta0, np.array([[1000.0], [1700.0], [1623.0], [10.0]]), dict(T_liquidus=1623.0))
    assert r["change_pct"] == 25.0 and r["liquid_pct"] == 50.0 and r["max_T"] == 1700.0 and r["finite"]


def test_c_oracle_matches_numpy_oracle():
    from oracle import c_oracle as C
    if not C.available():
        pytest.skip("oracle/_build/libpforacle.so not built (make -C oracle)")
    for (Nx, Ny, Nz, K, eps, lk, fl) in ((11, 7, 5, 6, 0.15, "gauss", True), (9, 8, 6, 3, 0.0, "rosen", False),
                                         (6, 5, 7, 4, 0.2, "quench", True)):
        a = args(Nx, Ny, Nz, K, anisotropy=eps, force_liquid=fl)
        pc = poly(Nx, Ny, Nz, K)
        L = {"gauss": gaussian(Nx, Ny, Nz), "rosen": rosenthal(Nx, Ny, Nz), "quench": quench(Nz)}[lk].as_dict()
        eta = random_eta(Nx * Ny * Nz, K, seed=K)
        T = O.laser_T(L, O.centroids(Nz, Ny, Nx, pc.mesh_h_xyz), 0.2)
        ref = O.explicit_euler(eta, 0.2, 0.23, T, a, pc.mesh_h_xyz, pc.unique_grain_directions)
        out = np.empty_like(eta)
        C.step(eta, out, C.make_params(a, pc.mesh_h_xyz, L), pc.unique_grain_directions, 0.2, 0.23, threads=2)
        assert np.max(np.abs(out - ref)) <= 1e-13 * np.max(np.abs(ref))
        out2 = np.empty_like(eta)
        C.step(eta, out2, C.make_params(a, pc.mesh_h_xyz, None), pc.unique_grain_directions, 0.2, 0.23, T=T)
        assert np.max(np.abs(out - out2)) <= 1e-13 * np.max(np.abs(ref))   # libm vs numpy exp in T only
        e32 = eta.astype(np.float32)
        o32 = np.empty_like(e32)
        C.step(e32, o32, C.make_params(a, pc.mesh_h_xyz, L), pc.unique_grain_directions, 0.2, 0.23)
        assert np.max(np.abs(o32 - ref)) <= 1e-5 * np.max(np.abs(ref))
