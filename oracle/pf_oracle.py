"""CPU FP64 oracle of the phase-field Allen-Cahn explicit step.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP path: only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import it.  The product path (``pfkit``) never calls it.

It restates, in NumPy FP64 and in the reference's own order of operations, the
phase-field path of JAX-AM (reference files cited per function; the reference
runs with jax x64 enabled, see applications/phase_field/simple/example.py:10 ->
jax_am/cfd/cfd_am.py:10).  The one deliberate difference: the reference obtains the
local free-energy gradient with ``jax.grad`` (allen_cahn.py:51); here it is the
closed form, which tests/test_oracle.py pins against autograd and finite
differences of the literal energy ``local_energy``.

Parameter dict ``P`` uses the reference's yaml keys (Nx, Ny, Nz, num_oris,
m_g, kappa_g, ad_hoc, L0, Qg, gas_const, T_liquidus, anisotropy) plus the
constants the reference hard-codes, made explicit: gamma (allen_cahn.py:43, 1),
T_clip (:56, 2000), zeta_sharpness (:57, 1e10), force_liquid (:142, off) and
liquid_shift (:124, 1.0).
"""
import numpy as np

DEFAULTS = dict(gamma=1.0, T_clip=2000.0, zeta_sharpness=1e10, force_liquid=False, liquid_shift=1.0,
                anisotropy=0.0, ad_hoc=1.0)


def _p(P, key):
    return P[key] if key in P else DEFAULTS[key]


# --------------------------------------------------------------------- local energy
def local_energy(eta, zeta, P):
    """Literal restatement of allen_cahn.py:42-49 (sum over all cells).

    eta: (N, K); zeta: (N, 1).  The (N, K, K) outer products are formed exactly as
    in the reference (vmap(np.outer)); use on small inputs only.
    """
    gamma = _p(P, 'gamma')
    outer = eta[:, :, None] * eta[:, None, :]
    e1 = np.sum(eta**4 / 4. - eta**2 / 2.)
    e2 = gamma * (np.sum(outer**2) - np.sum(eta**4))
    e3 = np.sum((1 - zeta.reshape(-1))**2 * np.sum(eta**2, axis=1).reshape(-1))
    return P['m_g'] * (e1 + e2 + e3)


def local_energy_grad(eta, zeta, P):
    """d local_energy / d eta in closed form (replaces jax.grad at allen_cahn.py:51):
    m_g * (eta^3 - eta + 4 gamma eta (S - eta^2) + 2 (1 - zeta)^2 eta), S = sum_j eta_j^2."""
    gamma = _p(P, 'gamma')
    S = np.sum(eta**2, axis=1, keepdims=True)
    z = zeta.reshape(-1, 1)
    return P['m_g'] * (eta**3 - eta + 4. * gamma * eta * (S - eta**2) + 2. * (1. - z)**2 * eta)


def zeta_of_T(T, P):
    """allen_cahn.py:56-57: clip T at T_clip, zeta = 0.5*(1 - tanh(theta*(T/T_l - 1)))."""
    Tc = np.where(T > _p(P, 'T_clip'), _p(P, 'T_clip'), T)
    return Tc, 0.5 * (1 - np.tanh(_p(P, 'zeta_sharpness') * (Tc / P['T_liquidus'] - 1)))


# --------------------------------------------------------------------- anisotropy
def anisotropy_factor(edges, grain_dirs, eps):
    """Restatement of update_anisotropy_helper, allen_cahn.py:28-39 (arccos path).

    edges: (..., K, 3) central differences (eta_+ - eta_-) per axis;
    grain_dirs: (3, K, 3) rotated cube axes (utils.py:104-109).  Returns (..., K).
    """
    K = edges.shape[-2]
    e = edges.reshape(-1, K, 3)
    with np.errstate(invalid='ignore', divide='ignore'):
        cosines = np.sum(grain_dirs[None, :, :, :] * e[:, None, :, :], axis=-1) \
            / (np.linalg.norm(e, axis=-1)[:, None, :])
        angles = np.arccos(cosines)
    angles = np.where(np.isfinite(angles), angles, 0.)
    angles = np.where(angles < np.pi / 2., angles, np.pi - angles)
    angles = np.min(angles, axis=1)
    a = 1. + eps * (np.cos(angles)**4 + np.sin(angles)**4)
    return a.reshape(edges.shape[:-1])


# --------------------------------------------------------------------- right-hand side
def neighbours(eta, P):
    """Edge-replicated shifts, allen_cahn.py:60-66.  Returns (c, -x, +x, -y, +y, -z, +z), each (Nz,Ny,Nx,K)."""
    c = np.reshape(eta, (P['Nz'], P['Ny'], P['Nx'], P['num_oris']))
    nx_ = np.concatenate((c[:, :, :1, :], c[:, :, :-1, :]), axis=2)
    px_ = np.concatenate((c[:, :, 1:, :], c[:, :, -1:, :]), axis=2)
    ny_ = np.concatenate((c[:, :1, :, :], c[:, :-1, :, :]), axis=1)
    py_ = np.concatenate((c[:, 1:, :, :], c[:, -1:, :, :]), axis=1)
    nz_ = np.concatenate((c[:1, :, :, :], c[:-1, :, :, :]), axis=0)
    pz_ = np.concatenate((c[1:, :, :, :], c[-1:, :, :, :]), axis=0)
    return c, nx_, px_, ny_, py_, nz_, pz_


def state_rhs(eta, T, P, h_xyz, grain_dirs=None):
    """Restatement of state_rhs, allen_cahn.py:53-85.  eta (N, K), T (N, 1) -> rhs (N, K)."""
    h_x, h_y, h_z = h_xyz
    K = P['num_oris']
    T, zeta = zeta_of_T(T, P)
    grad = local_energy_grad(eta, zeta, P) / P['ad_hoc']
    c, nx_, px_, ny_, py_, nz_, pz_ = neighbours(eta, P)
    eps = _p(P, 'anisotropy')
    if eps > 0.:
        aniso = anisotropy_factor(np.stack((px_ - nx_, py_ - ny_, pz_ - nz_), axis=-1), grain_dirs, eps)[..., None]
    else:
        aniso = 1.
    laplace_xyz = -(np.stack((((px_ - c) + (nx_ - c)) / h_x**2,
                              ((py_ - c) + (ny_ - c)) / h_y**2,
                              ((pz_ - c) + (nz_ - c)) / h_z**2), axis=-1)
                    * aniso * P['kappa_g'] * P['ad_hoc'])
    laplace = np.sum(laplace_xyz.reshape(-1, K, 3), axis=-1)
    Lg = P['L0'] * np.exp(-P['Qg'] / (T * P['gas_const']))
    return -Lg * (grad + laplace)


def force_eta(eta, T, P):
    """get_force_eta_fn, allen_cahn.py:117-128: eta <- 0 where T > T_l + shift (unclipped T)."""
    liquid = T > P['T_liquidus'] + _p(P, 'liquid_shift')
    return np.where(liquid, 0., eta)


def explicit_euler(eta, t_prev, t_crt, T, P, h_xyz, grain_dirs=None):
    """explicit_euler, allen_cahn.py:106-114 (+ the liquid reset when P['force_liquid'])."""
    h = t_crt - t_prev
    y = eta + h * state_rhs(eta, T, P, h_xyz, grain_dirs)
    if _p(P, 'force_liquid'):
        y = force_eta(y, T, P)
    return y


def rk4(eta, t_prev, t_crt, T, P, h_xyz, grain_dirs=None):
    """rk4, allen_cahn.py:88-101: classical Runge-Kutta with the temperature T held fixed over the
    step (the reference passes T through ode_params); no liquid reset."""
    h = t_crt - t_prev
    f = lambda y: state_rhs(y, T, P, h_xyz, grain_dirs)  # noqa: E731
    k1 = h * f(eta)
    k2 = h * f(eta + k1 / 2.)
    k3 = h * f(eta + k2 / 2.)
    k4 = h * f(eta + k3)
    return eta + 1. / 6 * (k1 + 2 * k2 + 2 * k3 + k4)


# --------------------------------------------------------------------- temperature
def centroids(Nz, Ny, Nx, h_xyz):
    """Cell centroids (N, 3) in (x, y, z), cells in C order (z, y, x), x fastest (utils.py:43-57)."""
    z, y, x = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing='ij')
    return np.stack(((x.reshape(-1) + 0.5) * h_xyz[0], (y.reshape(-1) + 0.5) * h_xyz[1],
                     (z.reshape(-1) + 0.5) * h_xyz[2]), axis=1)


def laser_T(L, cent, t):
    """Temperature (N, 1) of a laser descriptor dict at time t.

    rosenthal: simple/example.py:21-37 (point source at pos0 + vel*t*dir; multi_scan/example.py:41-47
    for a general direction; T_cap > 0 caps T as multi_scan does at :47); gaussian: moving Gaussian
    spot; quench: solidification/example.py:27-32.
    """
    model = L['model']
    if model == 'quench':
        T = L['T_ref'] + L['grad_T'] * cent[:, 2] - L['cool_rate'] * t
        return T[:, None]
    d = L.get('dir', (1.0, 0.0, 0.0))
    p0 = L['pos0']
    vt = L['vel'] * t
    X = cent[:, 0] - p0[0] - vt * d[0]
    Y = cent[:, 1] - p0[1] - vt * d[1]
    Z = cent[:, 2] - p0[2] - vt * d[2]
    if model == 'rosenthal':
        R = np.sqrt(X**2 + Y**2 + Z**2)
        proj = X * d[0] + Y * d[1] + Z * d[2]
        T = L['T_ambient'] + L['Q'] / (2 * np.pi * L['kappa_T']) / R * np.exp(-L['vel'] / (2 * L['alpha']) * (R + proj))
        if L.get('T_cap', 0.0) > 0.0:
            T = np.where(T > L['T_cap'], L['T_cap'], T)
        return T[:, None]
    if model == 'gaussian':
        g_sxy = 1.0 / (2.0 * L['sigma_xy'] * L['sigma_xy'])
        g_sz = 1.0 / (2.0 * L['sigma_z'] * L['sigma_z'])
        T = L['T_ambient'] + L['dT_peak'] * np.exp(-((X * X + Y * Y) * g_sxy + Z * Z * g_sz))
        return T[:, None]
    raise ValueError(f"unknown laser model {model!r}")


# --------------------------------------------------------------------- initial condition, inspection
def ini_cond(cell_ori_inds, K):
    """PFSolver.ini_cond, allen_cahn.py:145-153: one-hot eta."""
    n = len(cell_ori_inds)
    eta = np.zeros((n, K))
    eta[np.arange(n), cell_ori_inds] = 1.
    return eta


def inspect(eta, eta0, T, P):
    """The quantities printed by PFSolver.inspect_sol (allen_cahn.py:165-180)."""
    a0 = np.argmax(eta0, axis=1)
    a = np.argmax(eta, axis=1)
    return dict(change_pct=np.sum(a0 != a) / len(a) * 100.,
                liquid_pct=np.sum(T.reshape(-1) >= P['T_liquidus']) / len(a) * 100.,
                max_T=float(np.max(T)), finite=bool(np.all(np.isfinite(eta))))


def free_energy(eta, T, P, h_xyz):
    """Discrete free energy whose eta-gradient is the isotropic right-hand side / (-L):
    F_local = dV sum_c f_loc(eta_c, zeta_c) / ad_hoc (f_loc: allen_cahn.py:42-49),
    F_grad  = dV kappa_g ad_hoc / 2 sum_faces sum_i ((eta_i' - eta_i) / h)^2 (edge-replicated faces
    contribute 0).  Returns (F_local, F_grad)."""
    K = P['num_oris']
    gamma = _p(P, 'gamma')
    _, zeta = zeta_of_T(T, P)
    S = np.sum(eta**2, axis=1)
    s4 = np.sum(eta**4, axis=1)
    q = np.sum(eta**4 / 4. - eta**2 / 2., axis=1)
    f = P['m_g'] * (q + gamma * (S * S - s4) + (1. - zeta.reshape(-1))**2 * S) / P['ad_hoc']
    dV = h_xyz[0] * h_xyz[1] * h_xyz[2]
    c = eta.reshape(P['Nz'], P['Ny'], P['Nx'], K)
    g = (np.sum((c[:, :, 1:] - c[:, :, :-1])**2) / h_xyz[0]**2 +
         np.sum((c[:, 1:] - c[:, :-1])**2) / h_xyz[1]**2 +
         np.sum((c[1:] - c[:-1])**2) / h_xyz[2]**2)
    return dV * np.sum(f), dV * 0.5 * P['kappa_g'] * P['ad_hoc'] * g


def volumes(eta, h_xyz):
    """Per-orientation volume dV * sum_c eta_ci."""
    return h_xyz[0] * h_xyz[1] * h_xyz[2] * np.sum(eta, axis=0)


def run(eta0, ts, P, h_xyz, grain_dirs=None, T_fn=None, laser=None, cent=None):
    """Reference time loop (simple/example.py:52-67): for each step i, T = T(ts[i]) and
    eta <- explicit_euler(eta, ts[i], ts[i+1], T).  T from T_fn(t) or a laser dict."""
    eta = eta0
    for i in range(len(ts) - 1):
        T = T_fn(ts[i]) if T_fn is not None else laser_T(laser, cent, ts[i])
        eta = explicit_euler(eta, ts[i], ts[i + 1], T, P, h_xyz, grain_dirs)
    return eta
