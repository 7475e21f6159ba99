"""Parameter handling: the reference's yaml dict, plus the constants it hard-codes.

The reference loads ``pf_params.yaml`` into a plain dict (jax_am/common.py:24-31) and
mutates it at run time (Nx/Ny/Nz from the mesh, utils.py:47-49; data_dir from the
app).  This module keeps that dict as the interface and fills in, under explicit
keys, what the reference hard-codes:

=================  =======  =========================================
key                default  reference
=================  =======  =========================================
gamma              1.0      allen_cahn.py:43 ``gamma = 1``
T_clip             2000.0   allen_cahn.py:56
zeta_sharpness     1e10     allen_cahn.py:57
force_liquid       False    allen_cahn.py:142 (call commented out)
liquid_shift       1.0      allen_cahn.py:124
anisotropy         0.0      pf_params.yaml
=================  =======  =========================================
"""
import math

import yaml

from . import _lib

HARD_CODED = dict(gamma=1.0, T_clip=2000.0, zeta_sharpness=1e10, force_liquid=False, liquid_shift=1.0)


def yaml_parse(path):
    """Load a reference pf_params.yaml into a dict (jax_am/common.py:24-31; SafeLoader here)."""
    with open(path) as f:
        args = yaml.load(f, Loader=yaml.SafeLoader)
    return args


def with_defaults(pf_args):
    """Copy of pf_args with the hard-coded constants filled in where absent."""
    out = dict(HARD_CODED)
    out.setdefault("anisotropy", 0.0)
    out.setdefault("ad_hoc", 1.0)
    out.update(pf_args)
    return out


def mobility(T, pf_args):
    """Arrhenius mobility L0 exp(-Qg / (R T)) at T clipped to T_clip (allen_cahn.py:56,83)."""
    a = with_defaults(pf_args)
    T = min(T, a["T_clip"])
    return a["L0"] * math.exp(-a["Qg"] / (T * a["gas_const"]))


def stable_dt(pf_args, h_xyz, fraction=0.2, c_local=None):
    """Explicit-Euler step at `fraction` of the linear stability limit.

    dt_max = 2 / (L_max (4 kappa_max sum_d 1/h_d^2 + m_g c)), with L_max the mobility at
    T_clip, kappa_max = kappa_g ad_hoc (1 + anisotropy), and c a bound on the diagonal
    of the local-term Jacobian: 4 for a just-melted one-hot cell plus 4*gamma for the
    cross terms (c_local overrides).  Returns (dt, dt_max, c).
    """
    a = with_defaults(pf_args)
    c = 4.0 + 4.0 * a["gamma"] if c_local is None else c_local
    L_max = mobility(a["T_clip"], a)
    kap = a["kappa_g"] * a["ad_hoc"] * (1.0 + max(a["anisotropy"], 0.0))
    stiff = 4.0 * kap * sum(1.0 / (h * h) for h in h_xyz) + a["m_g"] / a["ad_hoc"] * c
    dt_max = 2.0 / (L_max * stiff)
    return fraction * dt_max, dt_max, c


def make_params(pf_args, h_xyz, laser=None):
    """Pack pf_args (+ a laser descriptor) into the C struct pf_params_t."""
    a = with_defaults(pf_args)
    p = _lib.ParamsT()
    p.nz, p.ny, p.nx = int(a["Nz"]), int(a["Ny"]), int(a["Nx"])
    p.num_oris = int(a["num_oris"])
    for d in range(3):
        p.h[d] = float(h_xyz[d])
    p.m_g = float(a["m_g"])
    p.gamma = float(a["gamma"])
    p.kappa_g = float(a["kappa_g"])
    p.ad_hoc = float(a["ad_hoc"])
    p.L0 = float(a["L0"])
    p.Qg = float(a["Qg"])
    p.gas_const = float(a["gas_const"])
    p.T_liquidus = float(a["T_liquidus"])
    p.T_clip = float(a["T_clip"])
    p.zeta_sharpness = float(a["zeta_sharpness"])
    p.anisotropy = float(a["anisotropy"])
    p.force_liquid = 1 if a["force_liquid"] else 0
    p.liquid_shift = float(a["liquid_shift"])
    if laser is not None:
        laser.fill(p.laser)
    else:
        p.laser.model = _lib.PF_T_BUFFER
        p.laser.T_dtype = _lib.PF_F64
    return p
