"""Seeded small problems shared by the CPU and GPU tests."""
import numpy as np

from pfkit.config import with_defaults
from pfkit.laser import Gaussian, Quench, Rosenthal
from pfkit.polycrystal import Polycrystal

# Reference physical constants (applications/phase_field/simple/pf_params.yaml) in units of
# micrometres and microseconds: L0, Qg, gas_const give L(T) in 1/us.
BASE = dict(m_g=1.0, kappa_g=1.0, ad_hoc=1.0, L0=3.5e3, Qg=1.4e5, gas_const=8.3, T_liquidus=1623.0,
            gamma=0.75, T_clip=2000.0, zeta_sharpness=1e10, force_liquid=False, liquid_shift=1.0)


def args(Nx, Ny, Nz, K, **kw):
    a = dict(BASE, Nx=Nx, Ny=Ny, Nz=Nz, num_oris=K)
    a.update(kw)
    return with_defaults(a)


def poly(Nx, Ny, Nz, K, grains=None, seed=3, h=(1.0, 1.0, 1.0)):
    return Polycrystal(Nx, Ny, Nz, h, grains or max(4, (Nx * Ny * Nz) // 200), K, seed=seed)


def random_eta(N, K, seed=0, lo=0.0, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, size=(N, K))


def gaussian(Nx, Ny, Nz, h=(1.0, 1.0, 1.0), vel=2.0, dT=1900.0):
    Lx, Ly, Lz = Nx * h[0], Ny * h[1], Nz * h[2]
    return Gaussian(pos0=(0.3 * Lx, 0.5 * Ly, Lz), vel=vel, dT_peak=dT, sigma_xy=0.25 * Ly, sigma_z=0.5 * Lz)


def rosenthal(Nx, Ny, Nz, h=(1.0, 1.0, 1.0), Q=1.6e6):
    """Rosenthal source in micrometre/microsecond units: Q/(2 pi kappa_T) ~ 9.4e3 K um puts the
    T_liquidus isotherm ~5 um from the source."""
    Lx, Ly, Lz = Nx * h[0], Ny * h[1], Nz * h[2]
    return Rosenthal(pos0=(0.3 * Lx, 0.5 * Ly, Lz), vel=0.5, Q=Q, alpha=5.2, kappa_T=27.0)


def quench(Nz, h=1.0):
    return Quench(T_ref=1623.0 - 0.3 * Nz * h * 20.0, grad_T=20.0, cool_rate=40.0)
