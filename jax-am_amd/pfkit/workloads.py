"""The benchmark workloads (seeded, synthetic microstructures).

Units: lengths in micrometres (dx = 1 um, so h = 1) and times in microseconds, so the
reference's Arrhenius constants L0 = 3.5e3, Qg = 1.4e5, gas_const = 8.3
(applications/phase_field/simple/pf_params.yaml) give the mobility in 1/us and a
laser speed of v m/s is v um/us.  m_g = 1, kappa_g = 1; the pair coefficient
"2 gamma" with gamma = 1.5 is the reference's 4*gamma, i.e. gamma = 0.75 in the
reference convention (allen_cahn.py:43-46).  Cells above T_liquidus + 1 are reset to
liquid (force_liquid, allen_cahn.py:117-128).  The laser is a Gaussian spot moving
along +x at the top surface (Rosenthal is available as pfkit.laser.Rosenthal);
sigma_xy = 0.176 Ly and a 1900 K peak rise give a melt pool 0.3 Ly wide.
dt = 0.2 of the explicit stability limit (config.stable_dt, c_local = 4 + 4 gamma).

  c0  64x64x32, K = 10, 50 grains, isotropic, 200 steps, FP64 (the CPU reference case)
  c1  256x256x64, K = 20, 500 grains, cubic anisotropy eps = 0.15, 1000 steps, FP32
  c2  c1 with BF16 storage (FP32 arithmetic)
  c3  64 instances, grids log-uniform 32^3 .. 128x128x64, K in {10, 16, 20},
      laser speed U[0.5, 2] um/us, 200-1000 steps, FP32
  c4  serving: Poisson arrivals, grids log-uniform 32^3 .. 256x256x64, K in [10, 20],
      100-2000 steps, FP32 or BF16 per request
"""
from dataclasses import dataclass, field

import numpy as np

from .config import stable_dt, with_defaults
from .laser import Gaussian
from .polycrystal import Polycrystal

PHYSICS = dict(m_g=1.0, kappa_g=1.0, ad_hoc=1.0, gamma=0.75, L0=3.5e3, Qg=1.4e5, gas_const=8.3,
               T_liquidus=1623.0, T_ambient=300.0, T_clip=2000.0, zeta_sharpness=1e10,
               force_liquid=True, liquid_shift=1.0)


@dataclass
class Workload:
    name: str
    pf_args: dict
    polycrystal: Polycrystal
    laser: object
    dt: float
    steps: int
    dtype: str
    extra: dict = field(default_factory=dict)

    @property
    def N(self):
        return self.polycrystal.n_cells

    @property
    def K(self):
        return int(self.pf_args["num_oris"])

    def instance(self, **kw):
        """The batch.Instance of this workload (one-hot start)."""
        from .batch import Instance
        return Instance(pf_args=self.pf_args, h_xyz=self.polycrystal.mesh_h_xyz,
                        grain_dirs=self.polycrystal.unique_grain_directions, laser=self.laser, dtype=self.dtype,
                        steps=self.steps, dt=self.dt, cell_ori_inds=self.polycrystal.cell_ori_inds, tag=self.name,
                        **kw)

    def ts(self, steps=None):
        n = self.steps if steps is None else steps
        return np.arange(n + 1) * self.dt   # np.arange time grid (simple/example.py:52)


def make(name, Nx, Ny, Nz, K, grains, steps, dtype, anisotropy, seed=1, vel=None, device_grains=False):
    """A workload; device_grains=True builds the grain map on the GPU (DevicePolycrystal)."""
    h = (1.0, 1.0, 1.0)
    a = with_defaults(dict(PHYSICS, Nx=Nx, Ny=Ny, Nz=Nz, num_oris=K, num_grains=grains,
                           anisotropy=anisotropy, domain_x=Nx * h[0], domain_y=Ny * h[1], domain_z=Nz * h[2]))
    dt, dt_max, c = stable_dt(a, h)
    if device_grains:
        from .polycrystal import DevicePolycrystal
        pc = DevicePolycrystal(Nx, Ny, Nz, h, grains, K, seed=seed)
    else:
        pc = Polycrystal(Nx, Ny, Nz, h, grains, K, seed=seed)
    Lx, Ly, Lz = Nx * h[0], Ny * h[1], Nz * h[2]
    if vel is None:
        vel = 0.6 * Lx / (steps * dt)   # crosses 0.6 Lx during the run, like simple/example.py:30
    laser = Gaussian(pos0=(0.2 * Lx, 0.5 * Ly, Lz), vel=vel, dT_peak=1900.0, sigma_xy=0.176 * Ly,
                     sigma_z=0.47 * Lz, T_ambient=PHYSICS["T_ambient"])
    return Workload(name, a, pc, laser, dt, steps, dtype, dict(dt_max=dt_max, c_local=c))


def c0():
    return make("c0", 64, 64, 32, 10, 50, 200, "float64", 0.0)


def c1(dtype="float32"):
    return make("c1" if dtype == "float32" else "c2", 256, 256, 64, 20, 500, 1000, dtype, 0.15)


def c2():
    return c1("bfloat16")


def _loguniform_dims(rng, lo, hi):
    """(Nx, Ny, Nz) with the cell count log-uniform in [prod(lo), prod(hi)]: one log-uniform
    scale applied to every axis between its bounds."""
    u = rng.random()
    return tuple(int(round(l * (h / l) ** u)) for l, h in zip(lo, hi))


def c3_instances(n=64, seed=3, device_grains=True):
    rng = np.random.default_rng(seed)
    out = []
    for j in range(n):
        Nx, Ny, Nz = _loguniform_dims(rng, (32, 32, 32), (128, 128, 64))
        K = int(rng.choice([10, 16, 20]))
        vel = float(rng.uniform(0.5, 2.0))
        steps = int(rng.integers(200, 1001))
        grains = max(8, (Nx * Ny * Nz) // 8000)
        out.append(make(f"c3_{j}", Nx, Ny, Nz, K, grains, steps, "float32", 0.15, seed=seed * 1000 + j, vel=vel,
                        device_grains=device_grains))
    return out


def c4_request(rng, j, device_grains=True):
    Nx, Ny, Nz = _loguniform_dims(rng, (32, 32, 32), (256, 256, 64))
    K = int(rng.integers(10, 21))
    steps = int(rng.integers(100, 2001))
    dtype = "float32" if rng.random() < 0.5 else "bfloat16"
    grains = max(8, (Nx * Ny * Nz) // 8000)
    return make(f"c4_{j}", Nx, Ny, Nz, K, grains, steps, dtype, 0.15, seed=100000 + j,
                vel=float(rng.uniform(0.5, 2.0)), device_grains=device_grains)
