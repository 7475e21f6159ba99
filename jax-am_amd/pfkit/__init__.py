"""pfkit: MI355X-native phase-field Allen-Cahn explicit step (JAX-AM jax_am/phase_field drop-in).

The compute path is libpfhip.so (hand-written HIP kernels for gfx950) called through
ctypes on PyTorch-ROCm tensors; see allen_cahn.PFSolver for the reference-shaped API.
"""
from .allen_cahn import PFSolver, StateRHS, explicit_euler, get_force_eta_fn, phase_field, rk4  # noqa: F401
from .config import make_params, stable_dt, with_defaults, yaml_parse  # noqa: F401
from .laser import Gaussian, Quench, Rosenthal, Toolpath  # noqa: F401
from .polycrystal import Field, Polycrystal  # noqa: F401

__version__ = "0.1.0"
