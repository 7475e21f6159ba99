"""One-way CFD -> phase-field coupling (applications/phase_field/coupled/example.py:31-42).

The reference's coupled loop linearly interpolates two CFD temperature snapshots in time
(``interpolate_T``: r = (t_cfd - t_pf) / dt_cfd, T = r T_past + (1 - r) T_future) and reorders
the CFD array (Nx, Ny, Nz) to the phase-field cell order (``convert_temperature``).  Here both
happen in one device pass (libpfhip.so pf_cfd_temperature); the result is the (N, 1) T buffer
that ``PFSolver.stepper`` takes in ``ode_params``.
"""
import torch

from . import _lib


def cfd_temperature(T_past, T_future, t_pf, t_cfd, dt_cfd, out=None):
    """T_past, T_future: (Nx, Ny, Nz) float64/float32 device tensors (CFD layout).  Returns the
    (Nz*Ny*Nx, 1) phase-field temperature on the same device."""
    if T_past.shape != T_future.shape or T_past.dim() != 3:
        raise ValueError("T_past and T_future must be (Nx, Ny, Nz) arrays of the same shape")
    if T_past.dtype != T_future.dtype or T_past.dtype not in (torch.float64, torch.float32):
        raise ValueError("temperatures must both be float64 or both float32")
    Tp, Tf = T_past.contiguous(), T_future.contiguous()
    Nx, Ny, Nz = Tp.shape
    if out is None:
        out = torch.empty((Nx * Ny * Nz, 1), dtype=Tp.dtype, device=Tp.device)
    lib = _lib.load()
    _lib.check(lib.pf_cfd_temperature(Tp.data_ptr(), Tf.data_ptr(), _lib.TORCH_TO_PF[Tp.dtype], Nz, Ny, Nx,
                                      float(t_pf), float(t_cfd), float(dt_cfd), out.data_ptr(),
                                      _lib.stream_handle(Tp.device)), "pf_cfd_temperature")
    return out
