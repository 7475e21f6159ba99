"""ctypes wrapper of the C/OpenMP oracle (pf_oracle.c -> _build/libpforacle.so).

TEST INFRASTRUCTURE ONLY (CPU baseline of bench.py and a cross-check of pf_oracle.py).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "libpforacle.so")
_MODELS = {"buffer": 0, "rosenthal": 1, "gaussian": 2, "quench": 3}


class _Params(ctypes.Structure):
    _fields_ = [("nz", ctypes.c_int), ("ny", ctypes.c_int), ("nx", ctypes.c_int), ("k", ctypes.c_int),
                ("h", ctypes.c_double * 3)] + \
               [(n, ctypes.c_double) for n in ("m_g", "gamma", "kappa_g", "ad_hoc", "L0", "Qg", "gas_const",
                                               "T_liquidus", "T_clip", "zeta_sharpness", "anisotropy",
                                               "liquid_shift")] + \
               [("force_liquid", ctypes.c_int), ("model", ctypes.c_int), ("T_ambient", ctypes.c_double),
                ("pos0", ctypes.c_double * 3), ("dir", ctypes.c_double * 3)] + \
               [(n, ctypes.c_double) for n in ("vel", "Q", "kappa_T", "alpha", "T_cap", "dT_peak", "sigma_xy",
                                               "sigma_z", "T_ref", "grad_T", "cool_rate")]


_lib = None


def available():
    return os.path.exists(LIB)


def load():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        _lib = ctypes.CDLL(LIB)
        for nm in ("pfo_step_f64", "pfo_step_f32"):
            fn = getattr(_lib, nm)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(_Params), ctypes.c_void_p,
                           ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        _lib.pfo_max_threads.restype = ctypes.c_int
    return _lib


def max_threads():
    return load().pfo_max_threads()


def make_params(P, h_xyz, laser=None):
    p = _Params()
    p.nz, p.ny, p.nx, p.k = int(P["Nz"]), int(P["Ny"]), int(P["Nx"]), int(P["num_oris"])
    for d in range(3):
        p.h[d] = h_xyz[d]
    for n in ("m_g", "kappa_g", "ad_hoc", "L0", "Qg", "gas_const", "T_liquidus"):
        setattr(p, n, float(P[n]))
    p.gamma = float(P.get("gamma", 1.0))
    p.T_clip = float(P.get("T_clip", 2000.0))
    p.zeta_sharpness = float(P.get("zeta_sharpness", 1e10))
    p.anisotropy = float(P.get("anisotropy", 0.0))
    p.liquid_shift = float(P.get("liquid_shift", 1.0))
    p.force_liquid = 1 if P.get("force_liquid", False) else 0
    if laser is None:
        p.model = 0
        return p
    p.model = _MODELS[laser["model"]]
    for n in ("T_ambient", "vel", "Q", "kappa_T", "alpha", "T_cap", "dT_peak", "sigma_xy", "sigma_z", "T_ref",
              "grad_T", "cool_rate"):
        if n in laser:
            setattr(p, n, float(laser[n]))
    for d in range(3):
        if "pos0" in laser:
            p.pos0[d] = laser["pos0"][d]
        p.dir[d] = laser.get("dir", (1.0, 0.0, 0.0))[d]
    return p


def step(eta_in, eta_out, params, axes, t_prev, t_crt, T=None, threads=0):
    """One step in place into eta_out (float64 -> pfo_step_f64, float32 -> pfo_step_f32)."""
    lib = load()
    fn = lib.pfo_step_f64 if eta_in.dtype == np.float64 else lib.pfo_step_f32
    assert eta_in.flags.c_contiguous and eta_out.flags.c_contiguous and eta_out.dtype == eta_in.dtype
    ax = np.ascontiguousarray(axes, dtype=np.float64) if axes is not None else None
    Tb = np.ascontiguousarray(T, dtype=np.float64).reshape(-1) if T is not None else None
    rc = fn(eta_in.ctypes.data, eta_out.ctypes.data, ctypes.byref(params),
            ax.ctypes.data if ax is not None else None, Tb.ctypes.data if Tb is not None else None,
            float(t_prev), float(t_crt), int(threads))
    if rc != 0:
        raise RuntimeError("pfo_step failed")
    return eta_out
