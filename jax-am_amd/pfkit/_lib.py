"""ctypes binding of libpfhip.so (include/pfhip.h).

The library is loaded after ``torch`` so that it binds to the HIP runtime torch has
already loaded (same soname, libamdhip64.so.7); a second runtime in the process is
detected and refused.  There is no fallback: without the built library every
compute entry point raises.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PFKIT_LIB") or os.path.join(_HERE, "libpfhip.so")

PF_F64, PF_F32, PF_BF16 = 0, 1, 2
PF_T_BUFFER, PF_T_ROSENTHAL, PF_T_GAUSSIAN, PF_T_QUENCH = 0, 1, 2, 3
REDUCE_SCRATCH_BYTES = 8 << 20   # PF_REDUCE_SCRATCH_BYTES: fixed scratch of pf_reduce_noalloc
PF_PATH_ROW = 10                  # values per row of a batch instance's source table
ZCHUNK_TAIL = 0x10000             # PF_ZCHUNK_TAIL: the lowest quarter of the planes in half-length chunks
ABI_VERSION = 2

TORCH_TO_PF = {torch.float64: PF_F64, torch.float32: PF_F32, torch.bfloat16: PF_BF16}
PF_TO_TORCH = {v: k for k, v in TORCH_TO_PF.items()}


PF_OK, PF_E_ARG, PF_E_UNSUPPORTED, PF_E_HIP, PF_E_STATE = 0, -1, -2, -3, -4


class PFError(RuntimeError):
    """A libpfhip.so call returned an error code (``code``: one of PF_E_*)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class LaserT(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int32), ("T_dtype", ctypes.c_int32), ("T_ambient", ctypes.c_double),
                ("pos0", ctypes.c_double * 3), ("dir", ctypes.c_double * 3), ("vel", ctypes.c_double),
                ("Q", ctypes.c_double), ("kappa_T", ctypes.c_double), ("alpha", ctypes.c_double),
                ("T_cap", ctypes.c_double), ("dT_peak", ctypes.c_double), ("sigma_xy", ctypes.c_double),
                ("sigma_z", ctypes.c_double), ("T_ref", ctypes.c_double), ("grad_T", ctypes.c_double),
                ("cool_rate", ctypes.c_double)]


class ParamsT(ctypes.Structure):
    _fields_ = [("nz", ctypes.c_int32), ("ny", ctypes.c_int32), ("nx", ctypes.c_int32),
                ("num_oris", ctypes.c_int32), ("h", ctypes.c_double * 3), ("m_g", ctypes.c_double),
                ("gamma", ctypes.c_double), ("kappa_g", ctypes.c_double), ("ad_hoc", ctypes.c_double),
                ("L0", ctypes.c_double), ("Qg", ctypes.c_double), ("gas_const", ctypes.c_double),
                ("T_liquidus", ctypes.c_double), ("T_clip", ctypes.c_double),
                ("zeta_sharpness", ctypes.c_double), ("anisotropy", ctypes.c_double),
                ("force_liquid", ctypes.c_int32), ("reserved0", ctypes.c_int32),
                ("liquid_shift", ctypes.c_double), ("laser", LaserT)]


class BatchInstT(ctypes.Structure):
    _fields_ = [("eta_dev", ctypes.c_void_p * 2), ("T_dev", ctypes.c_void_p), ("axes_dev", ctypes.c_void_p),
                ("dtype", ctypes.c_int32), ("tile", ctypes.c_int32), ("zchunk", ctypes.c_int32),
                ("steps", ctypes.c_int32), ("t0", ctypes.c_double), ("dt", ctypes.c_double), ("p", ParamsT),
                ("path_dev", ctypes.c_void_p)]


# name -> (restype, argtypes); every symbol declared in include/pfhip.h
_VP, _I32, _I64, _D, _SZ = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_size_t
SIGNATURES = {
    "pf_version": (ctypes.c_int, []),
    "pf_last_error": (ctypes.c_char_p, []),
    "pf_abi_sizes": (ctypes.c_int, [ctypes.POINTER(_SZ), ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)]),
    "pf_device_count": (ctypes.c_int, []),
    "pf_step": (ctypes.c_int, [_VP, _VP, _I32, _VP, ctypes.POINTER(ParamsT), _VP, _D, _D, _I32, _I32, _VP]),
    "pf_rhs": (ctypes.c_int, [_VP, _VP, _I32, _VP, ctypes.POINTER(ParamsT), _VP, _D, _VP]),
    "pf_tile_count": (ctypes.c_int, [_I32, _I32]),
    "pf_tile_shape": (ctypes.c_int, [_I32, _I32, _I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "pf_tile_lanes": (ctypes.c_int, [_I32, _I32, _I32]),
    "pf_cfd_temperature": (ctypes.c_int, [_VP, _VP, _I32, _I32, _I32, _I32, _D, _D, _D, _VP, _VP]),
    "pf_segment": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(ParamsT), _D, _VP, _VP, _VP, _VP, ctypes.c_int64,
                                  ctypes.POINTER(_I32), _VP]),
    "pf_segment_moments": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(ParamsT), _I32, _VP, _VP]),
    "pf_fast_path": (ctypes.c_int, [_I32, _I32]),
    "pf_temperature": (ctypes.c_int, [_VP, _I32, ctypes.POINTER(ParamsT), _D, _VP]),
    "pf_init_onehot": (ctypes.c_int, [_VP, _I32, _VP, _I64, _I32, _VP]),
    "pf_argmax": (ctypes.c_int, [_VP, _I32, _I64, _I32, _VP, _VP]),
    "pf_voronoi": (ctypes.c_int, [_VP, _I32, _VP, ctypes.POINTER(ParamsT), _VP, _VP]),
    "pf_reduce_workspace_bytes": (_SZ, [_I64, _I32]),
    "pf_reduce": (ctypes.c_int, [_VP, _I32, _VP, _VP, ctypes.POINTER(ParamsT), _D, _VP, _VP, _VP, _SZ, _VP]),
    "pf_reduce_noalloc": (ctypes.c_int, [_VP, _I32, _VP, _VP, ctypes.POINTER(ParamsT), _D, _VP, _VP, _VP]),
    "pf_batch_create": (_VP, [_I32, _I32, _VP]),
    "pf_batch_destroy": (None, [_VP]),
    "pf_batch_add": (ctypes.c_int, [_VP, ctypes.POINTER(BatchInstT)]),
    "pf_batch_remove": (ctypes.c_int, [_VP, _I32]),
    "pf_batch_run": (ctypes.c_int, [_VP, _I32, _I32]),
    "pf_batch_status": (ctypes.c_int, [_VP, ctypes.POINTER(_I32), ctypes.POINTER(_D), ctypes.POINTER(_I32), _I32]),
    "pf_batch_active": (ctypes.c_int, [_VP]),
}

_lib = None


def _hip_runtimes_mapped():
    paths = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    paths.add(os.path.realpath(line.split()[-1]))
    except OSError:
        pass
    return paths


def load():
    """Load libpfhip.so once (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PFError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                      "(or `make -C jax-am_amd/csrc`)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    runtimes = _hip_runtimes_mapped()
    if len(runtimes) > 1:
        raise PFError(f"two HIP runtimes are mapped in this process: {sorted(runtimes)}")
    sp, sl, sb = _SZ(), _SZ(), _SZ()
    lib.pf_abi_sizes(ctypes.byref(sp), ctypes.byref(sl), ctypes.byref(sb))
    for nm, got, want in (("pf_params_t", sp.value, ctypes.sizeof(ParamsT)),
                          ("pf_laser_t", sl.value, ctypes.sizeof(LaserT)),
                          ("pf_batch_inst_t", sb.value, ctypes.sizeof(BatchInstT))):
        if got != want:
            raise PFError(f"ABI mismatch: sizeof({nm}) is {got} in C, {want} in the binding")
    _lib = lib
    return lib


def check(rc, what=""):
    """Raise PFError with the library's message when rc is a PF_E* code."""
    if rc != 0:
        msg = load().pf_last_error().decode(errors="replace")
        raise PFError(f"{what}: {msg} (code {rc})" if what else f"{msg} (code {rc})", code=rc)
    return rc


def stream_handle(device=None):
    """hipStream_t of torch's current stream on `device` (as a c_void_p-compatible int)."""
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t):
    return None if t is None else t.data_ptr()
