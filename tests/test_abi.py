"""CPU checks of the C ABI: the library loads, exports every function include/pfhip.h
declares, the ctypes struct layouts match C, and argument validation fails cleanly
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

import pytest

from pfkit import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "pfhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(pf_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_version_and_struct_sizes():
    lib = _lib.load()
    assert lib.pf_version() == _lib.ABI_VERSION == 2
    sp, sl, sb = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
    lib.pf_abi_sizes(ctypes.byref(sp), ctypes.byref(sl), ctypes.byref(sb))
    assert (sp.value, sl.value, sb.value) == (ctypes.sizeof(_lib.ParamsT), ctypes.sizeof(_lib.LaserT),
                                             ctypes.sizeof(_lib.BatchInstT))


def test_header_constants_match_the_bindings():
    src = open(os.path.join(ROOT, "include", "pfhip.h")).read()
    defs = dict(re.findall(r"#define\s+(PF\w+)\s+(\S+)", src))
    assert int(defs["PFHIP_ABI_VERSION"]) == _lib.ABI_VERSION
    assert int(defs["PF_PATH_ROW"]) == _lib.PF_PATH_ROW
    assert int(defs["PF_ZCHUNK_TAIL"], 0) == _lib.ZCHUNK_TAIL


def test_tile_tables():
    lib = _lib.load()
    for dt in (_lib.PF_F32, _lib.PF_BF16):
        for k in list(range(1, 25)) + [32]:
            # three one-lane shapes; even K also two lane-pair shapes
            assert lib.pf_fast_path(dt, k) == 1 and lib.pf_tile_count(dt, k) == (5 if k % 2 == 0 else 3)
        tx, ty = ctypes.c_int32(), ctypes.c_int32()
        assert lib.pf_tile_shape(dt, 20, 0, ctypes.byref(tx), ctypes.byref(ty)) == 0
        assert (tx.value, ty.value) == (32, 8) and lib.pf_tile_lanes(dt, 20, 0) == 1
        assert lib.pf_tile_shape(dt, 20, 3, ctypes.byref(tx), ctypes.byref(ty)) == 0
        assert (tx.value, ty.value) == (16, 8) and lib.pf_tile_lanes(dt, 20, 3) == 2
        assert lib.pf_tile_lanes(dt, 19, 3) < 0
    assert lib.pf_tile_count(_lib.PF_F64, 20) == 4 and lib.pf_tile_count(_lib.PF_F64, 19) == 2
    assert lib.pf_fast_path(_lib.PF_F32, 25) == 0 and lib.pf_fast_path(_lib.PF_F32, 64) == 0
    assert lib.pf_tile_shape(_lib.PF_F32, 20, 7, None, None) != 0


def test_argument_validation_without_gpu():
    from pfkit.config import make_params
    from problems import args
    lib = _lib.load()
    p = make_params(args(8, 8, 8, 4), (1.0, 1.0, 1.0))   # PF_T_BUFFER with no T buffer
    rc = lib.pf_step(16, 32, _lib.PF_F32, None, p, None, 0.0, 1.0, -1, 0, None)
    assert rc == -1 and b"T buffer" in lib.pf_last_error()
    rc = lib.pf_step(16, 16, _lib.PF_F32, None, p, None, 0.0, 1.0, -1, 0, None)
    assert rc == -1 and b"alias" in lib.pf_last_error()
    p.nx = 0
    rc = lib.pf_step(16, 32, _lib.PF_F32, 3, p, None, 0.0, 1.0, -1, 0, None)
    assert rc == -1 and b"grid" in lib.pf_last_error()
    p.nx = 8
    # zchunk: a plane count below 2^16, optionally with PF_ZCHUNK_TAIL; other bits are refused
    rc = lib.pf_step(16, 32, _lib.PF_F32, 3, p, None, 0.0, 1.0, -1, 0x20000 | 4, None)
    assert rc == -1 and b"zchunk" in lib.pf_last_error()
    p.anisotropy = 0.1
    rc = lib.pf_step(16, 32, _lib.PF_F32, 3, p, None, 0.0, 1.0, -1, 0, None)
    assert rc == -1 and b"grain-direction" in lib.pf_last_error()
    with pytest.raises(_lib.PFError):
        _lib.check(-1, "x")
    assert lib.pf_reduce_workspace_bytes(1000, 10) > 0
    assert lib.pf_batch_create(0, 0, None) is None
    rc = lib.pf_step(16, 40, _lib.PF_F32, None, p, None, 0.0, 1.0, -1, 0, None)
    assert rc == -1 and b"aligned" in lib.pf_last_error()
    # the tiled kernels address a z-plane with 32-bit byte offsets and drop out-of-range buffer
    # stores at 2^31: a tiled step on 6000 x 6000 x 20 FP32 values (2.9 GB) per plane is refused
    # before any launch
    big = make_params(args(6000, 6000, 1, 20), (1.0, 1.0, 1.0))
    rc = lib.pf_step(16, 32, _lib.PF_F32, 3, big, None, 0.0, 1.0, 0, 0, None)
    assert rc == -1 and b"2^31 bytes" in lib.pf_last_error()


def test_noalloc_reduce_refuses_oversized_input_without_gpu():
    # pf_reduce_noalloc checks the reduction's scratch need on the host before any HIP call
    from pfkit.config import make_params
    from problems import args, gaussian
    lib = _lib.load()
    big = make_params(args(256, 256, 128, 256), (1.0, 1.0, 1.0), gaussian(256, 256, 128))
    assert lib.pf_reduce_workspace_bytes(256 * 256 * 128, 256) > _lib.REDUCE_SCRATCH_BYTES
    rc = lib.pf_reduce_noalloc(16, _lib.PF_F32, None, None, big, 0.0, 16, 16, None)
    assert rc == -1 and b"scratch" in lib.pf_last_error()
    # the c1 diagnostics (256 x 256 x 64 cells, K = 20) fit the fixed scratch
    assert lib.pf_reduce_workspace_bytes(256 * 256 * 64, 20) <= _lib.REDUCE_SCRATCH_BYTES
    rc = lib.pf_reduce_noalloc(16, _lib.PF_F32, None, None, big, 0.0, None, 16, None)
    assert rc == -1 and b"NULL" in lib.pf_last_error()
