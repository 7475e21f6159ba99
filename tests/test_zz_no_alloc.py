"""GPU test of pf_reduce_noalloc, the allocation-free variant of pf_reduce: called through the
C ABI it gives bitwise the FP64 diagnostics of pf_reduce with a caller workspace, on calls in a
row whose later inputs are smaller (the fixed scratch array then holds the previous call's
partial results), and refuses a reduction larger than its scratch without launching."""
import numpy as np
import pytest
import torch

from pfkit import _lib
from pfkit.config import make_params
from problems import args, gaussian, poly, random_eta

pytestmark = pytest.mark.gpu


def _case(nx, ny, nz, k, dtype, seed):
    a = args(nx, ny, nz, k, anisotropy=0.1)
    pc = poly(nx, ny, nz, k, seed=seed)
    p = make_params(a, pc.mesh_h_xyz, gaussian(nx, ny, nz))
    eta = torch.as_tensor(random_eta(nx * ny * nz, k, seed=seed), dtype=dtype, device="cuda")
    eta0 = torch.as_tensor(random_eta(nx * ny * nz, k, seed=seed + 1), dtype=dtype, device="cuda")
    return p, eta, eta0


def _reduce(lib, p, eta, a0, noalloc):
    k, n = p.num_oris, eta.shape[0]
    f = torch.full((4 + k,), float("nan"), dtype=torch.float64, device=eta.device)
    i = torch.full((k + 3,), -7, dtype=torch.int64, device=eta.device)
    st = _lib.stream_handle(eta.device)
    dt = _lib.TORCH_TO_PF[eta.dtype]
    if noalloc:
        rc = lib.pf_reduce_noalloc(eta.data_ptr(), dt, None, a0.data_ptr(), p, 0.37, f.data_ptr(), i.data_ptr(), st)
    else:
        ws = torch.empty(lib.pf_reduce_workspace_bytes(n, k), dtype=torch.uint8, device=eta.device)
        rc = lib.pf_reduce(eta.data_ptr(), dt, None, a0.data_ptr(), p, 0.37, f.data_ptr(), i.data_ptr(),
                           ws.data_ptr(), ws.numel(), st)
    _lib.check(rc, "pf_reduce_noalloc" if noalloc else "pf_reduce")
    torch.cuda.synchronize()
    return f.cpu().numpy(), i.cpu().numpy()


def test_noalloc_equals_workspace_reduce(cuda):
    lib = _lib.load()
    # largest input first, then smaller ones: each later call runs over stale partials
    for (nx, ny, nz, k, dtype, seed) in ((96, 80, 40, 20, torch.float32, 1), (20, 12, 8, 6, torch.float64, 2),
                                         (33, 17, 5, 11, torch.bfloat16, 3)):
        print(f"pf_reduce_noalloc case {nx}x{ny}x{nz} K={k} {dtype}", flush=True)
        p, eta, eta0 = _case(nx, ny, nz, k, dtype, seed)
        dev = torch.device("cuda", 0)
        n = nx * ny * nz
        a0 = torch.empty(n, dtype=torch.int32, device=dev)
        _lib.check(lib.pf_argmax(eta0.data_ptr(), _lib.TORCH_TO_PF[dtype], n, k, a0.data_ptr(),
                                 _lib.stream_handle(dev)), "pf_argmax")
        ref_f, ref_i = _reduce(lib, p, eta, a0, noalloc=False)
        got_f, got_i = _reduce(lib, p, eta, a0, noalloc=True)
        assert np.all(np.isfinite(ref_f)) and ref_i.min() >= 0
        assert np.array_equal(got_f.view(np.uint64), ref_f.view(np.uint64)), (got_f, ref_f)
        assert np.array_equal(got_i, ref_i)
        assert ref_i[:k].sum() == n


def test_noalloc_refuses_oversized_reduction(cuda):
    lib = _lib.load()
    big = make_params(args(256, 256, 128, 256), (1.0, 1.0, 1.0), gaussian(256, 256, 128))
    assert lib.pf_reduce_workspace_bytes(256 * 256 * 128, 256) > _lib.REDUCE_SCRATCH_BYTES
    rc = lib.pf_reduce_noalloc(16, _lib.PF_F32, None, None, big, 0.0, 16, 16, None)
    assert rc == -1 and b"scratch" in lib.pf_last_error()
