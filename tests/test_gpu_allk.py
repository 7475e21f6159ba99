"""Every register-tiled instantiation of the fused step (K = 1..24 and 32, all tile shapes, FP64 /
FP32 / BF16) against the generic-K kernel, bitwise, and the FP64 generic kernel against the oracle.

The grid (19 x 11 x 6) leaves partial tiles in x and y and a z-chunk shorter than the column; the
state mixes dense random values with channels that are zero over whole regions, so the wave-wide
zero-channel and zero-gradient skips, the cold and hot temperature paths and the liquid reset all
occur.  Tolerance of the oracle comparison: relative L-inf <= 1e-12 per step.
"""
import numpy as np
import pytest
import torch

from oracle import pf_oracle as O
from pfkit import _lib
from pfkit.allen_cahn import PFSolver
from problems import args, gaussian, poly, random_eta

pytestmark = pytest.mark.gpu

KS = list(range(1, 25)) + [32]


def _state(N, K):
    eta = random_eta(N, K, seed=100 + K)
    rng = np.random.default_rng(K)
    for k in range(K):
        if rng.random() < 0.5:  # this channel is zero on a contiguous third of the cells
            lo = rng.integers(0, N // 2)
            eta[lo:lo + N // 3, k] = 0.0
    return eta


@pytest.mark.parametrize("K", KS)
def test_every_tiled_instantiation_matches_generic_kernel(K, cuda):
    Nx, Ny, Nz = 19, 11, 6
    a = args(Nx, Ny, Nz, K, anisotropy=0.15, force_liquid=True)
    pc = poly(Nx, Ny, Nz, K)
    L = gaussian(Nx, Ny, Nz)
    eta = _state(Nx * Ny * Nz, K)
    t0, t1 = 0.1, 0.12
    lib = _lib.load()
    for dtype in (torch.float64, torch.float32, torch.bfloat16):
        s = PFSolver(a, pc, dtype=dtype, device=cuda, clean=False)
        e = torch.as_tensor(eta, dtype=dtype, device=cuda)

        def step(tile, zc):
            out = torch.empty_like(e)
            s.state_rhs.step(e, out, L, t0, t1, tile, zc)
            torch.cuda.synchronize()
            return out

        base = step(-2, 0)
        if dtype == torch.float64:
            cent = O.centroids(Nz, Ny, Nx, pc.mesh_h_xyz)
            ref = O.explicit_euler(eta, t0, t1, O.laser_T(L.as_dict(), cent, t0), a, pc.mesh_h_xyz,
                                   pc.unique_grain_directions)
            got = base.cpu().numpy()
            assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)), K
        nt = lib.pf_tile_count(_lib.TORCH_TO_PF[dtype], K)
        assert nt >= 2
        for tile in range(nt):
            for zc in (0, 4, 4 | _lib.ZCHUNK_TAIL):
                assert torch.equal(step(tile, zc), base), (dtype, K, tile, zc)
