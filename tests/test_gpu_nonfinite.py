"""Non-finite values on the GPU path: where they spread in one step, and how the server isolates
a solve that holds them.

The reference propagates a NaN through state_rhs (allen_cahn.py:53-85): every channel of a cell
with a non-finite channel becomes NaN (the local term multiplies 0 by a non-finite sum of
squares), and so does each neighbour's channel through the Laplacian; a NaN temperature makes
zeta and the mobility NaN for the whole cell.  The fused kernel skips channels that are zero on
their whole stencil, and must not skip them in such cells: the NaN pattern after a step must be
exactly the oracle's.  The reference aborts a run on non-finite values (allen_cahn.py:179-180);
the server fails that request alone."""
import numpy as np
import pytest
import torch

from oracle import pf_oracle as O
from pfkit import PFSolver, workloads
from pfkit.batch import BatchEngine
from pfkit.serve import Request, Server
from problems import args, gaussian, poly

pytestmark = pytest.mark.gpu


def _onehot(pc, K):
    eta = np.zeros((pc.n_cells, K))
    eta[np.arange(pc.n_cells), pc.cell_ori_inds] = 1.0
    return eta


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("eps", [0.0, 0.15])
def test_nan_in_eta_spreads_like_the_oracle(dtype, eps, cuda):
    Nx, Ny, Nz, K = 24, 16, 12, 6
    print(f"NaN in eta: {dtype} eps {eps}", flush=True)
    a = args(Nx, Ny, Nz, K, anisotropy=eps, force_liquid=True)
    pc = poly(Nx, Ny, Nz, K, seed=5)
    L = gaussian(Nx, Ny, Nz)
    eta = _onehot(pc, K)
    c = (6 * Ny + 7) * Nx + 9                  # an interior cell, away from the heated surface
    j = (int(pc.cell_ori_inds[c]) + 1) % K     # a channel that is zero around it
    eta[c, j] = np.nan
    t0, t1 = 0.2, 0.23
    T = O.laser_T(L.as_dict(), O.centroids(Nz, Ny, Nx, pc.mesh_h_xyz), t0)
    with np.errstate(invalid="ignore"):
        ref = O.explicit_euler(eta, t0, t1, T, a, pc.mesh_h_xyz, pc.unique_grain_directions)
    s = PFSolver(a, pc, dtype=dtype, device=cuda, clean=False)
    for tile, zc in ((s.tile, s.zchunk), (-2, 0)):   # the fused tiled kernel and the generic kernel
        x = torch.as_tensor(eta, dtype=dtype, device=cuda)
        out = torch.empty_like(x)
        s.state_rhs.step(x, out, L, t0, t1, tile, zc)
        got = out.double().cpu().numpy()
        assert np.array_equal(np.isnan(got), np.isnan(ref)), (tile, np.argwhere(np.isnan(got) != np.isnan(ref))[:5])
        assert np.count_nonzero(np.isnan(ref)) == K + 6   # the cell's K channels + channel j of 6 neighbours
        fin = ~np.isnan(ref)
        tol = 1e-12 if dtype == torch.float64 else 2e-6
        assert np.max(np.abs(got[fin] - ref[fin])) <= tol


def test_nan_temperature_spreads_like_the_oracle(cuda):
    Nx, Ny, Nz, K = 20, 12, 10, 4
    a = args(Nx, Ny, Nz, K, anisotropy=0.15, force_liquid=True)
    pc = poly(Nx, Ny, Nz, K, seed=6)
    eta = _onehot(pc, K)
    T = np.full((pc.n_cells, 1), 1000.0)
    T[77, 0] = np.nan
    t0, t1 = 0.0, 0.05
    with np.errstate(invalid="ignore"):
        ref = O.explicit_euler(eta, t0, t1, T, a, pc.mesh_h_xyz, pc.unique_grain_directions)
    for dtype in (torch.float64, torch.float32):
        s = PFSolver(a, pc, dtype=dtype, device=cuda, clean=False)
        (out, _), _ = s.stepper((torch.as_tensor(eta, dtype=dtype, device=cuda), t0), t1, [T])
        got = out.double().cpu().numpy()
        assert np.array_equal(np.isnan(got), np.isnan(ref)), dtype
        assert np.array_equal(np.flatnonzero(np.isnan(ref).any(axis=1)), [77])


def test_server_fails_the_nan_request_alone(cuda):
    wls = [workloads.make(f"n{j}", nx, ny, nz, k, 12, st, dt, 0.15, seed=70 + j, vel=2.0)
           for j, (nx, ny, nz, k, dt, st) in enumerate(((40, 32, 32, 10, "float32", 40), (36, 33, 32, 16, "bfloat16", 33),
                                                        (32, 32, 32, 12, "float32", 50), (48, 32, 32, 20, "float32", 27)))]
    solo = {}
    for w in wls:
        s = PFSolver(w.pf_args, w.polycrystal, dtype=w.dtype, clean=False)
        solo[w.name] = s.run(s.ini_cond(), w.ts(), w.laser)
    bad_eta = _onehot(wls[2].polycrystal, wls[2].K)
    bad_eta[4321, 3] = np.nan
    insts = [w.instance() for w in wls]
    insts[2] = wls[2].instance(eta0=bad_eta)
    reqs = [Request(rid=j, arrival=0.0, instance=i) for j, i in enumerate(insts)]
    eng = BatchEngine(max_instances=8)
    srv = Server(eng, 1 << 32, chunk=8, keep_results=True)
    srv.serve(reqs)
    eng.close()
    for r, w in zip(reqs, wls):
        print(f"request {r.rid}: {r.status} {r.error}", flush=True)
        if r.rid == 2:
            assert r.status == "failed" and "non-finite" in r.error and r.result is None
        else:
            assert r.status == "done" and torch.equal(r.result, solo[w.name]), r.rid
    assert srv.used == 0
