"""Grain segmentation (pfkit.segment on the GPU; oracle/segmentation.py restates
jax_am/phase_field/utils.py:217-310)."""
import numpy as np
import pytest

from oracle import pf_oracle as O
from oracle import segmentation as S
from problems import args, gaussian, poly


def test_oracle_pca_matches_sklearn():
    from sklearn.decomposition import PCA
    rng = np.random.default_rng(4)
    for n in (5, 9, 40):
        X = rng.normal(size=(n, 3)) * np.array([3.0, 1.0, 0.2]) + rng.normal(size=3) * 10
        ref = PCA(n_components=3).fit(X).explained_variance_
        np.testing.assert_allclose(S.pca_variances(X), ref, rtol=1e-10)


def test_oracle_bfs_order_by_hand():
    # 4 x 3 x 1 grid (x fastest), orientations:
    #   y=0: 0 0 1 1
    #   y=1: 0 1 1 0
    #   y=2: 0 0 0 0     cell 5 (x=1, y=1) is liquid
    ori = np.array([0, 0, 1, 1, 0, 1, 1, 0, 0, 0, 0, 0])
    nonliquid = np.ones(12, bool)
    nonliquid[5] = False
    g = S.bfs_grains(ori, nonliquid, 2, 1, 3, 4)
    # orientation 0: {0,1,4,8,9,10,11,7} from cell 0; orientation 1: {2,3,6} from cell 2
    assert [sorted(x) for x in g] == [[0, 1, 4, 7, 8, 9, 10, 11], [2, 3, 6]]
    # BFS visiting order of the first grain (x-, x+, y-, y+ neighbour order)
    assert g[0][:3] == [0, 1, 4]


def test_oracle_box_aspect_ratio_closed_form():
    Nx, Ny, Nz = 10, 6, 5
    ori = np.zeros(Nx * Ny * Nz, int)
    z, y, x = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    box = (x >= 2) & (x < 8) & (y >= 1) & (y < 4) & (z >= 1) & (z < 3)   # 6 x 3 x 2 cells
    ori[box.reshape(-1)] = 1
    grains, vols, ars = S.segment(ori, np.ones(len(ori), bool), 2, Nz, Ny, Nx, (1.0, 1.0, 1.0))
    i = [k for k, gg in enumerate(grains) if ori[gg[0]] == 1][0]
    assert vols[i] == 36.0
    expected = 2 * np.sqrt(6**2 - 1) / (np.sqrt(3**2 - 1) + np.sqrt(2**2 - 1))
    assert abs(ars[i] - expected) < 1e-12


def test_host_statistics_match_oracle_on_random_grains():
    from pfkit.segment import _covariances, aspect_ratios
    rng = np.random.default_rng(2)
    Nx, Ny, Nz, h = 9, 7, 6, (0.5, 1.0, 2.0)
    ori = rng.integers(0, 2, size=Nx * Ny * Nz)
    grains, vols, ars = S.segment(ori, np.ones(len(ori), bool), 2, Nz, Ny, Nx, h)
    mom = []
    for g in grains:
        iz, r = np.divmod(np.array(g), Ny * Nx)
        iy, ix = np.divmod(r, Nx)
        mom.append([len(g), ix.sum(), iy.sum(), iz.sum(), (ix * ix).sum(), (iy * iy).sum(), (iz * iz).sum(),
                    (ix * iy).sum(), (ix * iz).sum(), (iy * iz).sum()])
    mom = np.array(mom, dtype=np.int64)
    got = aspect_ratios(_covariances(mom, h), mom[:, 0])
    np.testing.assert_allclose(got, ars, rtol=1e-9)


def _state(Nx, Ny, Nz, K, seed):
    a = args(Nx, Ny, Nz, K)
    pc = poly(Nx, Ny, Nz, K, grains=max(4, Nx * Ny * Nz // 60), seed=seed)
    eta = O.ini_cond(pc.cell_ori_inds, K) * 0.9 + 0.05 * np.random.default_rng(seed).random((Nx * Ny * Nz, K))
    return a, pc, eta


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(20, 16, 12, 4, 1), (33, 9, 7, 3, 2), (1, 1, 30, 2, 3)])
def test_gpu_segmentation_matches_oracle(shape, cuda):
    import torch
    from pfkit import PFSolver
    from pfkit.segment import grains
    Nx, Ny, Nz, K, seed = shape
    a, pc, eta = _state(Nx, Ny, Nz, K, seed)
    L = gaussian(Nx, Ny, Nz)                    # melts a pool under the top surface
    t = 0.3
    T = O.laser_T(L.as_dict(), O.centroids(Nz, Ny, Nx, pc.mesh_h_xyz), t).reshape(-1)
    ref_g, ref_v, ref_a = S.segment(np.argmax(eta, axis=1), T < a["T_liquidus"], K, Nz, Ny, Nx, pc.mesh_h_xyz)
    s = PFSolver(a, pc, dtype=torch.float64, device=cuda, clean=False)
    got = grains(s, eta, L, t)
    assert len(got.root) == len(ref_g)
    assert got.root.tolist() == [min(g) for g in ref_g]
    assert got.n_cells.tolist() == [len(g) for g in ref_g]
    np.testing.assert_allclose(got.volumes, ref_v, rtol=1e-12)
    np.testing.assert_allclose(got.aspect_ratios, ref_a, rtol=1e-9)
    lab = got.labels.cpu().numpy()
    for g in ref_g:
        assert np.all(lab[g] == min(g))
    liquid = ~(T < a["T_liquidus"])
    assert np.all(lab[liquid] == -1)
