"""CPU restatement of the reference's grain segmentation (TEST INFRASTRUCTURE ONLY: only
``tests/`` may import it; the product path is ``pfkit.segment`` on the GPU).

jax_am/phase_field/utils.py:217-310 (compute_edges_in_order, BFS, get_aspect_ratio_inputs,
compute_aspect_ratios_and_vols), as driven by process_eta (:181-214):

* graph: the 6-neighbour face adjacency of the (Nz, Ny, Nx) grid (cells in C order);
* grains: breadth-first search from every unvisited non-liquid cell in increasing cell index,
  over neighbours with the same orientation index (argmax eta) that are non-liquid; grains are
  listed by orientation index, and within one orientation in the order they were found;
* per grain: volume = sum of cell volumes; aspect ratio = 2 l0 / (l1 + l2) with l = sqrt of
  the PCA explained variances (sklearn PCA(n_components=3): variances of the centred data,
  denominator n - 1, descending) of the volume-weighted cell centroids; grains of fewer than 5
  cells get aspect ratio 1.
"""
from collections import deque

import numpy as np


def neighbours(c, Nz, Ny, Nx):
    """Face neighbours of cell c, in the order x-, x+, y-, y+, z-, z+ (those inside the grid)."""
    iz, r = divmod(c, Ny * Nx)
    iy, ix = divmod(r, Nx)
    out = []
    if ix > 0:
        out.append(c - 1)
    if ix < Nx - 1:
        out.append(c + 1)
    if iy > 0:
        out.append(c - Nx)
    if iy < Ny - 1:
        out.append(c + Nx)
    if iz > 0:
        out.append(c - Nx * Ny)
    if iz < Nz - 1:
        out.append(c + Nx * Ny)
    return out


def bfs_grains(ori, nonliquid, K, Nz, Ny, Nx):
    """utils.py:236-265 with combined=True: list of grains (lists of cell indices)."""
    n = len(ori)
    visited = np.zeros(n, dtype=bool)
    grains = [[] for _ in range(K)]
    for i in range(n):
        if not visited[i] and nonliquid[i]:
            o = ori[i]
            g = []
            q = deque([i])
            visited[i] = True
            while q:
                s = q.popleft()
                g.append(s)
                for cn in neighbours(s, Nz, Ny, Nx):
                    if not visited[cn] and ori[cn] == o and nonliquid[cn]:
                        q.append(cn)
                        visited[cn] = True
            grains[o].append(g)
    return [g for per_ori in grains for g in per_ori]


def pca_variances(X):
    """sklearn.decomposition.PCA(n_components=3).fit(X).explained_variance_ (descending)."""
    Xc = X - X.mean(axis=0)
    s = np.linalg.svd(Xc, compute_uv=False)
    return s**2 / (len(X) - 1)


def aspect_ratios_and_vols(grains, volumes, centroids):
    """utils.py:268-310: ([grain volumes], [grain aspect ratios])."""
    vols, ars = [], []
    for g in grains:
        gv = volumes[g]
        vols.append(float(np.sum(gv)))
        if len(gv) < 5:
            ars.append(1.0)
        else:
            ev = pca_variances(centroids[g] * gv[:, None])
            ln = np.sqrt(ev)
            ars.append(float(2 * ln[0] / (ln[1] + ln[2])))
    return vols, ars


def segment(ori, nonliquid, K, Nz, Ny, Nx, h_xyz):
    """Grains and [volumes, aspect ratios] of an orientation map on the regular grid."""
    from .pf_oracle import centroids as cell_centroids
    cent = cell_centroids(Nz, Ny, Nx, h_xyz)
    volumes = np.full(len(ori), h_xyz[0] * h_xyz[1] * h_xyz[2])
    grains = bfs_grains(np.asarray(ori), np.asarray(nonliquid), K, Nz, Ny, Nx)
    vols, ars = aspect_ratios_and_vols(grains, volumes, cent)
    return grains, vols, ars
