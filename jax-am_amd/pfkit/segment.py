"""Grain segmentation and grain statistics on the GPU.

The reference post-processes a written solution with ``process_eta``
(jax_am/phase_field/utils.py:181-214): cells with T < T_liquidus are grouped into grains by a
breadth-first search over face neighbours with the same orientation index (argmax eta)
(:217-265); each grain gets its volume and an aspect ratio 2 l0 / (l1 + l2) from the PCA of its
volume-weighted cell centroids (sklearn, explained variances l^2, descending; grains of fewer
than 5 cells get 1) (:268-310).  Here the grains come from label propagation on the GPU
(libpfhip.so pf_segment) and the per-grain covariances from exact integer moments
(pf_segment_moments); only the 3x3 eigenvalue problems run on the host.  Grains are listed in
the reference's order: by orientation index, and within one orientation by their first cell.
"""
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib


@dataclass
class Grains:
    labels: torch.Tensor        # (N,) int32 on the device: smallest cell index of the grain, -1 liquid
    root: np.ndarray            # (G,) first (smallest) cell index of each grain
    ori: np.ndarray             # (G,) orientation index
    n_cells: np.ndarray         # (G,)
    volumes: np.ndarray         # (G,) grain volumes
    aspect_ratios: np.ndarray   # (G,)
    covariance: np.ndarray      # (G, 3, 3) covariance of the cell centroids (denominator n - 1)

    def eta_results(self):
        """process_eta's return value: [grain volumes, grain aspect ratios] (lists)."""
        return [self.volumes.tolist(), self.aspect_ratios.tolist()]


def _covariances(mom, h_xyz):
    """(G, 3, 3) centroid covariances from integer moments, exact up to the final division."""
    n = mom[:, 0].astype(object)
    s1 = [mom[:, 1].astype(object), mom[:, 2].astype(object), mom[:, 3].astype(object)]
    s2 = {(0, 0): mom[:, 4], (1, 1): mom[:, 5], (2, 2): mom[:, 6], (0, 1): mom[:, 7], (0, 2): mom[:, 8],
          (1, 2): mom[:, 9]}
    G = mom.shape[0]
    cov = np.zeros((G, 3, 3))
    denom = np.array([float(v * (v - 1)) if v > 1 else np.inf for v in n])
    for (i, j), sij in s2.items():
        num = n * sij.astype(object) - s1[i] * s1[j]           # Python integers: exact
        c = np.array([float(v) for v in num]) / denom * (h_xyz[i] * h_xyz[j])
        cov[:, i, j] = c
        cov[:, j, i] = c
    return cov


def aspect_ratios(cov, n_cells, min_cells=5):
    """utils.py:298-306 from the covariances: 2 l0 / (l1 + l2), l = sqrt(eigenvalues), descending."""
    ev = np.linalg.eigvalsh(cov)[:, ::-1]
    ln = np.sqrt(np.maximum(ev, 0.0))
    with np.errstate(divide="ignore", invalid="ignore"):
        ar = 2.0 * ln[:, 0] / (ln[:, 1] + ln[:, 2])
    return np.where(n_cells < min_cells, 1.0, ar)


def grains(solver, eta, T, t=0.0, max_passes=1 << 20):
    """Segment the state ``eta`` of ``solver`` (a PFSolver) at temperature ``T`` (laser descriptor
    evaluated at ``t``, or an (N,) / (N, 1) array)."""
    lib = _lib.load()
    rhs = solver.state_rhs
    N = rhs.N
    dev = solver.device
    ori = solver.argmax(rhs.check_state(eta))
    p, Tptr = rhs.params_for(T)
    labels = torch.empty(N, dtype=torch.int32, device=dev)
    cid = torch.empty(N, dtype=torch.int32, device=dev)
    root = torch.empty(N, dtype=torch.int32, device=dev)
    ws = torch.zeros(2, dtype=torch.int32, device=dev)
    ng = _lib.ctypes.c_int32(0)
    stream = _lib.stream_handle(dev)
    _lib.check(lib.pf_segment(ori.data_ptr(), Tptr, p, float(t), labels.data_ptr(), cid.data_ptr(), root.data_ptr(),
                              ws.data_ptr(), int(max_passes), _lib.ctypes.byref(ng), stream), "pf_segment")
    G = int(ng.value)
    mom = torch.zeros((max(G, 1), 10), dtype=torch.int64, device=dev)
    _lib.check(lib.pf_segment_moments(labels.data_ptr(), cid.data_ptr(), p, G, mom.data_ptr(), stream),
               "pf_segment_moments")
    roots = root[:G].cpu().numpy().astype(np.int64)
    mom_h = mom[:G].cpu().numpy()
    ori_r = ori.index_select(0, root[:G].long()).cpu().numpy() if G else np.zeros(0, np.int32)
    order = np.lexsort((roots, ori_r))
    roots, ori_r, mom_h = roots[order], ori_r[order], mom_h[order]
    h = tuple(float(v) for v in rhs.h_xyz)
    dV = h[0] * h[1] * h[2]
    n_cells = mom_h[:, 0].astype(np.int64)
    cov = _covariances(mom_h, h) if G else np.zeros((0, 3, 3))
    return Grains(labels=labels, root=roots, ori=ori_r.astype(np.int64), n_cells=n_cells, volumes=n_cells * dV,
                  aspect_ratios=aspect_ratios(cov, n_cells) if G else np.zeros(0), covariance=cov)
