"""Seeded polycrystal: the ``Field`` duck type the step consumes.

The reference builds its polycrystal with the external Neper binary (Voronoi
tessellation + hex meshing, jax_am/phase_field/neper.py:5-22) and orix random
orientations (utils.py:85-149).  Neither exists here, so this module generates the
same attributes directly on the regular grid:

- cells in C order (z, y, x), x fastest, centroids at (i + 1/2) h (utils.py:43-57);
- Voronoi grains: each cell takes the nearest of ``num_grains`` seeds drawn
  uniformly in the box (Neper ``-regularization 0`` Voronoi, neper.py:15-17);
- grain -> orientation: ``numpy.random.RandomState(seed).randint(num_oris, num_grains)``,
  the reference's first draw after ``onp.random.seed(1)`` (utils.py:16,39);
- orientations: uniform random unit quaternions (Shoemake's method), stored as
  (w, x, y, z) like orix ``Orientation.data``; ``unique_grain_directions`` (3, K, 3)
  are the rotated cube axes R e_x, R e_y, R e_z (utils.py:104-109, scipy convention
  with quaternion (x, y, z, w));
- ``unique_oris_rgb`` (3, K, 3): inverse-pole-figure colours for the sample x, y, z
  directions with cubic symmetry (utils.py:93-102 uses orix's TSL key; this is a
  simpler key on the same standard triangle, used only for output files).
"""
import numpy as np
from scipy.spatial import cKDTree


def random_quaternions(n, seed):
    """n uniform random unit quaternions (w, x, y, z) (K. Shoemake, Graphics Gems III, 1992)."""
    rng = np.random.default_rng(seed)
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    x, y = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2)
    z, w = b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)
    return np.stack((w, x, y, z), axis=1)


def quat_to_matrix(q):
    """Rotation matrices (n, 3, 3) of unit quaternions (w, x, y, z) (active rotation)."""
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack((
        np.stack((1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)), axis=1),
        np.stack((2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)), axis=1),
        np.stack((2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)), axis=1)), axis=1)


def grain_directions(q):
    """(3, K, 3): [axis k][orientation i] = R_i e_k (utils.py:104-109)."""
    R = quat_to_matrix(q)
    return np.stack((R[:, :, 0], R[:, :, 1], R[:, :, 2]))


def ipf_colors(q):
    """(3, K, 3) RGB for sample directions x, y, z: the crystal direction R^T v folded into the
    cubic standard triangle (sorted |components| a <= b <= c) coloured red [001], green [101],
    blue [111]."""
    R = quat_to_matrix(q)
    out = []
    for v in np.eye(3):
        d = np.abs(np.einsum('kji,j->ki', R, v))
        d = np.sort(d, axis=1)
        a, b, c = d[:, 0], d[:, 1], d[:, 2]
        rgb = np.stack((c - b, b - a, a), axis=1)
        rgb = rgb / np.maximum(np.max(rgb, axis=1, keepdims=True), 1e-12)
        out.append(rgb)
    return np.stack(out)


def cell_centroids(Nz, Ny, Nx, h_xyz):
    z, y, x = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing='ij')
    return np.stack(((x.reshape(-1) + 0.5) * h_xyz[0], (y.reshape(-1) + 0.5) * h_xyz[1],
                     (z.reshape(-1) + 0.5) * h_xyz[2]), axis=1)


class Polycrystal:
    """Attributes consumed by the step (allen_cahn.py:12,150-151; utils.py:55-62,149)."""

    def __init__(self, Nx, Ny, Nz, h_xyz, num_grains, num_oris, seed=1, quat=None, voronoi_seed=None):
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.mesh_h_xyz = tuple(float(h) for h in h_xyz)
        self.num_grains, self.num_oris = int(num_grains), int(num_oris)
        self.centroids = cell_centroids(self.Nz, self.Ny, self.Nx, self.mesh_h_xyz)
        L = np.array([self.Nx * self.mesh_h_xyz[0], self.Ny * self.mesh_h_xyz[1], self.Nz * self.mesh_h_xyz[2]])
        rng = np.random.default_rng(seed if voronoi_seed is None else voronoi_seed)
        self.seeds = rng.random((self.num_grains, 3)) * L
        _, self.cell_grain_inds = cKDTree(self.seeds).query(self.centroids)
        self.cell_grain_inds = self.cell_grain_inds.astype(np.int64)
        grain_oris_inds = np.random.RandomState(seed).randint(self.num_oris, size=self.num_grains)
        self.cell_ori_inds = np.take(grain_oris_inds, self.cell_grain_inds, axis=0).astype(np.int32)
        self.quat = random_quaternions(self.num_oris, seed) if quat is None else np.asarray(quat, dtype=np.float64)
        self.unique_grain_directions = grain_directions(self.quat)
        self.unique_oris_rgb = ipf_colors(self.quat)

    @property
    def n_cells(self):
        return self.Nx * self.Ny * self.Nz


class DevicePolycrystal:
    """The same polycrystal as ``Polycrystal`` (same seeds, grain orientations and quaternions for
    the same arguments), with the grain map computed on the GPU (pf_voronoi) and kept there:
    ``cell_ori_inds`` is a device int32 tensor.  For request-time initial states."""

    def __init__(self, Nx, Ny, Nz, h_xyz, num_grains, num_oris, seed=1, device=None):
        import torch

        from . import _lib
        from .config import make_params
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.mesh_h_xyz = tuple(float(h) for h in h_xyz)
        self.num_grains, self.num_oris = int(num_grains), int(num_oris)
        L = np.array([self.Nx * self.mesh_h_xyz[0], self.Ny * self.mesh_h_xyz[1], self.Nz * self.mesh_h_xyz[2]])
        self.seeds = np.random.default_rng(seed).random((self.num_grains, 3)) * L
        grain_oris_inds = np.random.RandomState(seed).randint(self.num_oris, size=self.num_grains)
        self.quat = random_quaternions(self.num_oris, seed)
        self.unique_grain_directions = grain_directions(self.quat)
        self.unique_oris_rgb = ipf_colors(self.quat)
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        sd = torch.as_tensor(self.seeds, dtype=torch.float64, device=dev)
        go = torch.as_tensor(grain_oris_inds.astype(np.int32), device=dev)
        out = torch.empty(self.n_cells, dtype=torch.int32, device=dev)
        p = make_params(dict(Nx=self.Nx, Ny=self.Ny, Nz=self.Nz, num_oris=self.num_oris, m_g=1.0, kappa_g=1.0,
                             L0=1.0, Qg=1.0, gas_const=1.0, T_liquidus=1.0), self.mesh_h_xyz)
        lib = _lib.load()
        _lib.check(lib.pf_voronoi(sd.data_ptr(), self.num_grains, go.data_ptr(), p, out.data_ptr(),
                                  _lib.stream_handle(dev)), "pf_voronoi")
        self.cell_ori_inds = out

    @property
    def n_cells(self):
        return self.Nx * self.Ny * self.Nz


class Field(Polycrystal):
    """Constructor in the reference's form ``Field(pf_args, ori2=None)`` (utils.py:19-26).

    Reads Nx, Ny, Nz (or derives them from domain_* and h), num_grains, num_oris and
    id (seed) from pf_args and, like the reference, writes Nx/Ny/Nz back into pf_args
    (utils.py:47-49).  ``ori2`` may be (K, 4) quaternions (w, x, y, z).
    """

    def __init__(self, pf_args, ori2=None, h=None):
        if all(k in pf_args for k in ("Nx", "Ny", "Nz")):
            Nx, Ny, Nz = pf_args["Nx"], pf_args["Ny"], pf_args["Nz"]
        else:
            Nx = round(pf_args["domain_x"] / h)
            Ny = round(pf_args["domain_y"] / h)
            Nz = round(pf_args["domain_z"] / h)
        h_xyz = (pf_args["domain_x"] / Nx, pf_args["domain_y"] / Ny, pf_args["domain_z"] / Nz) \
            if "domain_x" in pf_args else (1.0, 1.0, 1.0)
        super().__init__(Nx, Ny, Nz, h_xyz, pf_args["num_grains"], pf_args["num_oris"],
                         seed=pf_args.get("id", 1), quat=ori2)
        pf_args["Nx"], pf_args["Ny"], pf_args["Nz"] = self.Nx, self.Ny, self.Nz
        if "data_dir" in pf_args:
            self.write_quat(pf_args["data_dir"])

    def write_quat(self, data_dir):
        """utils.py:111-114: the orientations as quaternions (w, x, y, z) in numpy/pf/quat.npy."""
        import os
        folder = os.path.join(data_dir, "numpy/pf")
        os.makedirs(folder, exist_ok=True)
        path = os.path.join(folder, "quat.npy")
        np.save(path, self.quat)
        return path


def stack_layer(below, above):
    """Multi-layer build (utils.py:151-179, generate_new_mesh): the polycrystal ``above`` placed on
    top of ``below`` (same Nx, Ny, h and orientation set).  The result has Nz = below.Nz +
    above.Nz; its cells are those of ``below`` followed by those of ``above`` (C order, z slowest),
    with the upper layer's centroids shifted by below's height, as the reference concatenates
    its cell arrays."""
    import copy
    if (below.Nx, below.Ny) != (above.Nx, above.Ny) or tuple(below.mesh_h_xyz) != tuple(above.mesh_h_xyz):
        raise ValueError("layers must share Nx, Ny and the cell size")
    if below.num_oris != above.num_oris:
        raise ValueError("layers must share the orientation set")
    out = copy.copy(below)
    out.Nz = below.Nz + above.Nz
    dz = below.Nz * below.mesh_h_xyz[2]
    upper = np.array(above.centroids, dtype=np.float64)
    upper[:, 2] += dz
    out.centroids = np.concatenate((below.centroids, upper), axis=0)
    out.cell_ori_inds = np.concatenate((np.asarray(below.cell_ori_inds), np.asarray(above.cell_ori_inds)))
    if hasattr(below, "cell_grain_inds") and hasattr(above, "cell_grain_inds"):
        out.cell_grain_inds = np.concatenate((below.cell_grain_inds, above.cell_grain_inds + below.num_grains))
        out.num_grains = below.num_grains + above.num_grains
    return out
