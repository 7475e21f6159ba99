"""Writes the fixtures ref_<app>.npz of this directory from the reference's own parameter files
(run from the repository root in the build container: ``python tests/golden/make_reference_apps.py``).

Every other fixture and test of this repository uses micrometre/microsecond units with
m_g = kappa_g = 1.  These use the reference's applications at their own SI magnitudes: the yaml
files applications/phase_field/{simple,solidification,multi_scan}/pf_params.yaml are read as data
(yaml.safe_load; m_g 2.4e5, kappa_g 2.77e-6, L0 3.5e3, Qg 1.4e5, dt 2e-7 s, domains of 1e-4 to 1e-3 m),
with the temperature each app prescribes, restated here from the app source:

  simple          Rosenthal source, Q = 25 W, x0 = 0.2 Lx, vel = 0.6 Lx / t_OFF      simple/example.py:21-37
  grf             the same with power and speed ramped over the run                 simple/example_grf.py:21-46
  solidification  quench T = T_l + 5e5 z - 1e5 t, three orientations, anisotropy 0.15  solidification/example.py:21-34,40-59
  multi_scan      Rosenthal source along the yaml laser_path, T capped at 2000 K     multi_scan/example.py:33-84

The grid of each app is coarsened to a few thousand cells over the app's own domain (cell size
domain / N per axis); gamma = 1 and no liquid reset, as allen_cahn.py runs.  Each app's solve is
developed from the one-hot start (allen_cahn.py:145-153) along the app's own time grid to a
mid-run step i0 with the C/OpenMP FP64 oracle (oracle/pf_oracle.c), the developed state is
stored (rounded to FP32 values, so that it is exact in both precisions) with the temperature
field of the next STEPS steps, and the NumPy FP64 oracle's state after those steps is stored as
the expected output.  Neper and orix are absent: the grains are a seeded Voronoi map and the
orientations seeded random rotations, except for solidification, whose three orientations are the
app's axis-angle pairs (solidification/example.py:51-56) as unit quaternions
(cos(a/2), sin(a/2) n).
"""
import json
import os
import sys

import numpy as np
import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "jax-am_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from oracle import c_oracle as C  # noqa: E402
from oracle import pf_oracle as O  # noqa: E402
from pfkit.config import with_defaults  # noqa: E402
from pfkit.polycrystal import Polycrystal  # noqa: E402

REF = "/root/reference/applications/phase_field"
STEPS = 4
APPS = {
    # name: (yaml, (Nx, Ny, Nz), grains, position of i0 in the run)
    "simple": ("simple/pf_params.yaml", (32, 8, 4), 60, 0.4),
    "grf": ("simple/pf_params.yaml", (32, 8, 4), 60, 0.5),
    "solidification": ("solidification/pf_params.yaml", (40, 2, 40), 40, 0.3),
    "multi_scan": ("multi_scan/pf_params.yaml", (32, 8, 4), 60, None),
}


def axis_angle_quat(axes, angles):
    n = np.asarray(axes, dtype=np.float64)
    n = n / np.linalg.norm(n, axis=1, keepdims=True)
    a = np.asarray(angles, dtype=np.float64)
    return np.concatenate((np.cos(a / 2)[:, None], np.sin(a / 2)[:, None] * n), axis=1)


def read_path(pf):
    """multi_scan/example.py:52-84: per-step times, source positions and unit directions."""
    lp = pf["laser_path"]
    xc, yc, zc = (np.asarray(lp[k], dtype=np.float64) for k in ("x_pos", "y_pos", "z_pos"))
    ts, ps, ds = [], [], []
    t_pre = 0.
    for i in range(len(xc) - 1):
        moving = np.array([xc[i + 1] - xc[i], yc[i + 1] - yc[i], zc[i + 1] - zc[i]])
        dist = np.linalg.norm(moving)
        t_seg = np.arange(t_pre, t_pre + dist / pf["vel"], pf["dt"])
        n = len(t_seg)
        ts.append(t_seg)
        ps.append(np.stack((np.linspace(xc[i], xc[i + 1], n), np.linspace(yc[i], yc[i + 1], n),
                            np.linspace(zc[i], zc[i + 1], n)), axis=1))
        ds.append(np.repeat((moving / dist)[None, :], n, axis=0))
        t_pre = t_pre + dist / pf["vel"]
    return np.concatenate(ts), np.concatenate(ps), np.concatenate(ds)


def temperature_fn(name, pf, cent):
    """The app's temperature of step i (at the step's start time, as each app's loop passes it)."""
    Lx, Ly, Lz = pf["domain_x"], pf["domain_y"], pf["domain_z"]
    cx, cy, cz = cent[:, 0], cent[:, 1], cent[:, 2]
    if name == "multi_scan":
        ts, pos, dirs = read_path(pf)

        def T(i):
            X, Y, Z = cx - pos[i, 0], cy - pos[i, 1], cz - pos[i, 2]
            R = np.sqrt(X**2 + Y**2 + Z**2)
            proj = X * dirs[i, 0] + Y * dirs[i, 1] + Z * dirs[i, 2]
            v = pf["T_ambient"] + 25 / (2 * np.pi * 27) / R * np.exp(-pf["vel"] / (2 * 5.2e-6) * (R + proj))
            return np.where(v > 2000., 2000., v)
        return ts, T
    ts = np.arange(0., pf["t_OFF"] + 1e-10, pf["dt"])
    if name == "solidification":
        return ts, lambda i: pf["T_liquidus"] + 5.e5 * cz - 5.e5 * 0.2 * ts[i]
    if name == "simple":
        def T(i):
            t = ts[i]
            vel = 0.6 * Lx / pf["t_OFF"]
            X, Y, Z = cx - 0.2 * Lx - vel * t, cy - 0.5 * Ly, cz - Lz
            R = np.sqrt(X**2 + Y**2 + Z**2)
            return pf["T_ambient"] + 25 / (2 * np.pi * 27) / R * np.exp(-vel / (2 * 5.2e-6) * (R + X))
        return ts, T

    def T(i):  # grf
        t = ts[i]
        vel_avg = 0.6 * Lx / pf["t_OFF"]
        Q = t / pf["t_OFF"] * (15 - 35) + 35
        vel_max, vel_min = vel_avg * 1.5, vel_avg * 0.5
        vel = t / pf["t_OFF"] * (vel_max - vel_min) + vel_min
        X, Y, Z = cx - 0.2 * Lx - (vel + vel_min) / 2. * t, cy - 0.5 * Ly, cz - Lz
        R = np.sqrt(X**2 + Y**2 + Z**2)
        return pf["T_ambient"] + Q / (2 * np.pi * 27) / R * np.exp(-vel / (2 * 5.2e-6) * (R + X))
    return ts, T


def build(name):
    ypath, (Nx, Ny, Nz), grains, frac = APPS[name]
    with open(os.path.join(REF, ypath)) as f:
        pf = yaml.safe_load(f)
    pf = {k: v for k, v in pf.items()}
    pf.update(Nx=Nx, Ny=Ny, Nz=Nz, num_grains=grains)
    h = (pf["domain_x"] / Nx, pf["domain_y"] / Ny, pf["domain_z"] / Nz)
    quat = None
    if name == "solidification":
        quat = axis_angle_quat([[1., 0., 0.], [1., 0., 0.], [1., -1., 0.]],
                               [0., np.pi / 4, np.pi / 2 - np.arccos(np.sqrt(2) / np.sqrt(3))])
    pc = Polycrystal(Nx, Ny, Nz, h, grains, pf["num_oris"], seed=pf["id"], quat=quat)
    a = with_defaults(pf)
    cent = O.centroids(Nz, Ny, Nx, h)
    ts, T = temperature_fn(name, pf, cent)
    i0 = 15998 if name == "multi_scan" else int(frac * (len(ts) - 1))
    assert i0 + STEPS < len(ts)
    dirs = pc.unique_grain_directions
    params = C.make_params(a, h, None)
    x = O.ini_cond(pc.cell_ori_inds, pf["num_oris"])
    y = np.empty_like(x)
    threads = min(8, os.cpu_count() or 1)
    for i in range(i0):
        C.step(x, y, params, dirs, ts[i], ts[i + 1], T=T(i), threads=threads)
        x, y = y, x
    eta_start = x.astype(np.float32).astype(np.float64)
    Ts = np.stack([T(i0 + j) for j in range(STEPS)])
    eta = eta_start
    outs = []
    for j in range(STEPS):
        eta = O.explicit_euler(eta, ts[i0 + j], ts[i0 + j + 1], Ts[j][:, None], a, h,
                               dirs if a["anisotropy"] > 0 else None)
        outs.append(eta)
    moved = np.count_nonzero(np.abs(outs[-1] - eta_start) > 1e-9)
    print(f"{name}: grid {Nx}x{Ny}x{Nz} K={pf['num_oris']} h={h} i0={i0} of {len(ts) - 1} steps; "
          f"{moved} values move by > 1e-9 over {STEPS} steps; T {Ts.min():.0f}..{Ts.max():.0f} K")
    return dict(pf_args=json.dumps(pf), yaml=ypath, h_xyz=np.array(h), grain_dirs=dirs,
                quat=pc.quat, cell_ori_inds=pc.cell_ori_inds, i0=i0, ts=ts[i0:i0 + STEPS + 1], T=Ts,
                eta_start=eta_start.astype(np.float32), eta_step1=outs[0], eta_final=outs[-1])


if __name__ == "__main__":
    if not C.available():
        raise SystemExit("build oracle/_build/libpforacle.so first (make -C oracle)")
    for app in APPS:
        np.savez_compressed(os.path.join(HERE, f"ref_{app}.npz"), **build(app))
