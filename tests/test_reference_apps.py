"""The step at the reference applications' own magnitudes (fixtures tests/golden/ref_<app>.npz,
written by tests/golden/make_reference_apps.py from applications/phase_field/*/pf_params.yaml).

CPU: the FP64 oracle reproduces each fixture, and the temperature descriptors that stand in for
each app's temperature function (Rosenthal.simple_app, RosenthalRamp.grf_app,
Quench.solidification_app, Toolpath.from_args) give the app's temperature field to 1e-12.
GPU, through the drop-in solver: FP64 steps match the oracle to 1e-12 relative (max norm) with the
temperature as a descriptor and as the app's (N, 1) array; FP32 steps stay within 1e-5; a batched
FP64 instance driven by the app's source (a per-step source table for the ramp and the toolpath)
is bitwise the solo stepper loop; the device temperature field matches the app's.
"""
import glob
import json
import os
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import pf_oracle as O
from pfkit.laser import Quench, Rosenthal, RosenthalRamp, Toolpath

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(glob.glob(os.path.join(HERE, "ref_*.npz")))
IDS = [os.path.basename(f)[4:-4] for f in FILES]


def load(path):
    z = np.load(path, allow_pickle=False)
    g = {k: z[k] for k in z.files}
    g["name"] = os.path.basename(path)[4:-4]
    g["pf"] = json.loads(str(g["pf_args"]))
    g["h"] = tuple(float(v) for v in g["h_xyz"])
    g["eta0"] = g["eta_start"].astype(np.float64)
    g["cent"] = O.centroids(g["pf"]["Nz"], g["pf"]["Ny"], g["pf"]["Nx"], g["h"])
    return g


def source(g, j):
    """The descriptor standing in for the app's temperature function at step i0 + j."""
    pf, name = g["pf"], g["name"]
    if name == "simple":
        return Rosenthal.simple_app(pf)
    if name == "grf":
        return RosenthalRamp.grf_app(pf).at(float(g["ts"][j]))
    if name == "solidification":
        return Quench.solidification_app(pf)
    return Toolpath.from_args(pf).laser(int(g["i0"]) + j)


def table(g):
    """The app's source as a per-step table (None: one fixed descriptor)."""
    if g["name"] == "grf":
        return RosenthalRamp.grf_app(g["pf"])
    if g["name"] == "multi_scan":
        return Toolpath.from_args(g["pf"])
    return None


def test_reference_app_fixtures_present():
    assert IDS == ["grf", "multi_scan", "simple", "solidification"]


@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_oracle_reproduces_reference_app(path):
    g = load(path)
    a = dict(g["pf"])
    from pfkit.config import with_defaults
    a = with_defaults(a)
    assert a["gamma"] == 1.0 and not a["force_liquid"] and a["m_g"] == 2.4e5 and a["dt"] == 2e-7
    eta = g["eta0"]
    dirs = g["grain_dirs"] if a["anisotropy"] > 0 else None
    for j in range(len(g["ts"]) - 1):
        eta = O.explicit_euler(eta, g["ts"][j], g["ts"][j + 1], g["T"][j][:, None], a, g["h"], dirs)
        if j == 0:
            np.testing.assert_allclose(eta, g["eta_step1"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(eta, g["eta_final"], rtol=0, atol=1e-14)


@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_descriptor_gives_the_app_temperature(path):
    g = load(path)
    ts = g["ts"]
    if g["name"] == "multi_scan":   # the toolpath's own times are the app's read_path times
        np.testing.assert_allclose(Toolpath.from_args(g["pf"]).step_times()[g["i0"]:g["i0"] + len(ts)], ts,
                                   rtol=0, atol=1e-18)
    if g["name"] == "grf":
        np.testing.assert_array_equal(RosenthalRamp.grf_app(g["pf"]).step_times()[g["i0"]:g["i0"] + len(ts)], ts)
    for j in range(len(ts) - 1):
        T = O.laser_T(source(g, j).as_dict(), g["cent"], ts[j])[:, 0]
        np.testing.assert_allclose(T, g["T"][j], rtol=1e-12, atol=0)


def _poly(g):
    return SimpleNamespace(mesh_h_xyz=g["h"], unique_grain_directions=g["grain_dirs"],
                           cell_ori_inds=g["cell_ori_inds"], centroids=g["cent"], unique_oris_rgb=None)


def _run(g, dtype, cuda, buffer=False):
    import torch
    from pfkit import PFSolver
    s = PFSolver(g["pf"], _poly(g), dtype=dtype, device=cuda, clean=False)
    state = (s.state_rhs.check_state(g["eta0"]), float(g["ts"][0]))
    outs = []
    for j in range(len(g["ts"]) - 1):
        T = g["T"][j][:, None] if buffer else source(g, j)
        state, sol = s.stepper(state, float(g["ts"][j + 1]), [T])
        outs.append(sol)
    torch.cuda.synchronize()
    return s, outs


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_gpu_reference_app_steps(path, cuda):
    import torch
    g = load(path)
    print(f"reference app {g['name']}", flush=True)
    scale = max(1.0, np.max(np.abs(g["eta_final"])))
    for buffer in (False, True):
        _, outs = _run(g, torch.float64, cuda, buffer=buffer)
        e1 = np.max(np.abs(outs[0].cpu().numpy() - g["eta_step1"])) / scale
        eN = np.max(np.abs(outs[-1].cpu().numpy() - g["eta_final"])) / scale
        assert e1 <= 1e-12 and eN <= 1e-12, (buffer, e1, eN)
    _, outs32 = _run(g, torch.float32, cuda)
    assert np.max(np.abs(outs32[-1].double().cpu().numpy() - g["eta_final"])) <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_gpu_reference_app_batched_and_temperature(path, cuda):
    import torch
    from pfkit.batch import BatchEngine, Instance
    g = load(path)
    s, outs = _run(g, torch.float64, cuda)
    n = len(g["ts"]) - 1
    i0 = int(g["i0"])
    tab = table(g)
    eng = BatchEngine(max_instances=2)
    if tab is None:
        inst = Instance(pf_args=g["pf"], h_xyz=g["h"], grain_dirs=g["grain_dirs"], laser=source(g, 0),
                        dtype="float64", steps=n, dt=float(g["ts"][1] - g["ts"][0]), t0=float(g["ts"][0]),
                        eta0=g["eta0"])
    else:
        # the table's steps i0 .. i0 + n, as a table of its own
        sub = SimpleNamespace(step_times=lambda: tab.step_times()[i0:i0 + n + 1], laser=lambda i: tab.laser(i0 + i))
        inst = Instance(pf_args=g["pf"], h_xyz=g["h"], grain_dirs=g["grain_dirs"], laser=None, dtype="float64",
                        steps=n, dt=0.0, eta0=g["eta0"], toolpath=sub)
    slot = eng.add(inst)
    eng.run_to_completion(chunk=2)
    _, _, par = eng.status()
    got = eng.state(slot, par[slot])
    if tab is None:
        # a fixed step: t_prev = t0 + i dt differs from the app's arange grid by rounding only
        assert np.max(np.abs(got.cpu().numpy() - g["eta_final"])) <= 1e-12 * max(1.0, np.max(np.abs(g["eta_final"])))
    else:
        assert torch.equal(got, outs[-1])
    eng.close()
    for j in range(n):
        T = s.temperature(source(g, j), float(g["ts"][j]))[:, 0]
        np.testing.assert_allclose(T, g["T"][j], rtol=1e-12, atol=0)
