"""Golden fixtures (tests/golden/golden_*.npz, written by tests/golden/make_golden.py).

The CPU tests re-run the FP64 oracle on each fixture's inputs and require its stored
output, so any change of the oracle shows up here.  The GPU tests run the HIP path through
the drop-in solver on the same inputs: FP64 within 1e-12 relative (max norm) per step and
after the whole fixture trajectory, FP32 within 2e-5 absolute, BF16 storage within the
BF16 rounding of each step (stochastic: one BF16 ulp) plus 2e-2 absolute after the trajectory, and the FP64
diagnostics (free energy, orientation volumes) within 1e-12 relative.
"""
import glob
import json
import os
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import pf_oracle as O
from pfkit.laser import Gaussian, Quench, Rosenthal

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(glob.glob(os.path.join(HERE, "golden_*.npz")))
IDS = [os.path.basename(f)[7:-4] for f in FILES]


def load(path):
    z = np.load(path, allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["pf_args"] = json.loads(str(d["pf_args"]))
    ld = json.loads(str(d["laser"]))
    kind = ld.pop("model")
    for k in ("pos0", "dir"):
        if k in ld:
            ld[k] = tuple(ld[k])
    d["laser_obj"] = {"rosenthal": Rosenthal, "gaussian": Gaussian, "quench": Quench}[kind](**ld)
    d["laser_dict"] = dict(ld, model=kind)
    return d


def test_fixtures_present():
    assert len(FILES) >= 4


@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_oracle_reproduces_fixture(path):
    g = load(path)
    a, h = g["pf_args"], tuple(g["h_xyz"])
    cent = O.centroids(a["Nz"], a["Ny"], a["Nx"], h)
    dirs = g["grain_dirs"] if a["anisotropy"] > 0 else None
    eta1 = O.run(g["eta0"], g["ts"][:2], a, h, dirs, laser=g["laser_dict"], cent=cent)
    etaN = O.run(g["eta0"], g["ts"], a, h, dirs, laser=g["laser_dict"], cent=cent)
    np.testing.assert_allclose(eta1, g["eta_step1"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(etaN, g["eta_final"], rtol=0, atol=1e-14)
    f = O.free_energy(etaN, O.laser_T(g["laser_dict"], cent, g["ts"][-1]), a, h)
    np.testing.assert_allclose(f, g["free_energy"], rtol=1e-13)
    np.testing.assert_allclose(O.volumes(etaN, h), g["volumes"], rtol=1e-13)


def _solver(g, dtype, cuda):
    from pfkit import PFSolver
    pc = SimpleNamespace(mesh_h_xyz=tuple(float(v) for v in g["h_xyz"]),
                         unique_grain_directions=g["grain_dirs"], cell_ori_inds=g["cell_ori_inds"])
    return PFSolver(dict(g["pf_args"]), pc, dtype=dtype, device=cuda, clean=False)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_gpu_f64_matches_fixture(path, cuda):
    import torch
    g = load(path)
    s = _solver(g, torch.float64, cuda)
    eta1 = s.run(g["eta0"], g["ts"][:2], g["laser_obj"]).cpu().numpy()
    etaN = s.run(g["eta0"], g["ts"], g["laser_obj"])
    assert _rel(eta1, g["eta_step1"]) <= 1e-12
    assert _rel(etaN.cpu().numpy(), g["eta_final"]) <= 1e-12
    d = s.diagnostics(etaN, g["laser_obj"], float(g["ts"][-1]))
    np.testing.assert_allclose([d["F_local"], d["F_grad"]], g["free_energy"], rtol=1e-12)
    np.testing.assert_allclose(d["volumes"], g["volumes"], rtol=1e-12, atol=1e-12)
    assert d["nonfinite"] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_gpu_f32_and_bf16_match_fixture(path, cuda):
    import torch
    g = load(path)
    s32 = _solver(g, torch.float32, cuda)
    got32 = s32.run(g["eta0"], g["ts"], g["laser_obj"]).double().cpu().numpy()
    assert np.max(np.abs(got32 - g["eta_final"])) <= 2e-5
    s16 = _solver(g, torch.bfloat16, cuda)
    e0 = torch.as_tensor(g["eta0"]).to(torch.bfloat16)
    got1 = s16.run(e0, g["ts"][:2], g["laser_obj"]).double().cpu().numpy()
    ref1 = O.run(e0.double().numpy(), g["ts"][:2], g["pf_args"], tuple(g["h_xyz"]),
                 g["grain_dirs"] if g["pf_args"]["anisotropy"] > 0 else None, laser=g["laser_dict"],
                 cent=O.centroids(g["pf_args"]["Nz"], g["pf_args"]["Ny"], g["pf_args"]["Nx"],
                                  tuple(g["h_xyz"])))
    # stochastic rounding to BF16 lands on one of the two BF16 neighbours: within 1 ulp (2^-7 relative)
    assert np.all(np.abs(got1 - ref1) <= 2.0 ** -7 * np.abs(ref1) + 2e-6)
    gotN = s16.run(e0, g["ts"], g["laser_obj"]).double().cpu().numpy()
    assert np.max(np.abs(gotN - g["eta_final"])) <= 2e-2 + 2.0 ** -7
