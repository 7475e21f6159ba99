"""GPU tests of the batched engine and the continuous-batching server: every solve's
result is bitwise the same alone (PFSolver.run, one launch per step) and in a batch."""
import numpy as np
import pytest
import torch

from pfkit import PFSolver, workloads
from pfkit.batch import BatchEngine
from pfkit.polycrystal import DevicePolycrystal, Polycrystal
from pfkit.serve import Request, Server

pytestmark = pytest.mark.gpu


def test_device_voronoi_matches_cpu(cuda):
    for dims in ((37, 21, 13), (64, 32, 16)):
        cpu = Polycrystal(*dims, (1.0, 1.0, 1.0), 40, 10, seed=7)
        dev = DevicePolycrystal(*dims, (1.0, 1.0, 1.0), 40, 10, seed=7, device=cuda)
        assert np.array_equal(dev.cell_ori_inds.cpu().numpy(), cpu.cell_ori_inds)
        assert np.array_equal(dev.unique_grain_directions, cpu.unique_grain_directions)


def _mixed(n_steps_scale=1):
    spec = [  # (Nx, Ny, Nz, K, dtype, aniso, steps)
        (40, 24, 12, 10, "float32", 0.15, 23),
        (33, 17, 9, 20, "float32", 0.15, 17),
        (24, 20, 16, 3, "float32", 0.0, 31),
        (30, 18, 10, 20, "bfloat16", 0.15, 12),
        (20, 12, 10, 25, "float32", 0.15, 9),    # generic-K kernel
        (16, 16, 8, 10, "float64", 0.15, 14),
    ]
    out = []
    for j, (nx, ny, nz, k, dt, eps, st) in enumerate(spec):
        w = workloads.make(f"t{j}", nx, ny, nz, k, 12, st * n_steps_scale, dt, eps, seed=50 + j, vel=2.0)
        out.append(w)
    return out


def _solo(w):
    s = PFSolver(w.pf_args, w.polycrystal, dtype=w.dtype, clean=False, tile=-1, zchunk=0)
    return s.run(s.ini_cond(), w.ts(), w.laser)


def test_batched_equals_solo_bitwise(cuda):
    wls = _mixed()
    solo = [_solo(w) for w in wls]
    eng = BatchEngine(max_instances=16)
    slots = [eng.add(w.instance()) for w in wls]
    eng.run_to_completion(chunk=5)       # graph replays of 5 steps + eager remainder
    steps_left, t, par = eng.status()
    for w, sl, ref in zip(wls, slots, solo):
        print(f"instance {w.name} slot {sl}", flush=True)
        assert steps_left[sl] == 0
        assert t[sl] == pytest.approx(w.ts()[-1], rel=0, abs=1e-12)
        assert torch.equal(eng.state(sl, par[sl]), ref), w.name
    eng.close()


def test_batched_zchunk_layouts_equal_solo_bitwise(cuda):
    # the same instances batched with explicit z-chunks, with and without the tail split
    from pfkit import _lib
    wls = _mixed()
    solo = [_solo(w) for w in wls]
    for zc in (3, 4 | _lib.ZCHUNK_TAIL, 8 | _lib.ZCHUNK_TAIL):
        print(f"zchunk {zc}", flush=True)
        eng = BatchEngine(max_instances=16)
        slots = [eng.add(w.instance(zchunk=zc)) for w in wls]
        eng.run_to_completion(chunk=5)
        _, _, par = eng.status()
        for w, sl, ref in zip(wls, slots, solo):
            assert torch.equal(eng.state(sl, par[sl]), ref), (w.name, zc)
        eng.close()


def test_batch_slots_come_and_go(cuda):
    wls = _mixed()
    solo = {w.name: _solo(w) for w in wls}
    eng = BatchEngine(max_instances=8)
    first = {eng.add(wls[0].instance()): wls[0], eng.add(wls[1].instance()): wls[1]}
    eng.run(6, 3)
    # admit more while the first two are half way; remove nothing yet
    later = {eng.add(w.instance()): w for w in wls[2:4]}
    eng.run_to_completion(chunk=4)
    _, _, par = eng.status()
    for sl, w in {**first, **later}.items():
        assert torch.equal(eng.state(sl, par[sl]), solo[w.name]), w.name
        eng.remove(sl)
    # reuse the freed slots
    sl = eng.add(wls[5].instance())
    eng.run_to_completion(chunk=7)
    assert torch.equal(eng.state(sl), solo[wls[5].name])
    eng.close()


def test_server_results_same_alone_and_in_the_mix(cuda):
    wls = _mixed(n_steps_scale=3)
    solo = [_solo(w) for w in wls]
    reqs = [Request(rid=j, arrival=0.002 * j, instance=w.instance()) for j, w in enumerate(wls)]
    eng = BatchEngine(max_instances=4)      # fewer slots than requests: admission waits for slots
    budget = 3 * max(r.instance.state_bytes() for r in reqs)
    srv = Server(eng, budget, chunk=4, keep_results=True)
    srv.serve(reqs)
    for r, ref in zip(reqs, solo):
        assert r.done is not None and r.latency >= 0
        assert torch.equal(r.result, ref), r.rid
    eng.close()


def test_retired_classes_free_their_index(cuda):
    # 72 distinct (dtype, K) classes one after another: more than the engine's 64 class indices,
    # so retired classes must give their index (and their cached graphs) back; every instance
    # still runs the arithmetic of its own solo run
    eng = BatchEngine(max_instances=4)
    keep = workloads.make("keep", 12, 10, 6, 7, 4, 400, "float32", 0.15, seed=90, vel=2.0)
    ks = eng.add(keep.instance())
    n = 0
    for dtype in ("float32", "bfloat16", "float64"):
        for k in range(1, 25):
            w = workloads.make(f"cls{dtype}{k}", 10, 8, 4, k, 4, 3, dtype, 0.15, seed=k, vel=2.0)
            sl = eng.add(w.instance())
            eng.run(3, 3)
            _, _, par = eng.status()
            assert torch.equal(eng.state(sl, par[sl]), _solo(w)), w.name
            eng.remove(sl)
            n += 1
    assert n == 72
    eng.run_to_completion(chunk=8)
    _, _, par = eng.status()
    assert torch.equal(eng.state(ks, par[ks]), _solo(keep))
    eng.close()
