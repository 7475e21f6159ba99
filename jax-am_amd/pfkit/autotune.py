"""Tile-shape autotuner for the fused step.

A candidate is (tile, zchunk): the workgroup's (TX, TY) column shape (pf_tile_shape)
and the number of z-planes one workgroup marches.  ``tune`` times every candidate with
HIP events on the real problem size, on a state that has already been advanced a few
steps (the kernel skips all-zero channels per wavefront, so a one-hot start would time
an unrepresentative amount of work), and records the fastest per key (dtype, K,
anisotropy, shape class).  Shape class: the grid (Nz, Ny, Nx) rounded up to powers of
two, so nearby sizes share an entry.

``lookup`` consults, in order, the user cache ($PFKIT_AUTOTUNE_CACHE, default
``_autotune_cache.json`` next to this file; written by ``tune``) and the table shipped
with the library (``tiles_mi355x.json``, produced on an MI355X by
``python -m pfkit.autotune --sweep``), else returns the library defaults (-1, 0).
"""
import argparse
import json
import os
import sys
import threading

import numpy as np
import torch

from . import _lib

_HERE = os.path.dirname(os.path.abspath(__file__))
_CACHE_PATH = os.environ.get("PFKIT_AUTOTUNE_CACHE", os.path.join(_HERE, "_autotune_cache.json"))
_TABLE_PATH = os.path.join(_HERE, "tiles_mi355x.json")
_lock = threading.Lock()
_cache = None
_table = None


def _p2(n):
    n = int(n)
    p = 1
    while p < n:
        p *= 2
    return p


def key(pf_dtype, K, aniso, Nz, Ny, Nx):
    return f"d{pf_dtype}_k{int(K)}_a{int(bool(aniso))}_{_p2(Nz)}x{_p2(Ny)}x{_p2(Nx)}"


def _read(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _load():
    global _cache, _table
    if _cache is None:
        _cache = _read(_CACHE_PATH)
    if _table is None:
        _table = _read(_TABLE_PATH).get("entries", {})
    return _cache, _table


# Shape class whose entries are used for throughput (batched) execution: when many instances
# share the GPU, parallelism is abundant and the shapes that are fastest on a large grid win,
# whatever the size of each instance.
THROUGHPUT_CLASS = (64, 256, 256)


def batch_key(pf_dtype, K, aniso):
    return f"batch_d{pf_dtype}_k{int(K)}_a{int(bool(aniso))}"


def lookup(pf_dtype, pf_args, aniso, throughput=False):
    """(tile, zchunk) for this problem: user cache, then the shipped table, then (-1, 0).
    throughput=True (instances run in a batch): the entry tuned on batches of this (dtype, K,
    anisotropy), else the large-grid class."""
    keys = []
    if throughput:
        keys.append(batch_key(pf_dtype, pf_args["num_oris"], aniso))
    dims = THROUGHPUT_CLASS if throughput else (pf_args["Nz"], pf_args["Ny"], pf_args["Nx"])
    keys.append(key(pf_dtype, pf_args["num_oris"], aniso, *dims))
    e = None
    with _lock:
        cache, table = _load()
        for k in keys:
            e = cache.get(k) or table.get(k)
            if e is not None:
                break
    if e is None:
        return -1, 0
    return int(e["tile"]), int(e["zchunk"])


def candidates(pf_dtype, K, Nz):
    lib = _lib.load()
    nt = lib.pf_tile_count(pf_dtype, K)
    zcs = sorted({z for z in (0, 4, 8, 12, 16, 24, 32, Nz) if z <= Nz})
    # the same chunks with the lowest quarter of the planes in half-length chunks (dispatched last)
    zcs += [z | _lib.ZCHUNK_TAIL for z in (8, 12, 16, 24) if z < Nz]
    return [(t, z) for t in range(max(nt, 1)) for z in zcs]


def time_candidates(solver_rhs, eta, T, t_prev, dt, reps=20, warmup=3):
    """{(tile, zchunk): ms per step} for every candidate: each runs warmup + reps consecutive
    steps from this state (ping-pong buffers, time advancing by dt), as a solve does."""
    times = {}
    stream = torch.cuda.current_stream()
    bufs = [torch.empty_like(eta), torch.empty_like(eta)]
    for cand in candidates(solver_rhs.pf_dtype, solver_rhs.K, int(solver_rhs.pf_args["Nz"])):
        tile, zc = cand
        bufs[0].copy_(eta)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for i in range(warmup + reps):
            if i == warmup:
                e0.record(stream)
            t = t_prev + i * dt
            solver_rhs.step(bufs[i % 2], bufs[(i + 1) % 2], T, t, t + dt, tile, zc)
        e1.record(stream)
        e1.synchronize()
        times[cand] = e0.elapsed_time(e1) / reps
    return times


def tune(solver_rhs, eta, T, t_prev, dt, reps=20, warmup=3, store=True):
    """Time every candidate for this instance; returns (best (tile, zchunk), {cand: ms}).
    With store=True the choice goes into the user cache."""
    times = time_candidates(solver_rhs, eta, T, t_prev, dt, reps, warmup)
    best = min(times, key=times.get)
    if store:
        a = solver_rhs.pf_args
        with _lock:
            cache, _ = _load()
            cache[key(solver_rhs.pf_dtype, solver_rhs.K, solver_rhs.aniso, a["Nz"], a["Ny"], a["Nx"])] = dict(
                tile=best[0], zchunk=best[1], ms=times[best])
            try:
                with open(_CACHE_PATH, "w") as f:
                    json.dump(cache, f, indent=1, sort_keys=True)
            except OSError:
                pass
    return best, times


# ---------------------------------------------------------------- table sweep
# Shape classes of the benchmark workloads: the 256x256x64 grid, and the classes that grids
# drawn log-uniformly between 32^3 and 256x256x64 with Nx = Ny fall into.  Each class is
# tuned on a grid of 3/4 of its upper bound per axis (at least 32).
SWEEP_CLASSES = [(64, 256, 256), (32, 32, 32), (64, 64, 64), (64, 128, 128)]


def _dtype(name):
    return {"float32": torch.float32, "bfloat16": torch.bfloat16, "float64": torch.float64}[name]


def _tune_one(dtype, K, aniso, Nz, Ny, Nx, states=(20, 400), log=None):
    """Best candidate by the summed step time on the states reached after each of ``states`` steps
    of the workload (early and developed melt pool: the per-cell cost follows the hot region)."""
    from . import workloads
    from .allen_cahn import PFSolver
    wl = workloads.make("tune", Nx, Ny, Nz, K, max(8, Nx * Ny * Nz // 8000), 1000, dtype, 0.15 if aniso else 0.0)
    s = PFSolver(wl.pf_args, wl.polycrystal, dtype=dtype, clean=False, tile=-1, zchunk=0)
    ts = wl.ts(max(states) + 1)
    eta, done, times = s.ini_cond(), 0, {}
    for n in states:
        eta = s.run(eta, ts[done:n + 1], wl.laser)
        done = n
        for cand, ms in time_candidates(s.state_rhs, eta, wl.laser, ts[n], wl.dt).items():
            times[cand] = times.get(cand, 0.0) + ms / len(states)
    best = min(times, key=times.get)
    k = key(s.state_rhs.pf_dtype, K, aniso, Nz, Ny, Nx)
    entry = dict(tile=best[0], zchunk=best[1], ms=round(times[best], 5), default_ms=round(times[(0, 0)], 5),
                 tuned_on=[Nz, Ny, Nx])
    if tuple(states) != (20, 400):
        entry["states"] = list(states)
    if log:
        log(f"{k}: tile {best[0]} zchunk {best[1]} {times[best]:.4f} ms (default {times[(0, 0)]:.4f} ms)")
    return k, entry


def _tune_batch(dtype, K, aniso, n_inst=12, warm=30, steps=60, seed=7, log=None):
    """Best (tile, zchunk) for batched execution: a batch of n_inst instances of this (dtype, K,
    anisotropy) with grids drawn log-uniformly between 32^3 and 96x96x64, advanced together by the
    batch engine (graphs of 20 steps); each candidate is timed over `steps` steps after `warm`."""
    import time
    from . import workloads
    from .batch import BatchEngine
    rng = np.random.default_rng(seed)
    wls = []
    for j in range(n_inst):
        u = rng.random()
        Nx = int(round(32 * 3.0 ** u)); Nz = int(round(32 * 2.0 ** u))
        wls.append(workloads.make(f"tb{j}", Nx, Nx, Nz, K, max(8, Nx * Nx * Nz // 8000), 10 ** 6, dtype,
                                  0.15 if aniso else 0.0, seed=seed * 100 + j, device_grains=True))
    pf = _lib.TORCH_TO_PF[_dtype(dtype)]
    lib = _lib.load()
    times = {}
    for cand in [(t, z) for t in range(max(lib.pf_tile_count(pf, K), 1)) for z in (0, 8, 16)]:
        eng = BatchEngine(max_instances=n_inst + 4)
        for w in wls:
            eng.add(w.instance(tile=cand[0], zchunk=cand[1]))
        eng.run(warm, 20)
        eng.stream.synchronize()
        t0 = time.perf_counter()
        eng.run(steps, 20)
        eng.stream.synchronize()
        times[cand] = (time.perf_counter() - t0) / steps * 1e3
        eng.close()
    best = min(times, key=times.get)
    k = batch_key(pf, K, aniso)
    entry = dict(tile=best[0], zchunk=best[1], ms=round(times[best], 5), default_ms=round(times[(0, 0)], 5),
                 tuned_on=f"batch of {n_inst} instances, 32^3 .. 96x96x64")
    if log:
        log(f"{k}: tile {best[0]} zchunk {best[1]} {times[best]:.4f} ms/step (default {times[(0, 0)]:.4f})")
    return k, entry


# shapes tuned exactly: the c1/c2 grid in FP32, BF16 and FP64, isotropic FP32, and the c0 grid
EXACT_JOBS = [("float32", 20, True, 64, 256, 256), ("bfloat16", 20, True, 64, 256, 256),
              ("float64", 20, True, 64, 256, 256), ("float32", 20, False, 64, 256, 256),
              ("float64", 10, False, 32, 64, 64)]


def retune_exact(table_path, out_path, states=(20, 400, 800), log=None):
    """Re-tune only the exact-shape entries of an existing table on the states reached after
    each of ``states`` steps (a 1000-step run spends most of its steps in the developed states)
    and write the merged table; the entries record the states they were tuned on."""
    with open(table_path) as f:
        doc = json.load(f)
    for dtype, K, aniso, Nz, Ny, Nx in EXACT_JOBS:
        k, e = _tune_one(dtype, K, aniso, Nz, Ny, Nx, states=states, log=log)
        doc["entries"][k] = e
    with open(out_path, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    return doc


def _class_jobs(Ks):
    """Size-class jobs: each class tuned on a grid of 3/4 of its upper bound per axis (at least 32)."""
    jobs = []
    for (Nz, Ny, Nx) in SWEEP_CLASSES:
        for K in Ks:
            for dtype in ("float32", "bfloat16"):
                jobs.append((dtype, K, True) + tuple(max(32, (3 * v) // 4) for v in (Nz, Ny, Nx)))
    return jobs


def retune_classes(table_path, out_path, Ks=(10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20), log=None):
    """Re-tune only the size-class entries of an existing table (exact-shape and batch entries are
    kept) and write the merged table."""
    with open(table_path) as f:
        doc = json.load(f)
    exact = {key(_lib.TORCH_TO_PF[_dtype(d)], K, an, *dims) for d, K, an, *dims in EXACT_JOBS}
    for dtype, K, aniso, Nz, Ny, Nx in _class_jobs(Ks):
        k = key(_lib.TORCH_TO_PF[_dtype(dtype)], K, aniso, Nz, Ny, Nx)
        if k in exact:
            continue
        k, e = _tune_one(dtype, K, aniso, Nz, Ny, Nx, log=log)
        doc["entries"][k] = e
    with open(out_path, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    return doc


def sweep(out_path, Ks=(10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20), log=None):
    """Tune the shipped table: c1/c2/FP64 shapes plus the batched/serving classes."""
    entries = {}
    jobs = list(EXACT_JOBS) + _class_jobs(Ks)
    for dtype, K, aniso, Nz, Ny, Nx in jobs:
        if key(_lib.TORCH_TO_PF[_dtype(dtype)], K, aniso, Nz, Ny, Nx) in entries:
            continue
        k, e = _tune_one(dtype, K, aniso, Nz, Ny, Nx, log=log)
        entries[k] = e
    for K in Ks:
        for dtype in ("float32", "bfloat16"):
            k, e = _tune_batch(dtype, K, True, log=log)
            entries[k] = e
    doc = {"device": torch.cuda.get_device_name(0), "library_abi": int(_lib.load().pf_version()),
           "method": "single-instance entries: fastest of all (tile, zchunk) candidates by the mean time of 20 "
                     "consecutive ping-pong steps (after 3 warm-up steps) run from the states reached after 20 and "
                     "400 steps of the workload's one-hot start; batch_* entries: fastest candidate for a batch "
                     "of 12 instances advanced together by the batch engine",
           "entries": entries}
    with open(out_path, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    return doc


def main(argv=None):
    ap = argparse.ArgumentParser(description="tile autotuner of the fused step")
    ap.add_argument("--sweep", action="store_true", help="tune the shipped table's shape classes")
    ap.add_argument("--exact", action="store_true",
                    help="re-tune only the exact-shape entries of the shipped table (states after 20, 400, 800 steps)")
    ap.add_argument("--classes", action="store_true",
                    help="re-tune only the size-class entries of the shipped table")
    ap.add_argument("--out", default=_TABLE_PATH)
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    if args.sweep:
        torch.cuda.set_device(0)
        sweep(args.out, log=log)
    elif args.exact:
        torch.cuda.set_device(0)
        retune_exact(_TABLE_PATH, args.out, log=log)
    elif args.classes:
        torch.cuda.set_device(0)
        retune_classes(_TABLE_PATH, args.out, log=log)


if __name__ == "__main__":
    main()
