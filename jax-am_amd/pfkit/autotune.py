"""Tile-shape autotuner for the fused step.

A candidate is (tile, zchunk): the workgroup's (TX, TY) column shape (pf_tile_shape)
and the number of z-planes one workgroup marches.  ``tune`` times every candidate with
HIP events on the real problem size and stores the fastest per key
(dtype, K, anisotropy, shape class) in a JSON cache next to this file (or at
$PFKIT_AUTOTUNE_CACHE).  ``lookup`` returns the cached choice, else the library
defaults (-1, 0).  Shape class: the grid (Nz, Ny, Nx) rounded up to powers of two, so
nearby sizes share an entry.
"""
import json
import os
import threading

import torch

from . import _lib

_CACHE_PATH = os.environ.get("PFKIT_AUTOTUNE_CACHE",
                             os.path.join(os.path.dirname(os.path.abspath(__file__)), "_autotune_cache.json"))
_lock = threading.Lock()
_cache = None


def _p2(n):
    n = int(n)
    p = 1
    while p < n:
        p *= 2
    return p


def key(pf_dtype, K, aniso, Nz, Ny, Nx):
    return f"d{pf_dtype}_k{int(K)}_a{int(bool(aniso))}_{_p2(Nz)}x{_p2(Ny)}x{_p2(Nx)}"


def _load():
    global _cache
    if _cache is None:
        try:
            with open(_CACHE_PATH) as f:
                _cache = json.load(f)
        except (OSError, ValueError):
            _cache = {}
    return _cache


def lookup(pf_dtype, pf_args, aniso):
    with _lock:
        c = _load()
        e = c.get(key(pf_dtype, pf_args["num_oris"], aniso, pf_args["Nz"], pf_args["Ny"], pf_args["Nx"]))
    if e is None:
        return -1, 0
    return int(e["tile"]), int(e["zchunk"])


def candidates(pf_dtype, K, Nz):
    lib = _lib.load()
    nt = lib.pf_tile_count(pf_dtype, K)
    zcs = sorted({z for z in (0, 4, 8, 11, 16, 22, 32, Nz) if z <= Nz})
    return [(t, z) for t in range(max(nt, 1)) for z in zcs]


def tune(solver_rhs, eta, T, t_prev, dt, reps=20, warmup=3, store=True):
    """Time every candidate for this instance; returns (best (tile, zchunk), {cand: ms})."""
    out = torch.empty_like(eta)
    times = {}
    for cand in candidates(solver_rhs.pf_dtype, solver_rhs.K, int(solver_rhs.pf_args["Nz"])):
        tile, zc = cand
        for _ in range(warmup):
            solver_rhs.step(eta, out, T, t_prev, t_prev + dt, tile, zc)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            solver_rhs.step(eta, out, T, t_prev, t_prev + dt, tile, zc)
        e1.record()
        e1.synchronize()
        times[cand] = e0.elapsed_time(e1) / reps
    best = min(times, key=times.get)
    if store:
        a = solver_rhs.pf_args
        with _lock:
            c = _load()
            c[key(solver_rhs.pf_dtype, solver_rhs.K, solver_rhs.aniso, a["Nz"], a["Ny"], a["Nx"])] = dict(
                tile=best[0], zchunk=best[1], ms=times[best])
            try:
                with open(_CACHE_PATH, "w") as f:
                    json.dump(c, f, indent=1, sort_keys=True)
            except OSError:
                pass
    return best, times
