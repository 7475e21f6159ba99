#!/usr/bin/env python3
"""Benchmark of the fused Allen-Cahn explicit step on one MI355X.

Workload (default): 256x256x64 grid, K = 20 orientations, 500 Voronoi grains, cubic
anisotropy 0.15, moving Gaussian laser, FP32 state and arithmetic.  One step = one
launch of the fused step kernel over the whole grid (eta already in HBM).

    python bench.py [--gpus 1] [--steps 1000] [--warmup 20] [--quick]

Prints ONE JSON line on stdout (progress goes to stderr).  Metric: cell-orientation
updates per second (N*K per step).  ``roofline`` prices the step kernel against
MI355X HBM (8 TB/s) with 8 algorithmic bytes per FP32 update (read + write eta),
averaging the kernel's duration with HIP events recorded on the stream the kernel runs
on.  ``cpu_baseline`` times the C/OpenMP FP64 restatement of the reference step on
this host's cores for a bounded number of steps of the same grid.  Extra keys report
BF16 storage, the FP64 mode, a dense-state step, the 1000-step error of FP32/BF16 against
the FP64 mode, more CPU legs, and (unless --quick) the batched and serving paths.  Before the
timed FP32 steps ~0.3 s of untimed FP64-mode steps (other buffers) are enqueued, so that the
warm-up and timed steps run at the GPU's working clocks.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "jax-am_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = {"float32": 8.0, "bfloat16": 4.0, "float64": 16.0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(solver, eta0, laser, ts, warmup, steps, tile=None, zchunk=None):
    """Run `warmup` untimed steps then time exactly `steps` steps. Returns (wall_s, event_ms, final)."""
    import torch
    tile = solver.tile if tile is None else tile
    zchunk = solver.zchunk if zchunk is None else zchunk
    a = eta0.clone()
    b = torch.empty_like(a)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # the events' first record creates them (~70 us on the first timing of a process): done here,
    # before the warm-up steps, that one-time cost stays out of the timed window
    e0.record(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    e0.elapsed_time(e1)
    i = 0
    for _ in range(warmup):
        solver.state_rhs.step(a, b, laser, ts[i], ts[i + 1], tile, zchunk)
        a, b = b, a
        i += 1
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        solver.state_rhs.step(a, b, laser, ts[i], ts[i + 1], tile, zchunk)
        a, b = b, a
        i += 1
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    return wall, e0.elapsed_time(e1), a


def settle_gpu(solver, eta0, wl, seconds=0.2):
    """Enqueue ~`seconds` of the timed kernel itself (steps of the same solver on their own copy of
    the state) without waiting for it, so that the warm-up and timed steps that follow start with
    the GPU at the clocks and power state of this kernel: a few warm-up steps after an idle host
    phase otherwise time the clock ramp, and settling with other work (FP64-mode steps were tried)
    leaves the first timed steps in that work's power state."""
    import torch
    a = eta0.clone()
    b = torch.empty_like(a)
    n = max(1, int(seconds / 1.5e-4))
    ts = wl.ts(n)
    for i in range(n):
        solver.state_rhs.step(a, b, wl.laser, ts[i], ts[i + 1], solver.tile, solver.zchunk)
        a, b = b, a
    return a, b


def zchunk_layout(zchunk, nz):
    """The z-chunk argument in words: 16 | PF_ZCHUNK_TAIL on 64 planes -> "48 planes in chunks of 16,
    the lowest 16 in chunks of 8" (the rule of set_tiles in pf_abi.hip)."""
    from pfkit import _lib
    zc = zchunk & 0xFFFF
    if zc <= 0 or not zchunk & _lib.ZCHUNK_TAIL:
        return f"chunks of {zc} planes" if zc > 0 else "default chunks"
    zc = min(zc, nz)
    zs = (zc + 1) // 2
    nbot = (nz // 4 + zs - 1) // zs * zs
    if nbot > nz - zc:
        nbot = (nz - zc) // zs * zs
    if zs >= zc or nbot <= 0:
        return f"chunks of {zc} planes"
    return f"{nz - nbot} planes in chunks of {zc}, the lowest {nbot} in chunks of {zs}"


def load_traffic(name, tile, zchunk):
    """HBM bytes per launch from the rocprofv3 counter passes summarised in profiles/, or None when
    they were measured on another tile shape or z-chunk than the one timed here."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    try:
        with open(path) as f:
            tr = json.load(f)
    except (OSError, ValueError):
        return None
    if tr.get("tile") != tile or tr.get("zchunk") != zchunk:
        return None
    return tr


def cpu_legs(wl, seconds):
    """More CPU baselines (reported beside cpu_baseline): the NumPy FP64 oracle on the whole
    64x64x32, K = 10, 200-step CPU case (single process), and the C/OpenMP restatement in FP32 on
    the timed grid for ~`seconds`."""
    import numpy as np
    from oracle import c_oracle as C
    from oracle import pf_oracle as O
    from pfkit import workloads
    out = {}
    w0 = workloads.c0()
    P, h = w0.pf_args, w0.polycrystal.mesh_h_xyz
    cent = O.centroids(P["Nz"], P["Ny"], P["Nx"], h)
    eta = O.ini_cond(w0.polycrystal.cell_ori_inds, w0.K)
    ts = w0.ts()
    t0 = time.perf_counter()
    eta = O.run(eta, ts, P, h, None, laser=w0.laser.as_dict(), cent=cent)
    el = time.perf_counter() - t0
    out["c0_numpy_fp64"] = dict(value=w0.N * w0.K * w0.steps / el, unit="cell*orientation updates/s", cores=1,
                                kind="port", seconds=el,
                                sample=f"64x64x32 grid, K=10, {w0.steps} steps (the whole run), FP64, "
                                       "oracle/pf_oracle.py (NumPy)")
    if C.available():
        cores = len(os.sched_getaffinity(0))
        params = C.make_params(wl.pf_args, wl.polycrystal.mesh_h_xyz, wl.laser.as_dict())
        e = np.zeros((wl.N, wl.K), dtype=np.float32)
        e[np.arange(wl.N), wl.polycrystal.cell_ori_inds] = 1.0
        o = np.empty_like(e)
        axes = wl.polycrystal.unique_grain_directions
        ts = wl.ts(10000)
        C.step(e, o, params, axes, ts[0], ts[1], threads=cores)
        e, o = o, e
        t0 = time.perf_counter()
        n = 0
        while True:
            C.step(e, o, params, axes, ts[n + 1], ts[n + 2], threads=cores)
            e, o = o, e
            n += 1
            el = time.perf_counter() - t0
            if el >= seconds or n >= 200:
                break
        out["c1_c_fp32"] = dict(value=wl.N * wl.K * n / el, unit="cell*orientation updates/s", cores=cores,
                                kind="port", sample=f"{wl.polycrystal.Nz}x{wl.polycrystal.Ny}x{wl.polycrystal.Nx}"
                                                    f"x{wl.K} grid, FP32, {n} steps after 1 warm-up step, {el:.1f} s, "
                                                    f"OpenMP over {cores} threads (oracle/pf_oracle.c)")
    return out


def dense_state(wl, steps, warmup):
    """The FP32 step on a dense state (eta ~ U[0, 1], seeded), where no channel is zero on its
    whole stencil, so the zero-channel skip never applies."""
    import torch
    from pfkit import PFSolver
    s = PFSolver(wl.pf_args, wl.polycrystal, dtype="float32", clean=False)
    g = torch.Generator(device="cuda").manual_seed(7)
    eta0 = torch.rand((wl.N, wl.K), generator=g, device="cuda", dtype=torch.float32)
    ts = wl.ts(warmup + steps)
    w, e, _ = time_steps(s, eta0, wl.laser, ts, warmup, steps)
    k = e / 1e3 / steps
    return dict(ms_per_step=w / steps * 1e3, kernel_ms=k * 1e3, value=wl.N * wl.K * steps / w,
                roofline_frac=BYTES_PER_UPDATE["float32"] * wl.N * wl.K / k / 1e9 / HBM_PEAK_GBS,
                steps=steps, warmup=warmup, tile=s.tile, zchunk=s.zchunk,
                state="eta ~ U[0,1] (torch.rand, seed 7), FP32")


def cpu_baseline(wl, seconds):
    """C/OpenMP FP64 restatement (oracle/pf_oracle.c) on this host, bounded to ~`seconds`."""
    import numpy as np
    from oracle import c_oracle as C
    cores = len(os.sched_getaffinity(0))
    if not C.available():
        return None
    P = wl.pf_args
    params = C.make_params(P, wl.polycrystal.mesh_h_xyz, wl.laser.as_dict())
    eta = np.zeros((wl.N, wl.K))
    eta[np.arange(wl.N), wl.polycrystal.cell_ori_inds] = 1.0
    out = np.empty_like(eta)
    axes = wl.polycrystal.unique_grain_directions
    ts = wl.ts(10000)
    C.step(eta, out, params, axes, ts[0], ts[1], threads=cores)   # warm-up (page-in, thread pool)
    eta, out = out, eta
    t0 = time.perf_counter()
    n = 0
    while True:
        C.step(eta, out, params, axes, ts[n + 1], ts[n + 2], threads=cores)
        eta, out = out, eta
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 200:
            break
    return dict(value=wl.N * wl.K * n / el, unit="cell*orientation updates/s", cores=cores, kind="port",
                sample=f"{wl.polycrystal.Nz}x{wl.polycrystal.Ny}x{wl.polycrystal.Nx}x{wl.K} grid, FP64, {n} "
                       f"steps after 1 warm-up step, {el:.1f} s, OpenMP over {cores} threads "
                       f"(oracle/pf_oracle.c, {os.uname().machine})")


def accuracy(wl, steps):
    """FP32 and BF16 vs the FP64 mode after `steps` steps from the one-hot initial state."""
    import numpy as np
    import torch
    from pfkit import PFSolver
    ts = wl.ts(steps)
    res = {}
    finals = {}
    for dt in ("float64", "float32", "bfloat16"):
        s = PFSolver(wl.pf_args, wl.polycrystal, dtype=dt, clean=False)
        finals[dt] = (s, s.run(s.ini_cond(), ts, wl.laser))
    s64, ref = finals["float64"]
    am_ref = s64.argmax(ref)
    vol_ref = s64.diagnostics(ref, wl.laser, float(ts[-1]))["volumes"]
    for dt in ("float32", "bfloat16"):
        s, e = finals[dt]
        d = (e.double() - ref).abs().max().item()
        vol = s.diagnostics(e, wl.laser, float(ts[-1]))["volumes"]
        rel = np.abs(vol - vol_ref) / np.maximum(np.abs(vol_ref), 1e-300)
        agree = (s.argmax(e) == am_ref).double().mean().item()
        res[dt] = dict(steps=steps, max_abs_deta=d, max_rel_volume_err=float(rel.max()),
                       argmax_agreement=agree)
    torch.cuda.synchronize()
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--quick", action="store_true", help="step throughput + roofline + CPU baseline only")
    ap.add_argument("--only-step", action="store_true",
                    help="only the warm-up and timed FP32 steps, without clock settling (for profilers)")
    ap.add_argument("--only", choices=("bf16", "f64", "batched"), default=None,
                    help="only the timed c1 BF16 or FP64-mode steps, or only the c3 batched run (for profilers)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--tile", type=int, default=None)
    ap.add_argument("--zchunk", type=int, default=None)
    ap.add_argument("--calib-copy", type=int, default=0, metavar="N",
                    help="only run N device copies of one FP32 eta array (335.5 MB read + written each); "
                         "the profiler's byte counters on them calibrate the traffic figure")
    args = ap.parse_args()
    if args.calib_copy:
        import torch
        x = torch.ones(256 * 256 * 64 * 20, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(args.calib_copy):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy": args.calib_copy, "bytes_read_per_copy": x.numel() * 4,
                          "bytes_written_per_copy": x.numel() * 4}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the phase-field step runs on a single MI355X here.", file=sys.stderr)
        sys.exit(2)

    import torch
    from pfkit import PFSolver, workloads
    torch.cuda.set_device(0)
    if args.only == "batched":
        from pfkit import batchbench
        log("c3: 64 instances batched and one at a time...")
        r = batchbench.run_batched(log)
        print(json.dumps({"only": "batched", **{k: r[k] for k in ("batched_updates_per_s", "one_at_a_time_updates_per_s",
                                                                  "speedup_vs_one_at_a_time", "bitwise_equal_to_solo")}}))
        return
    log("building workload c1 (256x256x64x20)...")
    wl = workloads.c1()
    if args.only in ("bf16", "f64"):
        dt = {"bf16": "bfloat16", "f64": "float64"}[args.only]
        so = PFSolver(wl.pf_args, wl.polycrystal, dtype=dt, clean=False)
        ts = wl.ts(args.warmup + args.steps)
        log(f"timing {args.steps} {dt} steps after {args.warmup} warm-up steps (tile {so.tile}, zchunk {so.zchunk})")
        w, e, _ = time_steps(so, so.ini_cond(), wl.laser, ts, args.warmup, args.steps)
        print(json.dumps({"only": args.only, "ms_per_step": w / args.steps * 1e3, "kernel_ms": e / args.steps,
                          "roofline_frac": BYTES_PER_UPDATE[dt] * wl.N * wl.K / (e / 1e3 / args.steps) / 1e9
                          / HBM_PEAK_GBS, "tile": so.tile, "zchunk": so.zchunk}))
        return
    acc = None
    if not args.only_step and not args.quick:
        # the 1000-step accuracy runs (FP64, FP32, BF16) come first: besides their own result they
        # bring the GPU to its working clocks before the timed steps
        log(f"accuracy after {wl.steps} steps vs the FP64 mode...")
        acc = accuracy(wl, wl.steps)
    s = PFSolver(wl.pf_args, wl.polycrystal, dtype="float32", clean=False, tile=args.tile, zchunk=args.zchunk)
    eta0 = s.ini_cond()
    ts = wl.ts(args.warmup + args.steps)
    log(f"timing {args.steps} FP32 steps after {args.warmup} warm-up steps (tile {s.tile}, zchunk {s.zchunk})")
    # untimed, still running on the GPU when the warm-up steps are enqueued (profiler runs skip it,
    # so that every launch of the step kernel they record is a warm-up or timed step)
    keep = None if args.only_step else settle_gpu(s, eta0, wl)
    wall, ev_ms, _ = time_steps(s, eta0, wl.laser, ts, args.warmup, args.steps)
    del keep
    upd = wl.N * wl.K
    value = upd * args.steps / wall
    kern_s = ev_ms / 1e3 / args.steps
    achieved = BYTES_PER_UPDATE["float32"] * upd / kern_s / 1e9
    tr = load_traffic("c1_f32", s.tile, s.zchunk)
    out = {
        "metric": "cell*orientation updates/s (fused Allen-Cahn explicit step)",
        "value": value,
        "unit": "cell*orientation updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded Voronoi polycrystal, one-hot start, moving Gaussian laser)",
        "config": {"workload": "c1: 256x256x64 grid, K=20, 500 grains, cubic anisotropy 0.15, FP32",
                   "cells": wl.N, "orientations": wl.K, "tile": s.tile, "zchunk": s.zchunk,
                   "zchunk_layout": zchunk_layout(s.zchunk, int(wl.pf_args["Nz"])),
                   "dt_us": wl.dt, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (tr or {}).get("bytes_per_launch"),
                     "traffic_measured_at": None if tr is None else {"tile": tr["tile"], "zchunk": tr["zchunk"],
                                                                     "summary": tr.get("summary")},
                     "kernel_ms": kern_s * 1e3,
                     "algorithmic_bytes_per_launch": BYTES_PER_UPDATE["float32"] * upd},
    }
    if not args.only_step:
        log("BF16 storage throughput...")
        s2 = PFSolver(wl.pf_args, wl.polycrystal, dtype="bfloat16", clean=False)
        w2, e2, _ = time_steps(s2, s2.ini_cond(), wl.laser, ts, args.warmup, args.steps)
        k2 = e2 / 1e3 / args.steps
        out["bf16"] = {"value": upd * args.steps / w2, "ms_per_step": w2 / args.steps * 1e3,
                       "roofline_frac": BYTES_PER_UPDATE["bfloat16"] * upd / k2 / 1e9 / HBM_PEAK_GBS,
                       "tile": s2.tile, "zchunk": s2.zchunk}
        log("FP64 mode throughput...")
        s3 = PFSolver(wl.pf_args, wl.polycrystal, dtype="float64", clean=False)
        n3 = max(10, args.steps // 10)
        w3, e3, _ = time_steps(s3, s3.ini_cond(), wl.laser, ts, 3, n3)
        out["f64"] = {"value": upd * n3 / w3, "ms_per_step": w3 / n3 * 1e3,
                      "roofline_frac": BYTES_PER_UPDATE["float64"] * upd / (e3 / 1e3 / n3) / 1e9 / HBM_PEAK_GBS}
        log("FP32 step on a dense state...")
        out["dense_state"] = dense_state(wl, max(20, args.steps // 10), 5)
        if acc is None:
            log(f"accuracy after {wl.steps} steps vs the FP64 mode...")
            acc = accuracy(wl, wl.steps)
        out["accuracy_vs_f64"] = acc
    if not args.only_step:
        log(f"CPU baseline (~{args.cpu_seconds:.0f} s)...")
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_seconds)
        if not args.quick:
            log("more CPU legs (NumPy FP64 on the 64x64x32 case, C FP32 on this grid)...")
            out["cpu_legs"] = cpu_legs(wl, args.cpu_seconds / 2)
    if not args.quick and not args.only_step:
        try:
            from pfkit import batchbench
            log("batched (c3) and serving (c4)...")
            out.update(batchbench.run_all(log))
        except ImportError:
            pass
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
