"""Measurements of the batched engine (64 uneven instances) and the continuous-batching
server (Poisson arrivals), used by bench.py.  Each returns a JSON-serialisable dict."""
import time

import numpy as np
import torch

from . import workloads
from .allen_cahn import PFSolver
from .batch import BatchEngine
from .serve import Request, Server, poisson_arrivals, summarize


def run_batched(log, n=64, chunk=20, seed=3):
    """c3: n instances advanced together vs the same instances one at a time (PFSolver.run,
    one eager launch per step).  The one-at-a-time results double as the reference: every
    batched final state must be bitwise equal to its solo run."""
    insts = workloads.c3_instances(n, seed=seed)
    updates = sum(w.N * w.K * w.steps for w in insts)
    solvers, eta0s = [], []
    for w in insts:
        s = PFSolver(w.pf_args, w.polycrystal, dtype="float32", clean=False)
        solvers.append(s)
        eta0s.append(s.ini_cond())
    # untimed: one step of each solver, so that every kernel (each K and tile shape) is loaded
    # before the timed loop, as the batched leg's warm-up does for its kernels
    for s, e, w in zip(solvers, eta0s, insts):
        s.run(e, w.ts(1), w.laser)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    alone = [s.run(e, w.ts(), w.laser) for s, e, w in zip(solvers, eta0s, insts)]
    torch.cuda.synchronize()
    t_alone = time.perf_counter() - t0
    log(f"  c3 one-at-a-time: {t_alone:.3f} s")
    eng = BatchEngine(max_instances=n + 8)
    slots = [eng.add(w.instance()) for w in insts]
    eng.stream.synchronize()
    eng.run(chunk, chunk)               # warm-up: graph capture for this class set, then reset
    eng.stream.synchronize()
    for sl in slots:
        eng.remove(sl)
    slots = [eng.add(w.instance()) for w in insts]
    eng.stream.synchronize()
    t0 = time.perf_counter()
    eng.run_to_completion(chunk)
    eng.stream.synchronize()
    t_batch = time.perf_counter() - t0
    log(f"  c3 batched: {t_batch:.3f} s")
    _, _, par = eng.status()
    equal = all(torch.equal(eng.state(sl, par[sl]), a) for sl, a in zip(slots, alone))
    eng.close()
    return dict(instances=n, total_updates=updates, chunk=chunk,
                batched_updates_per_s=updates / t_batch, one_at_a_time_updates_per_s=updates / t_alone,
                speedup_vs_one_at_a_time=t_alone / t_batch, batched_s=t_batch, one_at_a_time_s=t_alone,
                bitwise_equal_to_solo=bool(equal),
                classes=sorted({(w.K, w.dtype) for w in insts}))


def run_serving(log, rates=(4.0, 8.0, 16.0, 24.0, 32.0), n_per_level=24, chunk=16, budget_gb=1.0,
                sample_every=8):
    """c4: for each arrival rate, Poisson arrivals of n_per_level requests served with
    continuous batching; reports throughput and latency percentiles, and the FP64-mode error
    of a sample of the requests (each re-run alone in FP64)."""
    out = []
    for level, lam in enumerate(rates):
        rng = np.random.default_rng(1000 + level)
        arr = poisson_arrivals(lam, n_per_level, rng)
        wls = [workloads.c4_request(rng, level * 1000 + j) for j in range(n_per_level)]
        reqs = [Request(rid=j, arrival=float(arr[j]), instance=w.instance(), meta=dict(wl=w))
                for j, w in enumerate(wls)]
        torch.cuda.synchronize()
        eng = BatchEngine(max_instances=128)
        srv = Server(eng, budget_gb * 2**30, chunk=chunk,
                     keep_results=True)
        srv.serve(reqs)
        st = summarize(reqs, lambda r: r.instance.N * r.instance.K * r.instance.steps)
        st.update(rate_per_s=lam, hbm_budget_gb=budget_gb, deferred_admissions=srv.deferred_for_memory,
                  deferred_for_capacity=srv.deferred_for_capacity)
        errs = []
        for r in reqs[::sample_every]:
            if not r.ok:
                continue
            w = r.meta["wl"]
            s64 = PFSolver(w.pf_args, w.polycrystal, dtype="float64", clean=False)
            ref = s64.run(s64.ini_cond(), w.ts(), w.laser)
            got = r.result.double()
            agree = (s64.argmax(ref) == s64.argmax(r.result.to(ref.dtype))).double().mean().item()
            errs.append(dict(rid=r.rid, dtype=w.dtype, cells=w.N, K=w.K, steps=w.steps,
                             max_abs_deta=(got - ref).abs().max().item(), argmax_agreement=agree))
        for r in reqs:
            r.result = None
        st["fp64_check"] = errs
        eng.close()
        log(f"  c4 rate {lam}/s: {st['throughput']:.3e} upd/s, p50 {st['p50_latency_s']:.3f} s, "
            f"p99 {st['p99_latency_s']:.3f} s")
        out.append(st)
    return out


def run_all(log):
    return {"batched_c3": run_batched(log), "serving_c4": run_serving(log)}
