"""Continuous-batching server for phase-field solves.

Requests arrive over time (``Request.arrival``, seconds after start).  The loop admits
every arrived request for which a batch slot and HBM budget are free, advances all
resident solves together by ``chunk`` steps (one graph replay on the engine), then
retires finished solves and frees their memory before admitting more: solves come and
go between chunks while the others keep running, and state stays on the device from
admission to completion.  Latency of a request = completion time - arrival time.

The engine is any object with the BatchEngine interface (add / remove / run / status /
state), so the scheduling logic also runs against a CPU stand-in in the tests.
"""
import collections
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    arrival: float
    instance: object            # batch.Instance (or a stand-in with .steps and .state_bytes())
    admitted: float = None
    done: float = None
    slot: int = -1
    result: object = None       # final state tensor (when keep_results)
    meta: dict = field(default_factory=dict)

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


class Server:
    def __init__(self, engine, hbm_budget_bytes, chunk=16, keep_results=False, clock=time.perf_counter,
                 sleep=time.sleep):
        self.engine = engine
        self.budget = int(hbm_budget_bytes)
        self.chunk = int(chunk)
        self.keep_results = keep_results
        self.clock = clock
        self.sleep = sleep
        self.used = 0
        self.active = {}
        self.deferred_for_memory = 0   # admissions postponed because the budget was full

    def _admit(self, pending, now):
        while pending and pending[0].arrival <= now:
            r = pending[0]
            need = r.instance.state_bytes()
            if need > self.budget:
                raise ValueError(f"request {r.rid} needs {need} bytes, more than the whole budget")
            if self.used + need > self.budget or len(self.active) >= self.engine.max_instances:
                self.deferred_for_memory += 1
                return
            pending.popleft()
            r.slot = self.engine.add(r.instance)
            r.admitted = now
            self.used += need
            self.active[r.slot] = r

    def _retire(self, now):
        steps_left, _, parity = self.engine.status()
        finished = []
        for slot, r in list(self.active.items()):
            if steps_left[slot] <= 0:
                r.done = now
                if self.keep_results:
                    r.result = self.engine.state(slot, parity[slot])
                self.engine.remove(slot)
                self.used -= r.instance.state_bytes()
                del self.active[slot]
                finished.append(r)
        return finished

    def serve(self, requests):
        """Run until every request is done; returns the requests with timestamps filled."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        t0 = self.clock()
        while pending or self.active:
            now = self.clock() - t0
            self._admit(pending, now)
            if not self.active:
                if pending:
                    self.sleep(max(0.0, pending[0].arrival - (self.clock() - t0)))
                continue
            self.engine.run(self.chunk, self.chunk)
            self._retire(self.clock() - t0)
        return requests


def summarize(requests, updates_of):
    """Throughput (updates/s over the makespan) and latency percentiles."""
    lat = np.array([r.latency for r in requests])
    t_end = max(r.done for r in requests)
    t_begin = min(r.arrival for r in requests)
    total = float(sum(updates_of(r) for r in requests))
    return dict(requests=len(requests), makespan_s=t_end - t_begin, updates=total,
                throughput=total / max(t_end - t_begin, 1e-12),
                p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
                mean_latency_s=float(lat.mean()), max_latency_s=float(lat.max()))


def poisson_arrivals(rate, n, rng):
    """n arrival times of a Poisson process of the given rate (per second)."""
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
