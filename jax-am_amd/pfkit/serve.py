"""Continuous-batching server for phase-field solves.

Requests arrive over time (``Request.arrival``, seconds after start).  The loop admits
every arrived request for which a batch slot and HBM budget are free, advances all
resident solves together by ``chunk`` steps (one graph replay on the engine), waits for
that chunk, then retires finished solves and frees their memory before admitting more:
solves come and go between chunks while the others keep running, and state stays on the
device from admission to completion.  Latency of a request = completion time - arrival
time, with the completion stamped after the chunk that finished it has run on the GPU.

Failures stay with their request; the other requests keep running:
  - a request larger than the whole budget, or one the engine rejects as invalid, fails at
    admission (``status == "failed"``, ``error`` says why);
  - a solve whose state holds a non-finite value fails when it finishes (or at the next
    check every ``check_interval`` steps), as the reference aborts its run on one
    (allen_cahn.py:179-180: "Found np.inf or np.nan in pf_sol");
  - ``cancel(rid)`` (from any thread) and ``Request.deadline`` (seconds after arrival) end a
    request at the next chunk boundary (``status == "cancelled"``) and free its slot and budget.

The engine is any object with the BatchEngine interface (add / remove / run / status /
state / nonfinite / max_instances), so the scheduling logic also runs against a CPU
stand-in in the tests.
"""
import collections
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from ._lib import PF_E_STATE, PFError


@dataclass
class Request:
    rid: int
    arrival: float
    instance: object            # batch.Instance (or a stand-in with .steps and .state_bytes())
    deadline: float = None      # seconds after arrival; past it the request is cancelled
    admitted: float = None
    done: float = None          # completion (or failure / cancellation) time
    slot: int = -1
    status: str = "pending"     # pending | running | done | failed | cancelled
    error: str = None
    result: object = None       # final state tensor (when keep_results)
    meta: dict = field(default_factory=dict)

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival

    @property
    def ok(self):
        return self.status == "done"


class Server:
    def __init__(self, engine, hbm_budget_bytes, chunk=16, keep_results=False, clock=time.perf_counter,
                 sleep=time.sleep, check_interval=None):
        self.engine = engine
        self.budget = int(hbm_budget_bytes)
        self.chunk = int(chunk)
        self.keep_results = keep_results
        self.clock = clock
        self.sleep = sleep
        self.check_interval = check_interval  # steps between non-finite checks (None: at completion only)
        self.used = 0
        self.active = {}
        self.deferred_for_memory = 0    # admissions postponed because the budget was full
        self.deferred_for_capacity = 0  # admissions postponed because the engine had no room
        self._cancel = set()
        self._lock = threading.Lock()
        self._since_check = 0

    def cancel(self, rid):
        """End request `rid` at the next chunk boundary (thread-safe)."""
        with self._lock:
            self._cancel.add(rid)

    def _cancelled(self, r):
        with self._lock:
            return r.rid in self._cancel

    @staticmethod
    def _end(r, status, now, error=None):
        r.status, r.done, r.error = status, now, error

    def _admit(self, pending, now):
        while pending and pending[0].arrival <= now:
            r = pending[0]
            if self._cancelled(r):
                pending.popleft()
                self._end(r, "cancelled", now, "cancelled before admission")
                continue
            need = r.instance.state_bytes()
            if need > self.budget:
                pending.popleft()
                self._end(r, "failed", now, f"needs {need} bytes of state, more than the whole HBM budget "
                                            f"({self.budget} bytes)")
                continue
            if self.used + need > self.budget:
                self.deferred_for_memory += 1
                return
            if len(self.active) >= self.engine.max_instances:
                self.deferred_for_capacity += 1
                return
            try:
                slot = self.engine.add(r.instance)
            except PFError as e:
                if e.code == PF_E_STATE and self.active:
                    # engine tables full (slots, work list, classes): wait for retirements
                    self.deferred_for_capacity += 1
                    return
                pending.popleft()
                self._end(r, "failed", now, str(e))
                continue
            except ValueError as e:
                pending.popleft()
                self._end(r, "failed", now, str(e))
                continue
            pending.popleft()
            r.slot, r.admitted, r.status = slot, now, "running"
            self.used += need
            self.active[slot] = r

    def _release(self, slot, r):
        self.engine.remove(slot)
        self.used -= r.instance.state_bytes()
        del self.active[slot]

    def _retire(self, now, steps_left, parity):
        check_all = self.check_interval is not None and self._since_check >= self.check_interval
        if check_all:
            self._since_check = 0
        finished = []
        for slot, r in list(self.active.items()):
            if self._cancelled(r):
                self._release(slot, r)
                self._end(r, "cancelled", now, "cancelled")
            elif r.deadline is not None and now - r.arrival > r.deadline:
                self._release(slot, r)
                self._end(r, "cancelled", now, f"deadline of {r.deadline} s exceeded")
            elif steps_left[slot] <= 0 or check_all:
                bad = self.engine.nonfinite(slot, parity[slot])
                if bad:
                    self._release(slot, r)
                    self._end(r, "failed", now, f"non-finite eta: {bad} values are inf or nan")
                elif steps_left[slot] <= 0:
                    if self.keep_results:
                        r.result = self.engine.state(slot, parity[slot])
                    self._release(slot, r)
                    self._end(r, "done", now)
                else:
                    continue
            else:
                continue
            finished.append(r)
        return finished

    def serve(self, requests):
        """Run until every request has ended; returns the requests with timestamps and status filled."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        t0 = self.clock()
        while pending or self.active:
            now = self.clock() - t0
            self._admit(pending, now)
            if not self.active:
                if pending and pending[0].arrival > now:
                    self.sleep(max(0.0, pending[0].arrival - (self.clock() - t0)))
                continue
            self.engine.run(self.chunk, self.chunk)
            self._since_check += self.chunk
            steps_left, _, parity = self.engine.status()   # waits for the chunk to finish on the GPU
            self._retire(self.clock() - t0, steps_left, parity)
        return requests


def summarize(requests, updates_of):
    """Throughput (updates/s of the completed requests over the makespan) and latency percentiles
    of the completed requests; failed and cancelled requests are counted apart."""
    ok = [r for r in requests if r.ok]
    out = dict(requests=len(requests), completed=len(ok),
               failed=sum(r.status == "failed" for r in requests),
               cancelled=sum(r.status == "cancelled" for r in requests))
    if not ok:
        return out
    lat = np.array([r.latency for r in ok])
    t_end = max(r.done for r in ok)
    t_begin = min(r.arrival for r in requests)
    total = float(sum(updates_of(r) for r in ok))
    out.update(makespan_s=t_end - t_begin, updates=total, throughput=total / max(t_end - t_begin, 1e-12),
               p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
               mean_latency_s=float(lat.mean()), max_latency_s=float(lat.max()))
    return out


def poisson_arrivals(rate, n, rng):
    """n arrival times of a Poisson process of the given rate (per second)."""
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
