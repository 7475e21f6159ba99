"""CPU tests of the server's admission and retirement logic against a stand-in engine
(no GPU): budget respected at every instant, FIFO admission, every request completes."""
import numpy as np
import pytest

from pfkit._lib import PF_E_ARG, PF_E_STATE, PFError
from pfkit.serve import Request, Server, poisson_arrivals, summarize


class FakeInst:
    def __init__(self, steps, nbytes):
        self.steps = steps
        self.nbytes = nbytes
        self.N, self.K = 10, 2

    def state_bytes(self):
        return self.nbytes


class FakeEngine:
    """CPU stand-in of BatchEngine.  `run` only enqueues; `status` is where the work "happens"
    (the clock advances by `chunk_time` per queued chunk there), as the real engine's status()
    waits for the GPU."""

    def __init__(self, max_instances=4, clock=None, chunk_time=0.0):
        self.max_instances = max_instances
        self.left = {}
        self.insts = {}
        self.order = []
        self.peak_slots = 0
        self.clock = clock
        self.chunk_time = chunk_time
        self.queued = 0

    def add(self, inst):
        if getattr(inst, "invalid", False):
            raise PFError("pf_batch_add: bad instance (code -1)", code=PF_E_ARG)
        if getattr(inst, "tiles", 0) and sum(getattr(i, "tiles", 0) for i in self.insts.values()) + inst.tiles > 8:
            raise PFError("pf_batch_add: batch work list full (max_tiles) (code -4)", code=PF_E_STATE)
        slot = next(s for s in range(self.max_instances) if s not in self.left)
        self.left[slot] = inst.steps
        self.insts[slot] = inst
        self.order.append(inst)
        self.peak_slots = max(self.peak_slots, len(self.left))
        return slot

    def remove(self, slot):
        del self.left[slot]
        del self.insts[slot]

    def run(self, steps, chunk):
        self.queued += 1
        for s in self.left:
            self.left[s] = max(0, self.left[s] - steps)

    def status(self):
        if self.clock is not None:
            self.clock.t += self.chunk_time * self.queued
        self.queued = 0
        sl = np.full(self.max_instances, -1, dtype=np.int32)
        for s, v in self.left.items():
            sl[s] = v
        return sl, np.zeros(self.max_instances), np.zeros(self.max_instances, dtype=np.int32)

    def nonfinite(self, slot, parity=None):
        return getattr(self.insts[slot], "nonfinite", 0)

    def state(self, slot, parity=None):
        return ("state", slot)


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001
        return self.t

    def sleep(self, d):
        self.t += d


def test_budget_fifo_and_completion():
    rng = np.random.default_rng(0)
    arr = poisson_arrivals(50.0, 40, rng)
    reqs = [Request(rid=j, arrival=float(arr[j]), instance=FakeInst(int(rng.integers(5, 60)), int(rng.integers(1, 5))))
            for j in range(40)]
    eng = FakeEngine(max_instances=6)
    clk = Clock()
    peak = [0]
    srv = Server(eng, hbm_budget_bytes=8, chunk=4, clock=clk, sleep=clk.sleep)
    orig_add = eng.add

    def add(inst):
        s = orig_add(inst)
        peak[0] = max(peak[0], srv.used + inst.state_bytes())
        return s

    eng.add = add
    srv.serve(reqs)
    assert all(r.done is not None and r.done >= r.arrival and r.admitted >= r.arrival for r in reqs)
    assert peak[0] <= 8 and srv.used == 0 and not srv.active
    assert [i for i in eng.order] == [r.instance for r in sorted(reqs, key=lambda r: r.arrival)]  # FIFO
    assert eng.peak_slots <= 6
    assert srv.deferred_for_memory > 0
    st = summarize(reqs, lambda r: r.instance.steps)
    assert st["p99_latency_s"] >= st["p50_latency_s"] > 0


def test_bad_requests_fail_alone():
    # an oversized request, one the engine rejects as invalid, and one whose state turns
    # non-finite each fail with a reason; the others complete
    eng = FakeEngine()
    clk = Clock()
    srv = Server(eng, hbm_budget_bytes=4, chunk=4, clock=clk, sleep=clk.sleep)
    big, bad, nan = FakeInst(10, 5), FakeInst(10, 1), FakeInst(12, 1)
    bad.invalid = True
    nan.nonfinite = 7
    reqs = [Request(rid=0, arrival=0.0, instance=FakeInst(9, 2)), Request(rid=1, arrival=0.0, instance=big),
            Request(rid=2, arrival=0.0, instance=bad), Request(rid=3, arrival=0.0, instance=nan),
            Request(rid=4, arrival=0.01, instance=FakeInst(6, 1))]
    srv.serve(reqs)
    assert [r.status for r in reqs] == ["done", "failed", "failed", "failed", "done"]
    assert "budget" in reqs[1].error and "bad instance" in reqs[2].error and "non-finite" in reqs[3].error
    assert all(r.done is not None for r in reqs) and srv.used == 0 and not srv.active
    st = summarize(reqs, lambda r: r.instance.steps)
    assert st["completed"] == 2 and st["failed"] == 3


def test_nonfinite_found_early_with_check_interval():
    eng = FakeEngine()
    clk = Clock()
    srv = Server(eng, hbm_budget_bytes=10, chunk=4, clock=clk, sleep=clk.sleep, check_interval=8)
    nan = FakeInst(1000, 1)
    nan.nonfinite = 1
    reqs = [Request(rid=0, arrival=0.0, instance=nan)]
    srv.serve(reqs)
    assert reqs[0].status == "failed" and eng.queued == 0 and len(eng.order) == 1
    assert srv.used == 0


def test_engine_capacity_defers_instead_of_failing():
    # the engine's work list holds 8 tiles: the third request waits for a retirement
    eng = FakeEngine(max_instances=8)
    clk = Clock()
    srv = Server(eng, hbm_budget_bytes=100, chunk=4, clock=clk, sleep=clk.sleep)
    insts = [FakeInst(8, 1), FakeInst(16, 1), FakeInst(4, 1)]
    for i in insts:
        i.tiles = 4
    reqs = [Request(rid=j, arrival=0.0, instance=i) for j, i in enumerate(insts)]
    srv.serve(reqs)
    assert all(r.status == "done" for r in reqs)
    assert srv.deferred_for_capacity > 0 and reqs[2].admitted >= reqs[0].done


def test_cancel_and_deadline_free_the_slot():
    eng = FakeEngine(max_instances=2)
    clk = Clock()
    srv = Server(eng, hbm_budget_bytes=2, chunk=4, clock=clk, sleep=clk.sleep)
    long_ = Request(rid=0, arrival=0.0, instance=FakeInst(10 ** 6, 1))
    late = Request(rid=1, arrival=0.0, instance=FakeInst(10 ** 6, 1), deadline=0.05)
    after = Request(rid=2, arrival=0.02, instance=FakeInst(8, 1))
    queued = Request(rid=3, arrival=0.03, instance=FakeInst(8, 1))
    orig_run = eng.run

    def run(steps, chunk):
        if clk.t > 0.2:
            srv.cancel(0)
        srv.cancel(3)
        orig_run(steps, chunk)

    eng.run = run
    srv.serve([long_, late, after, queued])
    assert long_.status == "cancelled" and late.status == "cancelled" and "deadline" in late.error
    assert after.status == "done" and after.admitted >= late.done
    assert queued.status == "cancelled" and queued.admitted is None
    assert srv.used == 0 and not eng.left


def test_latency_includes_the_last_chunk():
    # completion is stamped after status() has waited for the chunk that finished the request
    clk = Clock()
    eng = FakeEngine(clock=clk, chunk_time=1.0)
    srv = Server(eng, hbm_budget_bytes=10, chunk=4, clock=clk, sleep=clk.sleep)
    reqs = [Request(rid=0, arrival=0.0, instance=FakeInst(8, 1))]
    srv.serve(reqs)
    assert reqs[0].latency >= 2.0


def test_idle_server_waits_for_arrivals():
    eng = FakeEngine()
    clk = Clock()
    srv = Server(eng, hbm_budget_bytes=100, chunk=8, clock=clk, sleep=clk.sleep)
    reqs = [Request(rid=0, arrival=5.0, instance=FakeInst(16, 1)), Request(rid=1, arrival=9.0, instance=FakeInst(8, 1))]
    srv.serve(reqs)
    assert reqs[0].admitted >= 5.0 and reqs[1].admitted >= 9.0
    assert reqs[0].done < reqs[1].admitted
