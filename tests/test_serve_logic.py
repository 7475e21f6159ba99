"""CPU tests of the server's admission and retirement logic against a stand-in engine
(no GPU): budget respected at every instant, FIFO admission, every request completes."""
import numpy as np
import pytest

from pfkit.serve import Request, Server, poisson_arrivals, summarize


class FakeInst:
    def __init__(self, steps, nbytes):
        self.steps = steps
        self.nbytes = nbytes
        self.N, self.K = 10, 2

    def state_bytes(self):
        return self.nbytes


class FakeEngine:
    def __init__(self, max_instances=4):
        self.max_instances = max_instances
        self.left = {}
        self.order = []
        self.peak_slots = 0

    def add(self, inst):
        slot = next(s for s in range(self.max_instances) if s not in self.left)
        self.left[slot] = inst.steps
        self.order.append(inst)
        self.peak_slots = max(self.peak_slots, len(self.left))
        return slot

    def remove(self, slot):
        del self.left[slot]

    def run(self, steps, chunk):
        for s in self.left:
            self.left[s] = max(0, self.left[s] - steps)

    def status(self):
        sl = np.full(self.max_instances, -1, dtype=np.int32)
        for s, v in self.left.items():
            sl[s] = v
        return sl, np.zeros(self.max_instances), np.zeros(self.max_instances, dtype=np.int32)

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


def test_oversized_request_is_rejected():
    eng = FakeEngine()
    clk = Clock()
    srv = Server(eng, hbm_budget_bytes=4, chunk=4, clock=clk, sleep=clk.sleep)
    with pytest.raises(ValueError):
        srv.serve([Request(rid=0, arrival=0.0, instance=FakeInst(10, 5))])


def test_idle_server_waits_for_arrivals():
    eng = FakeEngine()
    clk = Clock()
    srv = Server(eng, hbm_budget_bytes=100, chunk=8, clock=clk, sleep=clk.sleep)
    reqs = [Request(rid=0, arrival=5.0, instance=FakeInst(16, 1)), Request(rid=1, arrival=9.0, instance=FakeInst(8, 1))]
    srv.serve(reqs)
    assert reqs[0].admitted >= 5.0 and reqs[1].admitted >= 9.0
    assert reqs[0].done < reqs[1].admitted
