"""CPU tests of the serving scheduler (admission under the HBM budget,
continuous batching, completion of every request) with a simulated device
backend and a simulated clock."""
import pytest

from torchcompactradius_amd.serve import Batch, Request, Scheduler, latency_stats, poisson_requests


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


class FakeBackend:
    """A device that takes `cost_per_point` seconds per point of a batch,
    runs up to `slots` batches at once and checks the budget invariant."""

    def __init__(self, clock, slots=2, cost_per_point=1e-7, bytes_per_point=1000):
        self.clock, self.slots, self.cost, self.bpp = clock, slots, cost_per_point, bytes_per_point
        self.inflight = 0
        self.max_inflight = 0
        self.completed = []
        self.arrived = []

    def footprint(self, r):
        return r.n * self.bpp

    def arrive(self, r):
        self.arrived.append(r.rid)

    def launch(self, b: Batch):
        self.inflight += 1
        self.max_inflight = max(self.max_inflight, self.inflight)
        assert self.inflight <= self.slots
        return self.clock.t + sum(r.n for r in b.requests) * self.cost

    def ready(self, handle):
        return self.clock.t >= handle

    def wait(self, handle):
        self.clock.t = max(self.clock.t, handle)

    def complete(self, b: Batch):
        assert self.clock.t >= b.handle
        self.inflight -= 1
        for r in b.requests:
            r.result = ("done", r.rid)
            self.completed.append(r.rid)


def _run(n_req=200, rate=2000.0, budget=5e8, slots=2, max_batch=16, seed=1):
    clock = FakeClock()
    be = FakeBackend(clock, slots=slots)
    sizes = [1000 * (1 + k % 50) for k in range(64)]
    reqs = poisson_requests(n_req, rate, sizes, seed, [s * 40 for s in sizes])
    sch = Scheduler(be, budget_bytes=budget, max_batch=max_batch, clock=clock, sleep=clock.sleep)
    done = sch.run(reqs)
    return sch, be, reqs, done


def test_every_request_completes_once_in_fifo_admission_order():
    sch, be, reqs, done = _run()
    assert sorted(r.rid for r in done) == list(range(len(reqs)))
    assert len(set(be.completed)) == len(reqs)
    assert all(r.done >= r.admitted >= r.arrival for r in reqs)
    # FIFO: admission order follows arrival order
    adm = sorted(reqs, key=lambda r: (r.admitted, r.rid))
    assert [r.rid for r in adm] == sorted(r.rid for r in reqs)


def test_budget_is_never_exceeded_and_released():
    budget = 2e8
    sch, be, reqs, done = _run(budget=budget)
    assert sch.peak_in_use <= budget
    assert sch.in_use == 0
    assert be.max_inflight <= be.slots


def test_batches_form_under_load_and_respect_max_batch():
    sch, _, _, _ = _run(rate=50000.0, max_batch=8)
    assert max(sch.batches) == 8
    assert sum(sch.batches) == 200
    sch2, _, _, _ = _run(rate=10.0)     # light load: requests mostly alone
    assert sum(1 for b in sch2.batches if b == 1) > 0.9 * len(sch2.batches)


def test_request_larger_than_budget_is_refused():
    with pytest.raises(ValueError):
        _run(budget=1e5)


def test_latency_stats():
    _, _, reqs, _ = _run()
    s = latency_stats(reqs)
    assert 0 <= s["p50_ms"] <= s["p99_ms"] <= s["max_ms"]
