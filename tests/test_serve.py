"""The continuous-batching loop's scheduling (zfp-rs_amd/serve.py), on the CPU
with a simulated clock and codec: admission under the HBM budget, requests
arriving during a batch joining the next one, every request completed once."""
import collections

import serve


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


class FakeCodec:
    """Each batch takes 1 ms plus 1 ms per MB; one call per request takes
    0.2 ms of overhead plus the same per-byte time."""

    def __init__(self, clock):
        self.clock = clock
        self.batches = []
        self.singles = 0

    def run(self, compress, decompress):
        reqs = compress + decompress
        self.batches.append([id(r) for r in reqs])
        self.clock.t += 1e-3 + sum(r.nbytes for r in reqs) / 1e9
        for r in reqs:
            r.size = r.nbytes // 4

    def run_one(self, r):
        self.batches.append([id(r)])
        self.singles += 1
        self.clock.t += 2e-4 + r.nbytes / 1e9
        r.size = r.nbytes // 4


def make_trace(n, gap, nbytes=1_000_000):
    return [serve.Request("compress" if i % 2 else "decompress", None, None, nbytes, 2 * nbytes,
                          i * gap) for i in range(n)]


def test_admit_fills_the_budget_in_order():
    reqs = collections.deque(make_trace(5, 0.0))
    batch, used = serve.admit(reqs, budget=5_000_000)
    assert len(batch) == 2 and used == 4_000_000 and len(reqs) == 3
    big = collections.deque([serve.Request("compress", None, None, 10, 10 ** 12, 0.0)])
    batch, used = serve.admit(big, budget=1)
    assert len(batch) == 1  # a request larger than the budget runs alone


def test_requests_arriving_during_a_batch_join_the_next():
    clock = FakeClock()
    codec = FakeCodec(clock)
    trace = make_trace(40, gap=1e-4)  # 10k requests/s, 1 MB each
    stats = serve.serve_batched(trace, codec, budget=10 ** 12, clock=clock, sleep=clock.sleep)
    seen = [i for b in codec.batches for i in b]
    assert sorted(seen) == sorted(id(r) for r in trace)  # each request exactly once
    assert stats["batches"] == len(codec.batches) < 40
    assert all(r.done >= r.arrival for r in trace)
    # arrival order is kept across batches: each batch holds later arrivals
    order = {id(r): k for k, r in enumerate(trace)}
    spans = [(min(order[i] for i in b), max(order[i] for i in b)) for b in codec.batches]
    assert all(a[1] < b[0] for a, b in zip(spans, spans[1:]))
    s = serve.summary(trace)
    assert s["requests"] == 40 and s["p99_ms"] >= s["p50_ms"] > 0


def test_a_lone_request_takes_the_single_array_call():
    """A batch of one request skips the batch call's planning: the loop codes
    it with the single-array call."""
    clock = FakeClock()
    codec = FakeCodec(clock)
    trace = make_trace(6, gap=0.1)  # far apart: each arrives to an idle loop
    stats = serve.serve_batched(trace, codec, budget=10 ** 12, clock=clock, sleep=clock.sleep)
    assert stats["batches"] == 6 and codec.singles == 6
    assert all(r.size == r.nbytes // 4 and r.done >= r.arrival for r in trace)


def test_budget_bounds_the_admitted_bytes():
    clock = FakeClock()
    codec = FakeCodec(clock)
    trace = make_trace(30, gap=0.0)
    stats = serve.serve_batched(trace, codec, budget=6_000_000, clock=clock, sleep=clock.sleep)
    assert stats["peak_bytes"] <= 6_000_000 and stats["max_batch"] == 3
    assert stats["batches"] == 10


def test_one_at_a_time_is_fifo():
    clock = FakeClock()
    codec = FakeCodec(clock)
    trace = make_trace(10, gap=5e-3)
    serve.serve_one_at_a_time(trace, codec, clock=clock, sleep=clock.sleep)
    dones = [r.done for r in trace]
    assert dones == sorted(dones) and all(r.done > r.arrival for r in trace)


class FakeAsyncCodec(FakeCodec):
    """Batches run side by side on a simulated device: a submitted batch ends
    1 ms plus 1 ms per MB after its submission; waiting moves the clock there."""

    def __init__(self, clock):
        super().__init__(clock)
        self.flights = []

    def submit(self, compress, decompress, slot):
        reqs = compress + decompress
        self.batches.append([id(r) for r in reqs])
        end = self.clock.t + 1e-3 + sum(r.nbytes for r in reqs) / 1e9
        for r in reqs:
            r.size = r.nbytes // 4
        self.flights.append((self.clock.t, end))
        return end

    def ready(self, end):
        return self.clock.t >= end

    def wait(self, end):
        self.clock.t = max(self.clock.t, end)


def test_overlapped_batches_take_arrivals_while_one_runs():
    clock = FakeClock()
    codec = FakeAsyncCodec(clock)
    trace = make_trace(40, gap=4e-4)  # 2500 requests/s, 1 MB each
    stats = serve.serve_batched(trace, codec, budget=10 ** 12, clock=clock, sleep=clock.sleep,
                                inflight=2)
    seen = [i for b in codec.batches for i in b]
    assert sorted(seen) == sorted(id(r) for r in trace)  # each request exactly once
    assert stats["max_inflight"] == 2 and stats["overlapped"] > 0
    # some batch was submitted while the one before it was still running
    assert any(s2 < e1 for (s1, e1), (s2, e2) in zip(codec.flights, codec.flights[1:]))
    assert all(r.done >= r.arrival for r in trace)
    order = {id(r): k for k, r in enumerate(trace)}
    spans = [(min(order[i] for i in b), max(order[i] for i in b)) for b in codec.batches]
    assert all(a[1] < b[0] for a, b in zip(spans, spans[1:]))
    over = serve.summary(trace)
    # the same trace with one batch at a time waits longer
    for r in trace:
        r.done = None
    clock2 = FakeClock()
    serve.serve_batched(trace, FakeCodec(clock2), budget=10 ** 12, clock=clock2, sleep=clock2.sleep)
    assert over["p99_ms"] <= serve.summary(trace)["p99_ms"]


def test_overlapped_batches_keep_to_the_budget():
    clock = FakeClock()
    codec = FakeAsyncCodec(clock)
    trace = make_trace(30, gap=1e-4)  # footprints of 2 MB
    stats = serve.serve_batched(trace, codec, budget=5_000_000, clock=clock, sleep=clock.sleep,
                                inflight=2)
    assert stats["peak_bytes"] <= 5_000_000
    assert sorted(i for b in codec.batches for i in b) == sorted(id(r) for r in trace)
    big = [serve.Request("compress", None, None, 10, 10 ** 9, 0.0)]
    stats = serve.serve_batched(big, FakeAsyncCodec(FakeClock()), budget=1, inflight=2,
                                clock=FakeClock(), sleep=lambda dt: None)
    assert stats["batches"] == 1  # a request larger than the budget still runs, alone
