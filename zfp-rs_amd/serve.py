"""Continuous-batching request loop over the batched codec.

Requests (compress or decompress one device-resident array through its own
stream) arrive on an open-loop trace.  The loop admits every request that has
arrived, first come first served, while the device bytes of the admitted
requests stay within a budget (the HBM the server may use), runs them as one
batched compress and one batched decompress, and stamps each request's
completion.  Requests that arrive while a batch is on the GPU join the next
batch as soon as it returns, so the device always works on whatever has
arrived; a request too large to share the budget runs alone.  A batch of one
request (the common case at low load) goes through the single-array call,
which skips the batch call's planning and table upload.

With `inflight` > 1 and a codec that can submit without waiting, the loop
keeps up to that many batches on the device at once (each on its own HIP
stream): requests that arrive while a batch runs are admitted and submitted
as soon as a slot is free and their bytes fit beside the batches in flight,
so their work overlaps the running batch instead of queueing behind it.

`serve_one_at_a_time` replays the same trace with one zfp_compress /
zfp_decompress per request, the path the batched loop is compared with.
"""
import collections
import time


class Request:
    """One job: `kind` is "compress" (array -> stream) or "decompress"
    (stream -> array).  The stream's cursor must be where the job starts."""

    __slots__ = ("kind", "stream", "array", "nbytes", "footprint", "arrival", "done", "size")

    def __init__(self, kind, stream, array, nbytes, footprint, arrival):
        if kind not in ("compress", "decompress"):
            raise ValueError(kind)
        self.kind = kind
        self.stream = stream
        self.array = array
        self.nbytes = nbytes          # uncompressed bytes of the array
        self.footprint = footprint    # device bytes the job holds while admitted
        self.arrival = arrival        # seconds from the trace start
        self.done = None
        self.size = None              # the call's return value (bytes)


class GpuCodec:
    """The batched entry points of libzfp_hip.so; `slots` HIP streams for
    batches in flight (zfp_hip_batch_submit / zfp_hip_batch_wait)."""

    def __init__(self, slots=2):
        import torch

        import zfp_hip as Z
        self.Z = Z
        self.streams = [torch.cuda.Stream() for _ in range(slots)]

    def submit(self, compress, decompress, slot):
        """Enqueue the batch's compress and decompress on slot `slot`'s
        stream; a group the asynchronous path cannot take runs at once
        through the batch call.  Returns what wait() needs."""
        Z = self.Z
        hs = self.streams[slot % len(self.streams)].cuda_stream
        ticket = []
        for reqs, encode in ((compress, True), (decompress, False)):
            if not reqs:
                continue
            b = Z.Batch([r.stream for r in reqs], [r.array for r in reqs])
            h = Z.lib().zfp_hip_batch_submit(1 if encode else 0, b.zs, b.fin, b.n,
                                             Z.BATCH_DEVICE_FIELDS, hs)
            if h:
                ticket.append((h, reqs, b))
            else:
                self.run(reqs if encode else [], [] if encode else reqs)
        return ticket

    def ready(self, ticket):
        return all(self.Z.lib().zfp_hip_batch_ready(h) for h, _, _ in ticket)

    def wait(self, ticket):
        for h, reqs, b in ticket:
            self.Z.lib().zfp_hip_batch_wait(h, b.sizes)
            for r, n in zip(reqs, b.sizes):
                r.size = n

    def run(self, compress, decompress):
        Z = self.Z
        for reqs, fn in ((compress, Z.lib().zfp_hip_compress_batch),
                         (decompress, Z.lib().zfp_hip_decompress_batch)):
            if not reqs:
                continue
            b = Z.Batch([r.stream for r in reqs], [r.array for r in reqs])
            fn(b.zs, b.fin, b.n, b.sizes, Z.BATCH_DEVICE_FIELDS)
            for r, n in zip(reqs, b.sizes):
                r.size = n

    def run_one(self, r):
        f = self.Z.make_field(r.array)
        L = self.Z.lib()
        fn = L.zfp_compress if r.kind == "compress" else L.zfp_decompress
        r.size = fn(r.stream.zs, self.Z.ctypes.byref(f))


def admit(pending, budget):
    """Take requests from the front of `pending` while their footprints fit
    in `budget`; the first one is always taken (it runs alone if too big)."""
    batch, used = [], 0
    while pending:
        r = pending[0]
        if batch and used + r.footprint > budget:
            break
        batch.append(pending.popleft())
        used += r.footprint
    return batch, used


def serve_batched(trace, codec, budget, clock=time.perf_counter, sleep=time.sleep, inflight=1):
    """Replay `trace` (requests sorted by arrival) through the batching loop.
    Returns per-batch statistics; each request gets `done` and `size`."""
    if inflight > 1 and hasattr(codec, "submit"):
        return _serve_overlapped(trace, codec, budget, clock, sleep, inflight)
    pending = collections.deque()
    todo = collections.deque(trace)
    t0 = clock()
    stats = {"batches": 0, "max_batch": 0, "peak_bytes": 0}
    while todo or pending:
        now = clock() - t0
        while todo and todo[0].arrival <= now:
            pending.append(todo.popleft())
        if not pending:
            sleep(max(0.0, todo[0].arrival - now))
            continue
        batch, used = admit(pending, budget)
        if len(batch) == 1 and hasattr(codec, "run_one"):
            # a lone request: the single-array call, without the batch
            # call's planning and table upload
            codec.run_one(batch[0])
        else:
            codec.run([r for r in batch if r.kind == "compress"],
                      [r for r in batch if r.kind == "decompress"])
        end = clock() - t0
        for r in batch:
            r.done = end
        stats["batches"] += 1
        stats["max_batch"] = max(stats["max_batch"], len(batch))
        stats["peak_bytes"] = max(stats["peak_bytes"], used)
    return stats


def _serve_overlapped(trace, codec, budget, clock, sleep, inflight, poll=20e-6):
    """The loop with up to `inflight` batches on the device: a new batch is
    submitted whenever a slot is free and the first waiting request fits in
    the budget beside the batches in flight (or nothing is in flight);
    otherwise the loop waits for the oldest batch."""
    pending = collections.deque()
    todo = collections.deque(trace)
    flying = collections.deque()  # (ticket, batch, bytes)
    used = 0
    slot = 0
    t0 = clock()
    stats = {"batches": 0, "max_batch": 0, "peak_bytes": 0, "max_inflight": 0, "overlapped": 0}
    while todo or pending or flying:
        now = clock() - t0
        while todo and todo[0].arrival <= now:
            pending.append(todo.popleft())
        if pending and len(flying) < inflight and (not flying or pending[0].footprint <= budget - used):
            batch, nbytes = admit(pending, budget - used if flying else budget)
            if len(batch) == 1 and not flying and hasattr(codec, "run_one"):
                # a lone request on an idle device: the single-array call
                codec.run_one(batch[0])
                batch[0].done = clock() - t0
                stats["batches"] += 1
                stats["max_batch"] = max(stats["max_batch"], 1)
                stats["peak_bytes"] = max(stats["peak_bytes"], nbytes)
                continue
            ticket = codec.submit([r for r in batch if r.kind == "compress"],
                                  [r for r in batch if r.kind == "decompress"], slot)
            slot = (slot + 1) % inflight
            stats["overlapped"] += 1 if flying else 0
            flying.append((ticket, batch, nbytes))
            used += nbytes
            stats["batches"] += 1
            stats["max_batch"] = max(stats["max_batch"], len(batch))
            stats["peak_bytes"] = max(stats["peak_bytes"], used)
            stats["max_inflight"] = max(stats["max_inflight"], len(flying))
            continue
        if flying:
            # with a slot free and more requests to come, poll the oldest
            # batch rather than block on it, so that arrivals are submitted
            # while it runs
            if len(flying) < inflight and todo and not codec.ready(flying[0][0]):
                sleep(min(poll, max(0.0, todo[0].arrival - now)))
                continue
            ticket, batch, nbytes = flying.popleft()
            codec.wait(ticket)
            end = clock() - t0
            for r in batch:
                r.done = end
            used -= nbytes
            continue
        sleep(max(0.0, todo[0].arrival - now))
    return stats


def serve_one_at_a_time(trace, codec, clock=time.perf_counter, sleep=time.sleep):
    """The same trace, one call per request, in arrival order."""
    t0 = clock()
    for r in trace:
        now = clock() - t0
        if r.arrival > now:
            sleep(r.arrival - now)
        codec.run_one(r)
        r.done = clock() - t0


def summary(trace):
    """Throughput (uncompressed GB/s over the replay) and latency percentiles (ms)."""
    import numpy as np
    lat = np.array([(r.done - r.arrival) * 1e3 for r in trace])
    span = max(r.done for r in trace) - min(r.arrival for r in trace)
    total = sum(r.nbytes for r in trace)
    return {"requests": len(trace), "GBps": round(total / span / 1e9, 3),
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "span_s": round(span, 4)}
