"""The batched workload at its configured size: 4096 arrays, 1-3-D, each axis
log-uniform in [3, 300] (mostly ragged, partial blocks everywhere), four
scalar types, fixed rate 4-24 / fixed precision 8-48 / fixed accuracy
1e-9..1e-2, each filled from its own seed with the k^-3 field.

Every array's stream words, size and decoded values from the batch call and
from the planned (graph-replayed) batch equal its single zfp_compress /
zfp_decompress call; 64 arrays sampled across the batch equal the oracle.
"""
import ctypes

import numpy as np
import pytest

import oracle_py as O
import zfp_hip as Z

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from synth import batch_workload  # noqa: E402

N = 4096


def ocfg(c):
    return O.config(c.minbits, c.maxbits, c.maxprec, c.minexp)


def test_batched_workload_at_full_size():
    print(f"[batch full size] generating {N} arrays", flush=True)
    work = batch_workload(N, seed=4096)
    arrays = [x for x, _, _ in work]
    cfgs = [c for _, c, _ in work]
    bufs = [torch.zeros((Z.maximum_size(c, x) + 7) // 8, dtype=torch.int64, device="cuda")
            for x, c, _ in work]
    refs = [torch.zeros_like(b) for b in bufs]
    # single calls
    rstreams = [Z.Stream(b, c) for b, c in zip(refs, cfgs)]
    want_sizes = [s.compress(x) for s, x in zip(rstreams, arrays)]
    for s in rstreams:
        s.rewind()
    want_out = [torch.zeros_like(x) for x in arrays]
    want_read = [s.decompress(o) for s, o in zip(rstreams, want_out)]
    torch.cuda.synchronize()
    print(f"[batch full size] single calls done, {sum(want_sizes) / 1e6:.1f} MB of streams", flush=True)
    # the batch call
    streams = [Z.Stream(b, c) for b, c in zip(bufs, cfgs)]
    outs = [torch.zeros_like(x) for x in arrays]
    batch = Z.Batch(streams, arrays, outs)
    assert batch.compress() == N
    assert list(batch.sizes) == want_sizes
    for b, r, n in zip(bufs, refs, want_sizes):
        assert torch.equal(b[: n // 8], r[: n // 8])
    for s in streams:
        s.rewind()
    assert batch.decompress() == N
    assert list(batch.sizes) == want_read
    for o, w in zip(outs, want_out):
        assert torch.equal(o.view(torch.uint8), w.view(torch.uint8))
    print("[batch full size] batch call equals single calls", flush=True)
    # the planned batch, replayed from its graph
    for s in streams:
        s.rewind()
    for b in bufs:
        b.zero_()
    plan = Z.BatchPlan(True, streams, arrays, Z.BATCH_DEVICE_FIELDS | Z.BATCH_NO_TUNE)
    done, sizes = plan.run()
    assert done == N and sizes == want_sizes
    for b, r, n in zip(bufs, refs, want_sizes):
        assert torch.equal(b[: n // 8], r[: n // 8])
    plan.close()
    print("[batch full size] planned batch equals single calls", flush=True)
    # 64 arrays across the batch against the oracle
    for k in range(0, N, N // 64):
        x = arrays[k].cpu().numpy()
        n, err, ow = O.compress(x, ocfg(cfgs[k]))
        assert err == 0 and n == want_sizes[k], k
        assert np.array_equal(refs[k][: n // 8].cpu().numpy().view(np.uint64), ow[: n // 8]), k
        ref, m, rerr = O.decompress(ow, x.shape, x.dtype, ocfg(cfgs[k]))
        assert rerr == 0
        u = np.uint32 if x.dtype.itemsize == 4 else np.uint64
        assert np.array_equal(want_out[k].cpu().numpy().view(u), ref.view(u)), k
    for s in streams + rstreams:
        s.close()
