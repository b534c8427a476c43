"""GPU: requests served by the continuous-batching loop give the same result
as each request run alone (bit for bit: stream words, sizes, decoded values)."""
import numpy as np
import pytest

import zfp_hip as Z
import serve

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from synth import grf_tiles, grf_window  # noqa: E402


def make_requests(n=40, seed=3):
    rng = np.random.default_rng(seed)
    tiles = grf_tiles(seed, "cuda", sizes=(1 << 14, 256, 64))
    types = [(Z.TYPE_INT32, np.int32, 2.0 ** 16), (Z.TYPE_INT64, np.int64, 2.0 ** 32),
             (Z.TYPE_FLOAT, np.float32, 1.0), (Z.TYPE_DOUBLE, np.float64, 1.0)]
    reqs, refs = [], []
    for i in range(n):
        dims = 1 + i % 3
        shape = tuple(int(rng.integers(3, 60)) for _ in range(dims))
        zt, npdt, scale = types[i % 4]
        mode = (i // 4) % 3
        cfg = (Z.ZfpConfig.fixed_rate(float(rng.integers(4, 25)), zt, dims) if mode == 0 else
               Z.ZfpConfig.fixed_precision(int(rng.integers(8, 49))) if mode == 1 else
               Z.ZfpConfig.fixed_accuracy(10.0 ** rng.uniform(-9, -2)))
        x = grf_window(tiles, shape, 100 + i, dtype=npdt, scale=scale)
        cap = Z.maximum_size(cfg, x)
        buf = torch.zeros((cap + 7) // 8, dtype=torch.int64, device="cuda")
        st = Z.Stream(buf, cfg)
        # the result of the request run alone
        ref_buf = torch.zeros_like(buf)
        ref = Z.Stream(ref_buf, cfg)
        n_ref = ref.compress(x)
        if i % 2:  # decompress request: its stream holds the array's code
            buf.copy_(ref_buf)
            ref.rewind()
            want = torch.empty_like(x)
            ref.decompress(want)
            req = serve.Request("decompress", st, torch.empty_like(x), x.numel() * x.element_size(),
                                cap, 0.0)
            refs.append(("decompress", n_ref, want, buf))
        else:
            req = serve.Request("compress", st, x, x.numel() * x.element_size(), cap, 0.0)
            refs.append(("compress", n_ref, ref_buf[: n_ref // 8].clone(), buf))
        ref.close()
        reqs.append(req)
    return reqs, refs


@pytest.mark.parametrize("inflight", [1, 2])
@pytest.mark.parametrize("rate", [2000.0, 0.0])
def test_each_request_is_the_same_alone_and_in_the_mix(rate, inflight):
    """One batch at a time, and two batches in flight on separate HIP streams
    (submitted without waiting, overlapping on the device)."""
    reqs, refs = make_requests()
    rng = np.random.default_rng(9)
    t = 0.0
    for r in reqs:
        if rate:
            t += rng.exponential(1.0 / rate)
        r.arrival = t
    stats = serve.serve_batched(reqs, serve.GpuCodec(), budget=1 << 40, inflight=inflight)
    torch.cuda.synchronize()
    assert stats["batches"] >= 1
    for r, (kind, n_ref, want, buf) in zip(reqs, refs):
        assert r.done is not None and r.done >= r.arrival
        if kind == "compress":
            assert r.size == n_ref
            assert torch.equal(buf[: n_ref // 8], want)
        else:
            assert r.size == n_ref
            assert torch.equal(r.array, want)
    for r in reqs:
        r.stream.close()


def test_submissions_on_two_streams_match_single_calls():
    """zfp_hip_batch_submit / zfp_hip_batch_wait directly: two batches in
    flight at once, each array's bytes and size equal to its single call."""
    reqs, refs = make_requests(n=24, seed=8)
    comp = [(r, ref) for r, ref in zip(reqs, refs) if r.kind == "compress"]
    codec = serve.GpuCodec(slots=2)
    half = len(comp) // 2
    t1 = codec.submit([r for r, _ in comp[:half]], [], 0)
    t2 = codec.submit([r for r, _ in comp[half:]], [], 1)
    assert len(t1) == 1 and len(t2) == 1  # both taken by the asynchronous path
    codec.wait(t2)
    codec.wait(t1)
    torch.cuda.synchronize()
    for r, (kind, n_ref, want, buf) in comp:
        assert r.size == n_ref
        assert torch.equal(buf[: n_ref // 8], want)
    for r in reqs:
        r.stream.close()
