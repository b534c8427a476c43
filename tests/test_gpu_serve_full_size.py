"""The served workload at its configured sizes: Poisson arrivals at 200 and
2000 requests/s, compress and decompress alternating, arrays log-uniform
from 16 KiB to 256 MiB with the batched workload's mix of dimensionality,
type and mode, fields and streams resident in HBM.

Through the continuous-batching loop (one batch at a time, and two batches in
flight on separate HIP streams), every request's result equals the request
run alone (stream words and size, or decoded values and bytes read), and the
smaller requests equal the oracle.
"""
import numpy as np
import pytest

import oracle_py as O
import serve
import zfp_hip as Z

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from synth import serving_workload  # noqa: E402


def ocfg(c):
    return O.config(c.minbits, c.maxbits, c.maxprec, c.minexp)


@pytest.mark.parametrize("rate", [200.0, 2000.0])
def test_served_workload_at_full_size(rate):
    work = serving_workload(24, rate, seed=2000 + int(rate))
    total = sum(x.numel() * x.element_size() for _, x, _, _ in work)
    print(f"[serve full size] {len(work)} requests, {total / 1e6:.0f} MB, {rate:g} req/s", flush=True)
    assert min(x.numel() * x.element_size() for _, x, _, _ in work) >= (16 << 10) * 0.5
    reqs, refs = [], []
    for kind, x, cfg, t in work:
        cap = Z.maximum_size(cfg, x)
        buf = torch.zeros((cap + 7) // 8, dtype=torch.int64, device="cuda")
        ref_buf = torch.zeros_like(buf)
        ref = Z.Stream(ref_buf, cfg)
        n_ref = ref.compress(x)
        st = Z.Stream(buf, cfg)
        nbytes = x.numel() * x.element_size()
        if kind == "decompress":
            buf.copy_(ref_buf)
            ref.rewind()
            want = torch.empty_like(x)
            m_ref = ref.decompress(want)
            reqs.append(serve.Request("decompress", st, torch.empty_like(x), nbytes, cap, t))
            refs.append((kind, m_ref, want, buf, x, cfg, ref_buf[: n_ref // 8]))
        else:
            reqs.append(serve.Request("compress", st, x, nbytes, nbytes + cap, t))
            refs.append((kind, n_ref, ref_buf[: n_ref // 8].clone(), buf, x, cfg, None))
        ref.close()
    torch.cuda.synchronize()
    codec = serve.GpuCodec()
    arrivals = [r.arrival for r in reqs]
    for inflight in (1, 2):
        for r, a, (kind, _, _, buf, _, _, code) in zip(reqs, arrivals, refs):
            r.stream.rewind()
            r.done = None
            r.arrival = a
            if kind == "compress":
                buf.zero_()
            else:
                r.array.zero_()
        torch.cuda.synchronize()
        stats = serve.serve_batched(reqs, codec, torch.cuda.mem_get_info()[0] // 2, inflight=inflight)
        torch.cuda.synchronize()
        assert stats["batches"] >= 1
        for r, (kind, n_ref, want, buf, _, _, _) in zip(reqs, refs):
            assert r.done is not None and r.size == n_ref, (kind, inflight)
            if kind == "compress":
                assert torch.equal(buf[: n_ref // 8], want)
            else:
                assert torch.equal(r.array.view(torch.uint8), want.view(torch.uint8))
        print(f"[serve full size] inflight={inflight}: {stats['batches']} batches, all equal alone",
              flush=True)
    # the requests up to 4 MiB against the oracle
    checked = 0
    for kind, n_ref, want, buf, x, cfg, code in refs:
        if x.numel() * x.element_size() > (4 << 20):
            continue
        xn = x.cpu().numpy()
        n, err, ow = O.compress(xn, ocfg(cfg))
        assert err == 0
        if kind == "compress":
            assert n == n_ref and np.array_equal(want.cpu().numpy().view(np.uint64), ow[: n // 8])
        else:
            ref, m, _ = O.decompress(ow, xn.shape, xn.dtype, ocfg(cfg))
            u = np.uint32 if xn.dtype.itemsize == 4 else np.uint64
            assert m == n_ref and np.array_equal(want.cpu().numpy().view(u), ref.view(u))
        checked += 1
    assert checked >= 3
    for r in reqs:
        r.stream.close()
