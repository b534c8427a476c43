"""GPU: planned batches (zfp_hip_batch_plan / zfp_hip_batch_run) replayed from
a captured graph give, on every replay, what single zfp_compress /
zfp_decompress calls give on the arrays' current contents: stream words,
sizes and decoded values.  The arrays change between replays (eager work and
synchronisations in between), the mix holds every scalar type and mode,
cursors inside a word, variable-rate arrays (offset index kept and used) and
large fixed-rate arrays launched on their own."""

import numpy as np
import pytest

import oracle_py as O
import zfp_hip as Z
from fields import random_walk, smooth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

ZROUND = {O.ROUND_NEVER: Z.ROUND_NEVER, O.ROUND_FIRST: Z.ROUND_FIRST, O.ROUND_LAST: Z.ROUND_LAST}
PREFIX_PART = 0x5a5a5a5a5a5a5a5a


def zcfg(oc):
    return Z.ZfpConfig(oc.minbits, oc.maxbits, oc.maxprec, oc.minexp,
                       Z.ZfpRounding(ZROUND[oc.rounding], bool(oc.tight)))


def capacity_words(arr, oc, start):
    dv = (O.ctypes.c_size_t * 3)(*(list(reversed(arr.shape)) + [0] * (3 - arr.ndim)))
    cap = O.lib().orc_maximum_size(oc, O.TYPE_OF[arr.dtype], dv)
    return (cap + start // 8 + 16) // 8


def new_stream(arr, oc, start):
    buf = torch.zeros(capacity_words(arr, oc, start), dtype=torch.int64, device="cuda")
    s = Z.Stream(buf, zcfg(oc))
    if start:
        s.write_bits(PREFIX_PART & ((1 << start) - 1), start)
    return s, buf


def cases(seed=31):
    rng = np.random.default_rng(seed)
    out = []
    types = [O.INT32, O.INT64, O.FLOAT, O.DOUBLE]
    for k in range(24):
        ztype = types[k % 4]
        dims = 1 + (k // 4) % 3
        shape = tuple(int(rng.integers(3, 36)) for _ in range(dims))
        mode = k % 4
        oc = (O.fixed_rate(float(rng.integers(3, 20)), ztype, dims) if mode == 0 else
              O.fixed_precision(int(rng.integers(8, 40))) if mode == 1 else
              O.fixed_accuracy(10.0 ** float(rng.uniform(-8, -2))) if mode == 2 else O.reversible())
        out.append((shape, O.DTYPES[ztype], oc, int(rng.choice([0, 13]))))
    # large fixed-rate arrays: launches of their own inside the batch
    out.append(((72, 96, 80), np.float32, O.fixed_rate(8, O.FLOAT, 3), 0))
    out.append(((40, 64, 72), np.float64, O.fixed_rate(12, O.DOUBLE, 3), 0))
    return out


def content(shape, dtype, seed):
    if np.dtype(dtype).kind == "f":
        return smooth(shape, dtype, seed, scale=10.0 ** (seed % 5 - 2))
    return random_walk(shape, dtype, seed)


def single(arr, oc, start):
    s, buf = new_stream(arr, oc, start)
    n = s.compress(torch.from_numpy(arr).cuda())
    torch.cuda.synchronize()
    w = buf.cpu().numpy().view(np.uint64)[: n // 8].copy()
    s.rewind()
    s.rseek(start)
    out = torch.zeros(arr.shape, dtype=torch.from_numpy(arr).dtype, device="cuda")
    m = s.decompress(out)
    s.close()
    return n, w, m, out.cpu().numpy()


@pytest.mark.parametrize("tune", [True, False])
def test_planned_batch_replays_match_single_calls(tune):
    cs = cases()
    streams, bufs, src, outs = [], [], [], []
    for i, (shape, dt, oc, start) in enumerate(cs):
        s, buf = new_stream(np.zeros(shape, dt), oc, start)
        streams.append(s)
        bufs.append(buf)
        src.append(torch.from_numpy(content(shape, dt, 500 + i)).cuda())
        outs.append(torch.zeros_like(src[-1]))
    flags = Z.BATCH_DEVICE_FIELDS | (0 if tune else Z.BATCH_NO_TUNE)
    enc = Z.BatchPlan(True, streams, src, flags)
    threads, slice_, ms = enc.tiling()
    assert threads in (64, 128, 256) and slice_ in (0, 64, 128, 256)
    assert (ms > 0) == tune
    if tune:
        # the tuning runs leave each variable-rate stream's offset index, as a
        # run does, so the decompress planned next reads through it
        for s, (shape, dt, oc, start) in zip(streams, cs):
            if oc.minbits != oc.maxbits:
                idx = s.index()
                nb = int(np.prod([(n + 3) // 4 for n in shape]))
                assert idx is not None and idx[0] == start and len(idx[1]) == nb, shape
    for s, (_, _, _, start) in zip(streams, cs):
        s.rewind()
        s.rseek(start)
    dec = Z.BatchPlan(False, streams, outs, flags)
    for rep in range(3):
        # new contents, written by eager work between the replays
        for i, (shape, dt, oc, start) in enumerate(cs):
            src[i].copy_(torch.from_numpy(content(shape, dt, 900 + 17 * rep + i)))
        torch.cuda.synchronize()
        done, sizes = enc.run()
        assert done == len(cs), (rep, sizes)
        torch.cuda.synchronize()
        for o in outs:
            o.zero_()
        ddone, dsizes = dec.run()
        assert ddone == len(cs), (rep, dsizes)
        for i, (shape, dt, oc, start) in enumerate(cs):
            arr = src[i].cpu().numpy()
            n, w, m, out = single(arr, oc, start)
            ctx = f"replay {rep} case {i}: {dt.__name__} {shape} start {start}"
            assert sizes[i] == n, ctx
            got = bufs[i].cpu().numpy().view(np.uint64)[: n // 8]
            assert np.array_equal(got, w), ctx
            assert dsizes[i] == m, ctx
            u = np.uint32 if np.dtype(dt).itemsize == 4 else np.uint64
            assert np.array_equal(outs[i].cpu().numpy().view(u), out.view(u)), ctx
    enc.close()
    dec.close()
    for s in streams:
        s.close()


def test_plan_refuses_what_it_cannot_replay():
    shape, dt, oc = (20, 24), np.float32, O.fixed_precision(20)
    s, _ = new_stream(np.zeros(shape, dt), oc, 0)
    host = content(shape, dt, 7)
    with pytest.raises(ValueError):
        Z.BatchPlan(True, [s], [host], 0)  # host-resident data (pointers queried)
    t = torch.from_numpy(host).cuda()
    with pytest.raises(ValueError):
        Z.BatchPlan(True, [s, s], [t, t], Z.BATCH_DEVICE_FIELDS)  # one stream twice
    plan = Z.BatchPlan(True, [s], [t], Z.BATCH_DEVICE_FIELDS | Z.BATCH_NO_TUNE)
    done, sizes = plan.run()
    assert done == 1 and sizes[0] > 0
    s.set_config(zcfg(O.fixed_precision(12)))  # no longer the planned configuration
    done, sizes = plan.run()
    assert done == 0 and sizes == [0]
    plan.close()
    s.close()
