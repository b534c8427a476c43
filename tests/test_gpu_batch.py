"""GPU: the batched entry points (zfp_hip_compress_batch /
zfp_hip_decompress_batch) give, for every array, what single zfp_compress /
zfp_decompress calls give — stream words, returned sizes, cursors, decoded
values — and those match the oracle.  The batch mixes scalar types,
dimensionalities, modes, ragged shapes and cursors inside a word, and includes
the cases the batch hands to the single-array path (host data, a stream used
twice) and index-less variable-rate streams (decoded by the batch's walk).
"""

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
# the oracle's prefix convention (oracle_py.compress with start_bits)
PREFIX_WORD = 0x0123456789abcdef
PREFIX_PART = 0x5a5a5a5a5a5a5a5a

def zcfg(oc):
    return Z.ZfpConfig(oc.minbits, oc.maxbits, oc.maxprec, oc.minexp,
                       Z.ZfpRounding(ZROUND[oc.rounding], bool(oc.tight)))

def make_cases(seed=11, n=48):
    rng = np.random.default_rng(seed)
    cases = []
    types = [O.INT32, O.INT64, O.FLOAT, O.DOUBLE]
    for k in range(n):
        ztype = types[k % 4]
        dims = 1 + (k // 4) % 3
        shape = tuple(int(rng.integers(3, 40)) for _ in range(dims))
        dt = O.DTYPES[ztype]
        if np.dtype(dt).kind == "f":
            arr = smooth(shape, dt, 1000 + k, scale=10.0 ** int(rng.integers(-3, 4)))
        else:
            arr = random_walk(shape, dt, 1000 + k)
        mode = k % 6
        if mode == 0:
            oc = O.fixed_rate(float(rng.integers(2, 25)), ztype, dims)
        elif mode == 1:
            oc = O.fixed_rate(8, ztype, dims, align=True)
        elif mode == 2:
            oc = O.fixed_precision(int(rng.integers(8, 49)))
        elif mode == 3:
            oc = O.fixed_accuracy(10.0 ** float(rng.uniform(-9, -2)))
        elif mode == 4:
            oc = O.reversible()
        else:
            rnd = [O.ROUND_FIRST, O.ROUND_LAST][k % 2]
            oc = O.config(1, 16658, int(rng.integers(10, 40)), -1074, rnd, bool(k % 3 == 0))
        start = int(rng.choice([0, 0, 13, 64, 77]))
        cases.append((arr, oc, start))
    return cases

def capacity_words(arr, oc, start):
    dv = (O.ctypes.c_size_t * 3)(*(list(reversed(arr.shape)) + [0] * (3 - arr.ndim)))
    cap = O.lib().orc_maximum_size(oc, O.TYPE_OF[arr.dtype], dv)
    return (cap + start // 8 + 16) // 8

def new_stream(arr, oc, start, nwords=None):
    buf = torch.zeros(nwords or capacity_words(arr, oc, start), dtype=torch.int64, device="cuda")
    s = Z.Stream(buf, zcfg(oc))
    for _ in range(start // 64):
        s.write_bits(PREFIX_WORD, 64)
    if start % 64:
        s.write_bits(PREFIX_PART & ((1 << (start % 64)) - 1), start % 64)
    return s, buf

def words_of(buf, nbytes):
    torch.cuda.synchronize()
    return buf.cpu().numpy().view(np.uint64)[: nbytes // 8]

def single_round_trip(arr, oc, start):
    s, buf = new_stream(arr, oc, start)
    n = s.compress(torch.from_numpy(arr).cuda())
    w = words_of(buf, n)
    s.rewind()
    s.rseek(start)
    out = torch.zeros(arr.shape, dtype=torch.from_numpy(arr).dtype, device="cuda")
    m = s.decompress(out)
    s.close()
    return n, w, m, out.cpu().numpy()

def test_batch_matches_single_calls_and_oracle():
    cases = make_cases()
    streams, bufs, src = [], [], []
    for arr, oc, start in cases:
        s, buf = new_stream(arr, oc, start)
        streams.append(s)
        bufs.append(buf)
        src.append(torch.from_numpy(arr).cuda())
    done, sizes = Z.compress_batch(streams, src)
    assert done == len(cases), sizes
    outs = [torch.zeros_like(t) for t in src]
    for s, (_, _, start) in zip(streams, cases):
        s.rewind()
        s.rseek(start)
    ddone, dsizes = Z.decompress_batch(streams, outs, Z.BATCH_DEVICE_FIELDS)
    assert ddone == len(cases), dsizes
    for i, (arr, oc, start) in enumerate(cases):
        ctx = f"case {i}: {arr.dtype} {arr.shape} {oc.minbits},{oc.maxbits},{oc.maxprec},{oc.minexp} start {start}"
        n, w, m, out = single_round_trip(arr, oc, start)
        assert sizes[i] == n, ctx
        assert np.array_equal(words_of(bufs[i], n), w), ctx
        assert dsizes[i] == m, ctx
        got = outs[i].cpu().numpy()
        u = np.uint32 if arr.dtype.itemsize == 4 else np.uint64
        assert np.array_equal(got.view(u), out.view(u)), ctx
        on, oerr, ow = O.compress(arr, oc, start_bits=start)
        assert oerr == 0 and on == n and np.array_equal(ow[: on // 8], w), ctx
        ref, rn, rerr = O.decompress(ow, arr.shape, arr.dtype, oc, start_bits=start)
        assert rerr == 0 and rn == m and np.array_equal(ref.view(u), got.view(u)), ctx
    for s in streams:
        s.close()

def test_batch_hands_over_what_it_cannot_take():
    """A host-resident array and a stream used twice go to the single-array
    calls, in input order; an index-less stream is walked; an undersized
    stream fails as zfp_compress does."""
    cases = make_cases(seed=5, n=8)
    arr0, oc0, _ = cases[3]  # variable-rate
    arr1, oc1, _ = cases[2]
    # reference: sequential single calls into one stream
    ref_s, ref_buf = new_stream(arr0, oc0, 0, nwords=4 * capacity_words(arr0, oc0, 0))
    n_a = ref_s.compress(torch.from_numpy(arr0).cuda())
    ref_s.set_config(zcfg(oc0))
    n_b = ref_s.compress(torch.from_numpy(arr0[..., ::-1].copy()).cuda())
    ref_words = words_of(ref_buf, n_b)
    ref_s.close()
    # batch: the same stream twice, a host array, an undersized stream
    dup, dup_buf = new_stream(arr0, oc0, 0, nwords=4 * capacity_words(arr0, oc0, 0))
    host_s, host_buf = new_stream(arr1, oc1, 13)
    assert n_a >= 16
    small, _ = new_stream(arr0, oc0, 0, nwords=n_a // 8 - 1)  # one word short
    arrays = [torch.from_numpy(arr0).cuda(), arr1, torch.from_numpy(arr0[..., ::-1].copy()).cuda(),
              torch.from_numpy(arr0).cuda()]
    done, sizes = Z.compress_batch([dup, host_s, dup, small], arrays)
    assert done == 3 and sizes[0] == n_a and sizes[2] == n_b and sizes[3] == 0
    assert np.array_equal(words_of(dup_buf, n_b), ref_words)
    hn, hw, _, _ = single_round_trip(arr1, oc1, 13)
    assert sizes[1] == hn and np.array_equal(words_of(host_buf, hn), hw)
    # index-less decode: a fresh handle over the same words has no index
    fresh = Z.Stream(dup_buf, zcfg(oc0))
    out = torch.zeros_like(arrays[0])
    d, ds = Z.decompress_batch([fresh], [out])
    assert d == 1 and ds[0] == n_a
    assert np.array_equal(out.cpu().numpy(), single_round_trip(arr0, oc0, 0)[3])
    for s in (dup, host_s, small, fresh):
        s.close()

@pytest.mark.parametrize("rate", [4, 8, 16])
def test_batch_of_fixed_rate_f32_cubes(rate):
    """Many same-class arrays: one launch covers them all."""
    arrays, streams, bufs = [], [], []
    oc = O.fixed_rate(rate, O.FLOAT, 3)
    for k in range(32):
        shape = (8 + k % 5, 12, 4 + 3 * (k % 7))
        arr = smooth(shape, np.float32, 77 + k, scale=1.0)
        s, buf = new_stream(arr, oc, 0)
        arrays.append(torch.from_numpy(arr).cuda())
        streams.append(s)
        bufs.append(buf)
    done, sizes = Z.compress_batch(streams, arrays, Z.BATCH_DEVICE_FIELDS)
    assert done == 32
    for k, t in enumerate(arrays):
        n, w, _, _ = single_round_trip(t.cpu().numpy(), oc, 0)
        assert sizes[k] == n and np.array_equal(words_of(bufs[k], n), w)
    for s in streams:
        s.close()


@pytest.mark.parametrize("rate,start", [(8, 0), (5, 13), (16, 64)])
def test_batch_with_large_fixed_rate_arrays(rate, start):
    """Large fixed-rate arrays in a batch take launches of their own (the
    LDS-staged kernels); their bytes and values match single calls."""
    oc = O.fixed_rate(rate, O.FLOAT, 3)
    shapes = [(64, 128, 128), (5, 6, 7), (72, 100, 70), (9, 9, 9)]
    arrays = [smooth(sh, np.float32, 300 + k, scale=10.0) for k, sh in enumerate(shapes)]
    streams, bufs = [], []
    for arr in arrays:
        s, buf = new_stream(arr, oc, start)
        streams.append(s)
        bufs.append(buf)
    src = [torch.from_numpy(a).cuda() for a in arrays]
    done, sizes = Z.compress_batch(streams, src, Z.BATCH_DEVICE_FIELDS)
    assert done == len(arrays)
    outs = [torch.zeros_like(t) for t in src]
    for s in streams:
        s.rewind()
        s.rseek(start)
    ddone, dsizes = Z.decompress_batch(streams, outs, Z.BATCH_DEVICE_FIELDS)
    assert ddone == len(arrays)
    for k, arr in enumerate(arrays):
        n, w, m, out = single_round_trip(arr, oc, start)
        assert sizes[k] == n and dsizes[k] == m, k
        assert np.array_equal(words_of(bufs[k], n), w), k
        assert np.array_equal(outs[k].cpu().numpy().view(np.uint32), out.view(np.uint32)), k
    for s in streams:
        s.close()
