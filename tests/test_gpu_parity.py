"""GPU parity: libzfp_hip.so through its C ABI against the oracle.

Bit-exact streams and bit-exact decoded values for every scalar type, 1-D to
3-D, every mode and rounding, partial blocks, a start offset inside a word,
host- and device-resident data; truncation and overflow errors as the
reference reports them (tests/truncated_stream.rs); the offset index and the
index-less decode of variable-rate streams.
"""
import numpy as np
import pytest

import oracle_py as O
import zfp_hip as Z
from fields import random_walk, smooth, specials
from synth import grf

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

PREFIX_WORD = 0x0123456789abcdef
PREFIX_PART = 0x5a5a5a5a5a5a5a5a
ZROUND = {O.ROUND_NEVER: Z.ROUND_NEVER, O.ROUND_FIRST: Z.ROUND_FIRST, O.ROUND_LAST: Z.ROUND_LAST}
TYPES = [O.INT32, O.INT64, O.FLOAT, O.DOUBLE]
SHAPES = {1: [(37,), (256,)], 2: [(9, 13), (33, 31)], 3: [(5, 6, 7), (17, 9, 12)]}


def zcfg(oc):
    return Z.ZfpConfig(oc.minbits, oc.maxbits, oc.maxprec, oc.minexp,
                       Z.ZfpRounding(ZROUND[oc.rounding], bool(oc.tight)))


def configs(ztype, dims):
    out = [("rate%g" % r, O.fixed_rate(r, ztype, dims)) for r in (1, 4, 8, 16, 27)]
    out.append(("rate8a", O.fixed_rate(8, ztype, dims, align=True)))
    out += [("prec%d" % p, O.fixed_precision(p)) for p in (5, 20, 32, 64)]
    out += [("acc%g" % t, O.fixed_accuracy(t)) for t in (1e-2, 1e-6, 1e-9)]
    out.append(("rev", O.reversible()))
    out += [("exp_small", O.config(1, 11, 64, -1074)), ("exp_mixed", O.config(100, 600, 20, -30)),
            ("exp_revmin", O.config(100, 16658, 64, -1075))]
    return out


def data_for(ztype, shape, kind, seed):
    dt = O.DTYPES[ztype]
    if kind == "smooth":
        return smooth(shape, dt, seed, scale=1e3 if np.dtype(dt).kind == "f" else 1.0)
    if kind == "walk":
        return random_walk(shape, dt, seed)
    return specials(shape, dt, seed)


def write_prefix(s, start_bits):
    for _ in range(start_bits // 64):
        s.write_bits(PREFIX_WORD, 64)
    if start_bits % 64:
        s.write_bits(PREFIX_PART & ((1 << (start_bits % 64)) - 1), start_bits % 64)


def gpu_roundtrip(arr, oc, start_bits=0, device=False):
    """Compress then decompress through the C ABI; returns (nbytes, words, decoded, read)."""
    cap = O.lib().orc_maximum_size(oc, O.TYPE_OF[arr.dtype],
                                   (O.ctypes.c_size_t * 3)(*(list(reversed(arr.shape)) + [0] * (3 - arr.ndim))))
    nwords = (cap + start_bits // 8 + 16) // 8
    if device:
        buf = torch.zeros(nwords, dtype=torch.int64, device="cuda")
        src = torch.from_numpy(np.ascontiguousarray(arr)).cuda()
        out = torch.zeros_like(src)
    else:
        buf = np.zeros(nwords, dtype=np.uint64)
        src = arr
        out = np.zeros_like(arr)
    s = Z.Stream(buf, zcfg(oc))
    write_prefix(s, start_bits)
    n = s.compress(src)
    s.rewind()
    s.rseek(start_bits)
    m = s.decompress(out)
    s.close()
    if device:
        torch.cuda.synchronize()
        words = buf.cpu().numpy().view(np.uint64)
        out = out.cpu().numpy()
    else:
        words = buf
    return n, words, out, m


def check(arr, oc, start_bits=0, device=False, ctx=""):
    on, oerr, ow = O.compress(arr, oc, start_bits=start_bits)
    assert oerr == 0, ctx
    n, words, out, m = gpu_roundtrip(arr, oc, start_bits, device)
    assert n == on, f"{ctx}: bytes {n} != {on}"
    assert np.array_equal(words[: n // 8], ow[: n // 8]), f"{ctx}: stream words differ"
    ref, rn, rerr = O.decompress(ow, arr.shape, arr.dtype, oc, start_bits=start_bits)
    assert rerr == 0 and m == rn, f"{ctx}: read {m} != {rn}"
    u = np.uint32 if arr.dtype.itemsize == 4 else np.uint64
    assert np.array_equal(out.view(u), ref.view(u)), f"{ctx}: decoded values differ"


@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("ztype", TYPES)
def test_modes_match_oracle(dims, ztype):
    for shape in SHAPES[dims]:
        kinds = ("smooth", "specials") if ztype in (O.FLOAT, O.DOUBLE) else ("smooth", "walk", "specials")
        for kind in kinds:
            arr = data_for(ztype, shape, kind, seed=(dims * 31 + ztype * 7 + len(kind)) & 0xffff)
            for name, oc in configs(ztype, dims):
                for start in (0, 13):
                    check(arr, oc, start, device=(start == 0 and kind == "smooth"),
                          ctx=f"{name} {kind} {shape} start={start}")


@pytest.mark.parametrize("ztype", TYPES)
def test_roundings_match_oracle(ztype):
    for dims in (1, 2, 3):
        arr = data_for(ztype, SHAPES[dims][1], "smooth", seed=5 + dims)
        for rnd in (O.ROUND_FIRST, O.ROUND_LAST):
            for tight in (False, True):
                for name, base in configs(ztype, dims):
                    oc = O.config(base.minbits, base.maxbits, base.maxprec, base.minexp, rnd, tight)
                    check(arr, oc, 0, ctx=f"{name} d{dims} rounding={rnd} tight={tight}")


def test_plain_64cube_rate8_bytes():
    """The plain small workload: 64^3 f32, rate 8, Never; byte-compare the stream."""
    arr = grf((64, 64, 64), seed=1)
    oc = O.fixed_rate(8, O.FLOAT, 3)
    check(arr, oc, 0, device=True, ctx="64^3 rate 8")
    n, _, _, _ = gpu_roundtrip(arr, oc, 0, device=True)
    assert n == 4096 * 64  # 4096 blocks x 512 bits


@pytest.mark.parametrize("rate", [4, 8, 16])
def test_large_fixed_rate_matches_oracle(rate):
    """128^3 slice of the large workload (zero blocks, subnormals), all roundings."""
    arr = grf((128, 128, 128), seed=rate, zero_blocks=0.05, subnormals=0.01)
    for rnd, tight in ((O.ROUND_NEVER, False), (O.ROUND_FIRST, False), (O.ROUND_LAST, False),
                       (O.ROUND_FIRST, True)):
        oc = O.fixed_rate(rate, O.FLOAT, 3)
        oc = O.config(oc.minbits, oc.maxbits, oc.maxprec, oc.minexp, rnd, tight)
        check(arr, oc, 0, device=True, ctx=f"rate {rate} rounding {rnd}/{tight}")


def test_header_then_stream_round_trip():
    """zfp_write_header + zfp_compress, then zfp_read_header + zfp_decompress."""
    import ctypes
    arr = grf((20, 24, 28), seed=3).astype(np.float64)
    L = Z.lib()
    buf = np.zeros(200000, dtype=np.uint64)
    s = Z.Stream(buf, Z.ZfpConfig.fixed_accuracy(1e-5))
    f = Z.make_field(arr)
    hb = L.zfp_write_header(s.zs, ctypes.byref(f), Z.ZFP_HEADER_FULL)
    assert hb == 32 + 52 + 12
    n = s.compress(arr)
    # oracle: the same header bits then the stream
    oc = O.fixed_accuracy(1e-5)
    on, oerr, ow = O.compress(arr, oc, start_bits=0)
    s.rewind()
    f2 = Z.zfp_field()
    s2 = Z.Stream(buf)
    assert L.zfp_read_header(s2.zs, ctypes.byref(f2), Z.ZFP_HEADER_FULL) == hb
    assert (f2.type, f2.nx, f2.ny, f2.nz) == (Z.TYPE_DOUBLE, 28, 24, 20)
    out = np.zeros_like(arr)
    f2.data = out.ctypes.data
    m = L.zfp_decompress(s2.zs, ctypes.byref(f2))
    assert m == n
    ref, _, _ = O.decompress(ow, arr.shape, arr.dtype, oc)
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    s.close()
    s2.close()


def test_variable_rate_index_and_walk():
    """Offset index equals the oracle's block offsets; decode works with and without it."""
    arr = grf((24, 36, 40), seed=9, scale=1e3).astype(np.float64)
    oc = O.fixed_accuracy(1e-6)
    on, oerr, ow = O.compress(arr, oc)
    offs = O.block_offsets(ow, arr.shape, arr.dtype, oc)
    buf = np.zeros(len(ow) + 8, dtype=np.uint64)
    s = Z.Stream(buf, zcfg(oc))
    n = s.compress(arr)
    assert n == on and np.array_equal(buf[: n // 8], ow[: n // 8])
    start, lens = s.index()
    assert start == 0 and np.array_equal(lens.astype(np.uint64), np.diff(offs))
    # decode through the index
    s.rewind()
    out = np.zeros_like(arr)
    assert s.decompress(out) == n
    ref, _, _ = O.decompress(ow, arr.shape, arr.dtype, oc)
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    # a fresh stream over the same bytes has no index: the GPU walks it
    s2 = Z.Stream(buf.copy(), zcfg(oc))
    out2 = np.zeros_like(arr)
    assert s2.decompress(out2) == n
    assert np.array_equal(out2.view(np.uint64), ref.view(np.uint64))
    # and an imported index gives the same
    s3 = Z.Stream(buf.copy(), zcfg(oc))
    assert s3.set_index(0, lens)
    out3 = np.zeros_like(arr)
    assert s3.decompress(out3) == n
    assert np.array_equal(out3.view(np.uint64), ref.view(np.uint64))
    for x in (s, s2, s3):
        x.close()


@pytest.mark.parametrize("ztype", TYPES)
def test_variable_rate_sparse_and_dense_workgroups(ztype):
    """A field that is constant but for a noisy patch: the workgroups over the
    constant part stage their words through LDS, those over the patch hold far
    more than the mean and are coded block by block from HBM; both give the
    oracle's bytes and values."""
    dt = O.DTYPES[ztype]
    rng = np.random.default_rng(40 + ztype)
    arr = np.zeros((64, 64, 64), dtype=dt)
    patch = (slice(18, 30), slice(0, 64), slice(0, 40))
    shape = arr[patch].shape
    if np.dtype(dt).kind == "f":
        arr[...] = 1.5
        arr[patch] = rng.standard_normal(shape).astype(dt) * 1e3
        cases = [("precision 28", O.fixed_precision(28)), ("accuracy 1e-3", O.fixed_accuracy(1e-3))]
    else:
        arr[...] = 7
        arr[patch] = rng.integers(-2 ** 20, 2 ** 20, size=shape).astype(dt)
        cases = [("precision 24", O.fixed_precision(24))]
    cases.append(("reversible", O.reversible()))
    for name, oc in cases:
        for device in (False, True):
            check(arr, oc, 0, device=device, ctx=f"{name} device={device}")


@pytest.mark.parametrize("mode", ["rate", "precision", "accuracy", "reversible"])
@pytest.mark.parametrize("device", [False, True])
def test_buffer_too_small_matches_reference_words(mode, device):
    """An undersized buffer: compress reports failure, and the words that fit are the
    reference's (writes past the end are dropped, core.rs:159-170), for fixed-rate and
    variable-rate streams (the latter through the write pass's staged workgroups, whose
    stores stop at the capacity)."""
    arr = grf((16, 16, 24), seed=4)
    if mode == "accuracy":
        arr = arr.astype(np.float64) * 1e3
    elif mode == "reversible":
        arr = np.cumsum(np.round(arr * 8).astype(np.int32), axis=2, dtype=np.int32)
    oc = {"rate": O.fixed_rate(8, O.FLOAT, 3), "precision": O.fixed_precision(20),
          "accuracy": O.fixed_accuracy(1e-4), "reversible": O.reversible()}[mode]
    on, oerr, ow = O.compress(arr, oc)
    assert oerr == 0
    for short in (on // 8 - 3, on // 16):
        if device:
            buf = torch.zeros(short, dtype=torch.int64, device="cuda")
            src = torch.from_numpy(arr).cuda()
        else:
            buf = np.zeros(short, dtype=np.uint64)
            src = arr
        s = Z.Stream(buf, zcfg(oc))
        with pytest.raises(Z.BufferTooSmall):
            s.compress(src)
        got = buf.cpu().numpy().view(np.uint64) if device else buf
        assert np.array_equal(got, ow[:short]), (mode, short)
        s.close()


def test_truncated_stream_is_reported():
    """Decoding that loads a word past the end fails; whole streams do not
    (tests/truncated_stream.rs)."""
    arr = grf((16, 16, 16), seed=6)
    for oc in (O.fixed_rate(8, O.FLOAT, 3), O.fixed_precision(16)):
        on, oerr, ow = O.compress(arr, oc)
        whole = ow[: on // 8].copy()
        out = np.zeros_like(arr)
        s = Z.Stream(whole, zcfg(oc))
        assert s.decompress(out) == on
        s.close()
        cut = ow[: on // 8 - 2].copy()
        s = Z.Stream(cut, zcfg(oc))
        with pytest.raises(Z.Truncated):
            s.decompress(np.zeros_like(arr))
        s.close()
        # the oracle agrees that the cut stream is truncated
        _, _, err = O.decompress(cut, arr.shape, arr.dtype, oc)
        assert err == 3


def test_missing_padding_is_not_truncated():
    """A stream missing only the padding a zero block skips over is whole."""
    arr = np.zeros((8, 8, 8), dtype=np.float32)
    oc = O.fixed_rate(8, O.FLOAT, 3)
    on, oerr, ow = O.compress(arr, oc)
    cut = ow[: on // 8 - 7].copy()  # the last block is one bit plus padding
    _, rn, err = O.decompress(cut, arr.shape, arr.dtype, oc)
    s = Z.Stream(cut, zcfg(oc))
    out = np.ones_like(arr)
    if err == 0:
        assert s.decompress(out) == rn
        assert not out.any()
    else:
        with pytest.raises(Z.Truncated):
            s.decompress(out)
    s.close()
