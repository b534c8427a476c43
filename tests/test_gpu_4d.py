"""GPU: 4-D fields and blocks through the C ABI against the oracle (the
reference's serial loop over its 4-D block coder, src/compress.rs:22-115,
src/codec/encode/dim4.rs, src/codec/decode/dim4.rs): stream words and decoded
values bit for bit, host and device memory, with and without the offset index;
the 4-D block-level entry points (zfp-rs-ffi/src/block.rs) likewise."""
import ctypes

import numpy as np
import pytest

import oracle_py as O
import zfp_hip as Z
from fields import random_walk, smooth, specials

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

TYPES = [O.INT32, O.INT64, O.FLOAT, O.DOUBLE]
NAMES = {O.INT32: "int32", O.INT64: "int64", O.FLOAT: "float", O.DOUBLE: "double"}


def zcfg(oc):
    kind = {O.ROUND_NEVER: Z.ROUND_NEVER, O.ROUND_FIRST: Z.ROUND_FIRST, O.ROUND_LAST: Z.ROUND_LAST}
    return Z.ZfpConfig(oc.minbits, oc.maxbits, oc.maxprec, oc.minexp,
                       Z.ZfpRounding(kind[oc.rounding], bool(oc.tight)))


def configs(ztype):
    out = [("rate4", O.fixed_rate(4, ztype, 4)), ("prec24", O.fixed_precision(24))]
    if ztype in (O.FLOAT, O.DOUBLE):
        out.append(("acc1e-5", O.fixed_accuracy(1e-5)))
    b = O.fixed_precision(14)
    out.append(("prec14_last", O.config(b.minbits, b.maxbits, b.maxprec, b.minexp, O.ROUND_LAST, True)))
    out.append(("rev", O.reversible()))
    return out


def field(ztype, shape, seed):
    dt = O.DTYPES[ztype]
    if np.dtype(dt).kind == "f":
        return smooth(shape, dt, seed, scale=1e3)
    return random_walk(shape, dt, seed)


def check(arr, oc, ctx):
    u = np.uint32 if arr.dtype.itemsize == 4 else np.uint64
    cap_words = arr.size * 2 + 4096
    ow, offs = O.compress_blocks(arr, oc, cap_words)
    nbytes = (int(offs[-1]) + 63) // 64 * 8
    ref = O.decompress_blocks(ow, arr.shape, arr.dtype, oc)
    for device in (False, True):
        c = f"{ctx} device={device}"
        buf = torch.zeros(cap_words, dtype=torch.int64, device="cuda") if device else np.zeros(cap_words, np.uint64)
        src = torch.from_numpy(arr).cuda() if device else arr
        s = Z.Stream(buf, zcfg(oc))
        assert s.compress(src) == nbytes, c
        words = buf.cpu().numpy().view(np.uint64) if device else buf
        assert np.array_equal(words[: nbytes // 8], ow[: nbytes // 8]), c
        assert not words[nbytes // 8:].any(), c
        # decode with the index the compress kept, then from a fresh stream
        # over the same words (the length walk)
        for fresh in (False, True):
            if fresh:
                s.close()
                s = Z.Stream(buf, zcfg(oc))
            s.rewind()
            out = torch.zeros_like(src) if device else np.zeros_like(arr)
            assert s.decompress(out) == nbytes, (c, fresh)
            got = out.cpu().numpy() if device else out
            assert np.array_equal(got.view(u), ref.view(u)), (c, fresh)
        s.close()


@pytest.mark.parametrize("ztype", TYPES)
def test_4d_fields_match_oracle(ztype):
    shape = (5, 3, 6, 9)  # w, z, y, x: partial blocks along every axis
    arr = field(ztype, shape, seed=31 + ztype)
    for name, oc in configs(ztype):
        check(arr, oc, f"{NAMES[ztype]} {name}")


def test_4d_special_values():
    for ztype in (O.FLOAT, O.DOUBLE):
        arr = specials((4, 4, 5, 8), O.DTYPES[ztype], seed=ztype)
        check(arr, O.reversible(), f"{NAMES[ztype]} specials rev")
        check(arr, O.fixed_precision(20), f"{NAMES[ztype]} specials prec20")


def test_4d_larger_field_fixed_and_variable():
    """32^4 f32 (4096 blocks, more than one workgroup) at rate 8 and precision 16."""
    arr = smooth((32, 32, 32, 32), np.float32, 9, scale=10.0)
    check(arr, O.fixed_rate(8, O.FLOAT, 4), "32^4 rate8")
    check(arr, O.fixed_precision(16), "32^4 prec16")


def fn(name, extra):
    f = getattr(Z.lib(), name)
    f.restype = ctypes.c_size_t
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + extra
    return f


@pytest.mark.parametrize("ztype", TYPES)
def test_4d_block_entry_points(ztype):
    """zfp_encode_block_<type>_4 and the strided and partial variants give the
    oracle's block bits; decoding restores the oracle's values."""
    name = NAMES[ztype]
    esz = np.dtype(O.DTYPES[ztype]).itemsize
    u = np.uint32 if esz == 4 else np.uint64
    big = field(ztype, (5, 5, 5, 5), seed=ztype)
    strides4 = [big.strides[3 - a] // esz for a in range(4)]
    part_args = [ctypes.c_size_t] * 4 + [ctypes.c_ssize_t] * 4
    for oc in (O.fixed_precision(18), O.reversible(), O.fixed_rate(6, ztype, 4)):
        # whole contiguous block
        blk = np.ascontiguousarray(big[:4, :4, :4, :4])
        st = O.OracleStream(8192)
        obits = O.encode_block(st, oc, ztype, 4, blk.reshape(-1), 0, [1, 4, 16, 64], [4] * 4)
        st.flush()
        ow = st.words
        buf = np.zeros(1024, np.uint64)
        s = Z.Stream(buf, zcfg(oc))
        assert fn(f"zfp_encode_block_{name}_4", [])(s.zs, blk.ctypes.data) == obits
        Z.lib().zfp_stream_flush(s.zs)
        assert np.array_equal(buf[: (obits + 63) // 64], ow[: (obits + 63) // 64])
        Z.lib().zfp_stream_rewind(s.zs)
        out = np.zeros_like(blk)
        got = fn(f"zfp_decode_block_{name}_4", [])(s.zs, out.ctypes.data)
        assert got == obits, (got, obits, Z.last_error())
        st.rewind()
        ref = np.zeros_like(blk)
        O.decode_block(st, oc, ztype, 4, ref.reshape(-1), 0, [1, 4, 16, 64], [4] * 4)
        assert np.array_equal(out.view(u), ref.view(u))
        s.close()
        # a partial block read in place through the big array's strides
        n = (3, 2, 4, 1)
        origin = big[1:, 1:, 1:, 1:]
        st = O.OracleStream(8192)
        obits = O.encode_block(st, oc, ztype, 4, big.reshape(-1), 1 + 5 + 25 + 125, strides4, list(n))
        st.flush()
        buf = np.zeros(1024, np.uint64)
        s = Z.Stream(buf, zcfg(oc))
        assert fn(f"zfp_encode_partial_block_strided_{name}_4", part_args)(
            s.zs, origin.ctypes.data, *n, *strides4) == obits
        Z.lib().zfp_stream_flush(s.zs)
        assert np.array_equal(buf[: (obits + 63) // 64], st.words[: (obits + 63) // 64])
        dst_big = np.zeros_like(big)
        Z.lib().zfp_stream_rewind(s.zs)
        assert fn(f"zfp_decode_partial_block_strided_{name}_4", part_args)(
            s.zs, dst_big[1:, 1:, 1:, 1:].ctypes.data, *n, *strides4) == obits
        st.rewind()
        ref_big = np.zeros_like(big)
        O.decode_block(st, oc, ztype, 4, ref_big.reshape(-1), 1 + 5 + 25 + 125, strides4, list(n))
        assert np.array_equal(dst_big.view(u), ref_big.view(u))
        s.close()


def test_batch_codes_4d_arrays_like_single_calls():
    """A batch holding 4-D arrays codes them (with the single-array call) and
    gives every stream the bytes and size of zfp_compress."""
    a3 = torch.from_numpy(field(O.FLOAT, (9, 10, 11), 3)).cuda()
    a4 = torch.from_numpy(field(O.FLOAT, (5, 3, 6, 9), 4)).cuda()
    cfg = Z.ZfpConfig.fixed_precision(18)
    arrays = [a3, a4, a3, a4]
    bufs = [torch.zeros(65536, dtype=torch.int64, device="cuda") for _ in arrays]
    streams = [Z.Stream(b, cfg) for b in bufs]
    _, sizes = Z._batch(Z.lib().zfp_hip_compress_batch, streams, arrays, Z.BATCH_DEVICE_FIELDS)
    for a, b, n in zip(arrays, bufs, sizes):
        ref = torch.zeros(65536, dtype=torch.int64, device="cuda")
        s = Z.Stream(ref, cfg)
        assert n > 0 and s.compress(a) == n
        s.close()
        assert torch.equal(b, ref)
    for s in streams:
        s.close()


def test_batches_with_3d_and_4d_arrays_in_one_stream():
    """Batched compress then batched decompress of streams that each carry two
    arrays in sequence, [3-D, 4-D] in one and [4-D, 3-D] in the other (a
    stream's later arrays follow its first, in input order), against the
    same sequence of single calls: words, sizes and decoded values."""
    a3 = torch.from_numpy(field(O.FLOAT, (9, 10, 11), 7)).cuda()
    a4 = torch.from_numpy(field(O.FLOAT, (5, 3, 6, 9), 8)).cuda()
    cfg = Z.ZfpConfig.fixed_precision(18)
    n = 65536
    bufs = [torch.zeros(n, dtype=torch.int64, device="cuda") for _ in range(2)]
    streams = [Z.Stream(b, cfg) for b in bufs]
    # input order: s0 <- a3, s1 <- a4, s0 <- a4, s1 <- a3
    order = [(0, a3), (1, a4), (0, a4), (1, a3)]
    _, sizes = Z._batch(Z.lib().zfp_hip_compress_batch, [streams[i] for i, _ in order],
                        [a for _, a in order], Z.BATCH_DEVICE_FIELDS)
    refs = [torch.zeros(n, dtype=torch.int64, device="cuda") for _ in range(2)]
    rstreams = [Z.Stream(b, cfg) for b in refs]
    want = [rstreams[i].compress(a) for i, a in order]
    assert sizes == want
    for b, r in zip(bufs, refs):
        assert torch.equal(b, r)
    # decompress both streams the same way, batched and single
    for s in streams + rstreams:
        s.rewind()
    outs = [torch.zeros_like(a) for _, a in order]
    _, dsizes = Z._batch(Z.lib().zfp_hip_decompress_batch, [streams[i] for i, _ in order], outs,
                         Z.BATCH_DEVICE_FIELDS)
    routs = [torch.zeros_like(a) for _, a in order]
    dwant = [rstreams[i].decompress(o) for (i, _), o in zip(order, routs)]
    assert dsizes == dwant
    for o, r in zip(outs, routs):
        assert torch.equal(o.view(torch.int32), r.view(torch.int32))
    for s in streams + rstreams:
        s.close()


def test_refused_plan_keeps_the_stream_index_and_overread():
    """A plan or submission that the batched launches cannot take (here: one
    array is 4-D) is refused with every stream as it was: cursor, offset index
    and overread state (nothing is changed before the batch is accepted)."""
    cfg = Z.ZfpConfig.fixed_precision(16)
    a3 = torch.from_numpy(field(O.FLOAT, (12, 12, 12), 9)).cuda()
    a4 = torch.from_numpy(field(O.FLOAT, (4, 4, 4, 4), 10)).cuda()
    buf = torch.zeros(8192, dtype=torch.int64, device="cuda")
    s = Z.Stream(buf, cfg)
    s.compress(a3)                       # leaves an offset index of 27 blocks
    L = Z.lib()
    before = L.zfp_stream_index_export(s.bs, None, 0, None)
    assert before == 27
    s.rewind()
    other = Z.Stream(torch.zeros(8192, dtype=torch.int64, device="cuda"), cfg)
    f3, f4 = Z.make_field(a3), Z.make_field(a4)
    SP, FP = ctypes.POINTER(Z.zfp_stream), ctypes.POINTER(Z.zfp_field)
    zs = (SP * 2)(s.zs, other.zs)
    fs = (FP * 2)(ctypes.pointer(f3), ctypes.pointer(f4))
    for encode in (1, 0):
        assert not L.zfp_hip_batch_plan(encode, zs, fs, 2, Z.BATCH_DEVICE_FIELDS)
        assert not L.zfp_hip_batch_submit(encode, zs, fs, 2, Z.BATCH_DEVICE_FIELDS, None)
        assert s.wtell() == 0 and s.rtell() == 0
        assert L.zfp_stream_index_export(s.bs, None, 0, None) == 27
    # the kept index still decodes the stream
    out = torch.zeros_like(a3)
    s.decompress(out)
    ref = torch.zeros_like(a3)
    s2 = Z.Stream(buf, cfg)
    s2.decompress(ref)                   # a fresh stream: no index, walks
    assert torch.equal(out.view(torch.int32), ref.view(torch.int32))
    for x in (s, s2, other):
        x.close()
