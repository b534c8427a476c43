"""GPU: the block-level entry points (zfp-rs-ffi/src/block.rs) against the
oracle.  Blocks coded one call after another at the stream's cursor give the
stream of a field whose blocks they are, side by side along x (the reference's
compress loop calls the same block coder, src/compress.rs:22-115), each call
returning its block's bits; decoding them back gives the oracle's values and
leaves the read cursor at the end of each block.  Strided and negative-stride
blocks code as their contiguous copies; partial blocks as a field of their
extent; host and device memory alike."""

import ctypes

import numpy as np
import pytest

import oracle_py as O
import zfp_hip as Z
from fields import random_walk, smooth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

NAMES = {O.INT32: "int32", O.INT64: "int64", O.FLOAT: "float", O.DOUBLE: "double"}
K = 6  # blocks per case


def zcfg(oc):
    return Z.ZfpConfig(oc.minbits, oc.maxbits, oc.maxprec, oc.minexp, Z.ZfpRounding.never())


def configs(ztype, dims):
    out = [("rate", O.fixed_rate(12, ztype, dims)), ("precision", O.fixed_precision(18))]
    if ztype in (O.FLOAT, O.DOUBLE):
        out.append(("accuracy", O.fixed_accuracy(1e-3)))
    out.append(("reversible", O.reversible()))
    return out


def fn(name, nargs_extra):
    f = getattr(Z.lib(), name)
    f.restype = ctypes.c_size_t
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + nargs_extra
    return f


def field_of(ztype, dims, seed):
    shape = (4,) * (dims - 1) + (4 * K,)
    if ztype in (O.FLOAT, O.DOUBLE):
        return smooth(shape, O.DTYPES[ztype], seed, scale=1e3)
    return random_walk(shape, O.DTYPES[ztype], seed)


def ptr_of(a):
    return a.data_ptr() if isinstance(a, torch.Tensor) else a.ctypes.data


def flushed_words(buf, s, nbytes):
    Z.lib().zfp_stream_flush(s.zs)
    w = buf.cpu().numpy() if isinstance(buf, torch.Tensor) else buf
    return w.view(np.uint64)[: nbytes // 8]


@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("ztype", [O.INT32, O.INT64, O.FLOAT, O.DOUBLE])
def test_blocks_in_order_match_the_field_stream(ztype, dims):
    name = NAMES[ztype]
    enc = fn(f"zfp_encode_block_{name}_{dims}", [])
    dec = fn(f"zfp_decode_block_{name}_{dims}", [])
    arr = field_of(ztype, dims, seed=11 * dims + ztype)
    u = np.uint32 if arr.dtype.itemsize == 4 else np.uint64
    for cname, oc in configs(ztype, dims):
        on, oerr, ow = O.compress(arr, oc)
        assert oerr == 0
        offs = O.block_offsets(ow, arr.shape, arr.dtype, oc)
        ref, _, _ = O.decompress(ow, arr.shape, arr.dtype, oc)
        blocks = [np.ascontiguousarray(arr[..., 4 * k:4 * k + 4]) for k in range(K)]
        for device in (False, True):
            ctx = f"{name} d{dims} {cname} device={device}"
            nw = on // 8 + 4
            buf = torch.zeros(nw, dtype=torch.int64, device="cuda") if device else np.zeros(nw, np.uint64)
            s = Z.Stream(buf, zcfg(oc))
            src = [torch.from_numpy(b).cuda() for b in blocks] if device else blocks
            for k in range(K):
                assert enc(s.zs, ptr_of(src[k])) == offs[k + 1] - offs[k], (ctx, k)
            assert Z.lib().stream_wtell(s.bs) == offs[K], ctx
            words = flushed_words(buf, s, on)
            assert np.array_equal(words, ow[: on // 8]), ctx
            # decode the blocks back one by one
            Z.lib().zfp_stream_rewind(s.zs)
            for k in range(K):
                out = torch.zeros_like(src[k]) if device else np.zeros_like(blocks[k])
                assert dec(s.zs, ptr_of(out)) == offs[k + 1] - offs[k], (ctx, k)
                assert Z.lib().stream_rtell(s.bs) == offs[k + 1], (ctx, k)
                got = out.cpu().numpy() if device else out
                want = np.ascontiguousarray(ref[..., 4 * k:4 * k + 4])
                assert np.array_equal(got.view(u), want.view(u)), (ctx, k)
            s.close()


@pytest.mark.parametrize("ztype", [O.FLOAT, O.INT64])
def test_strided_and_negative_stride_blocks(ztype):
    """A block read in place from a larger array (its strides) and from a
    reversed view (negative strides) codes as its contiguous copy."""
    name = NAMES[ztype]
    esz = np.dtype(O.DTYPES[ztype]).itemsize
    big = field_of(ztype, 3, seed=5)  # (4, 4, 4K)
    oc = O.fixed_precision(20)
    args3 = [ctypes.c_ssize_t] * 3
    enc = fn(f"zfp_encode_block_strided_{name}_3", args3)
    dec = fn(f"zfp_decode_block_strided_{name}_3", args3)
    for sel in (lambda a: a[..., 4:8], lambda a: a[:, ::-1, ::-1][..., 8:12]):
        view = sel(big)
        want_block = np.ascontiguousarray(view)
        on, _, ow = O.compress(want_block, oc)
        ref, _, _ = O.decompress(ow, want_block.shape, want_block.dtype, oc)
        st = [view.strides[2] // esz, view.strides[1] // esz, view.strides[0] // esz]
        buf = np.zeros(on // 8 + 4, np.uint64)
        s = Z.Stream(buf, zcfg(oc))
        bits = enc(s.zs, view.ctypes.data, *st)
        assert bits > 0 and Z.lib().stream_wtell(s.bs) == bits
        assert np.array_equal(flushed_words(buf, s, on), ow[: on // 8])
        # decode into a fresh array of the same layout: only the block's cells change
        dst_big = np.full_like(big, 7)
        dst = sel(dst_big)
        Z.lib().zfp_stream_rewind(s.zs)
        assert dec(s.zs, dst.ctypes.data, *st) == bits
        u = np.uint32 if esz == 4 else np.uint64
        assert np.array_equal(np.ascontiguousarray(dst).view(u), ref.view(u))
        mask = np.ones(big.shape, bool)
        sel(mask)[...] = False
        assert (dst_big[mask] == 7).all()
        s.close()


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_partial_blocks_code_as_their_extent(dims):
    """A partial block of extent (nx, ny, nz) codes as a field of that shape
    (one padded block); decoding writes only its cells."""
    ztype = O.DOUBLE
    esz = 8
    big = smooth((6, 6, 6)[:dims], np.float64, 3, scale=10.0)
    ext = {1: [(3,), (1,), (4,)], 2: [(3, 2), (1, 4), (2, 1)], 3: [(3, 2, 4), (1, 1, 1), (4, 3, 2)]}[dims]
    nargs = [ctypes.c_size_t] * dims + [ctypes.c_ssize_t] * dims
    enc = fn(f"zfp_encode_partial_block_strided_double_{dims}", nargs)
    dec = fn(f"zfp_decode_partial_block_strided_double_{dims}", nargs)
    strides = [big.strides[dims - 1 - a] // esz for a in range(dims)]
    for oc in (O.fixed_rate(16, ztype, dims), O.fixed_accuracy(1e-4), O.reversible()):
        for n in ext:  # n[a] = extent along axis a (x first)
            sl = tuple(slice(1, 1 + n[dims - 1 - i]) for i in range(dims))
            part = np.ascontiguousarray(big[sl])
            on, oerr, ow = O.compress(part, oc)
            assert oerr == 0
            ref, _, _ = O.decompress(ow, part.shape, part.dtype, oc)
            buf = np.zeros(on // 8 + 4, np.uint64)
            s = Z.Stream(buf, zcfg(oc))
            origin = big[tuple(slice(1, None) for _ in range(dims))]
            bits = enc(s.zs, origin.ctypes.data, *n, *strides)
            assert bits > 0
            assert np.array_equal(flushed_words(buf, s, on), ow[: on // 8]), (n, oc.maxbits)
            dst_big = np.full_like(big, -5.0)
            dst = dst_big[tuple(slice(1, None) for _ in range(dims))]
            Z.lib().zfp_stream_rewind(s.zs)
            assert dec(s.zs, dst.ctypes.data, *n, *strides) == bits
            assert np.array_equal(dst_big[sl].view(np.uint64), ref.view(np.uint64)), n
            mask = np.ones(big.shape, bool)
            mask[sl] = False
            assert (dst_big[mask] == -5.0).all()
            s.close()


def test_refused_blocks_return_zero():
    """Null blocks and partial lengths outside 1..4 code nothing."""
    buf = np.zeros(64, np.uint64)
    s = Z.Stream(buf, Z.ZfpConfig.fixed_precision(16))
    blk = np.ones(256, np.float32)
    assert fn("zfp_encode_block_float_3", [])(s.zs, None) == 0
    p2 = fn("zfp_encode_partial_block_strided_float_2", [ctypes.c_size_t] * 2 + [ctypes.c_ssize_t] * 2)
    assert p2(s.zs, blk.ctypes.data, 0, 4, 1, 4) == 0
    assert p2(s.zs, blk.ctypes.data, 4, 5, 1, 4) == 0
    assert Z.lib().stream_wtell(s.bs) == 0
    s.close()
