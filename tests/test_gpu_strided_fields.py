"""Whole fields in strided layouts on the GPU, against the oracle.

The reference codes a field through its strides on every path: as-is,
reversed, interleaved and permuted layouts (tests/c/endtoend.rs:34-178), and
negative strides from a start bit inside a word (tests/pipeline.rs:7-209).
Compressing a strided field gives the stream of its contiguous copy;
decompressing into one writes only the field's elements and leaves the gaps
between them as they were.  Fields whose strides may alias are decoded as the
reference's serial decode does (src/decompress.rs:157-163, 380-400): each
element keeps the value written last.
"""
import ctypes

import numpy as np
import pytest

import oracle_py as O
import zfp_hip as Z
from fields import smooth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

TYPES = [O.INT32, O.INT64, O.FLOAT, O.DOUBLE]
ZROUND = {O.ROUND_NEVER: Z.ROUND_NEVER, O.ROUND_FIRST: Z.ROUND_FIRST, O.ROUND_LAST: Z.ROUND_LAST}
PREFIX = 0x5a5a5a5a5a5a5a5a  # bits written before the field (as oracle_py.compress)


def zcfg(oc):
    return Z.ZfpConfig(oc.minbits, oc.maxbits, oc.maxprec, oc.minexp,
                       Z.ZfpRounding(ZROUND[oc.rounding], bool(oc.tight)))


def layouts(dims):
    """(name, strides) of a field of extents `dims` (x first)."""
    n = list(dims)
    nat = [1]
    for v in n[:-1]:
        nat.append(nat[-1] * v)
    out = [("as-is", nat), ("reversed", [-s for s in nat]), ("interleaved", [2 * s for s in nat])]
    if len(n) >= 2:
        # y fastest, then x, then z (a transposed store)
        perm = [n[1], 1] + ([n[0] * n[1]] if len(n) == 3 else [])
        out.append(("permuted", perm))
        out.append(("padded", [1, n[0] + 3] + ([(n[0] + 3) * n[1] + 5] if len(n) == 3 else [])))
        out.append(("mixed", [-1, n[0] + 2] + ([-(n[0] + 2) * n[1]] if len(n) == 3 else [])))
    return out


def span(dims, strides):
    lo = sum(s * (v - 1) for v, s in zip(dims, strides) if s < 0)
    hi = sum(s * (v - 1) for v, s in zip(dims, strides) if s > 0)
    return lo, hi


def addresses(dims, strides, lo):
    """Buffer index of each element of the field, in numpy order (x last)."""
    idx = np.zeros(tuple(reversed(dims)), dtype=np.int64) - lo
    for a, (v, s) in enumerate(zip(dims, strides)):
        shape = [1] * len(dims)
        shape[len(dims) - 1 - a] = v
        idx = idx + (np.arange(v, dtype=np.int64) * s).reshape(shape)
    return idx


def zfield(ztype, dims, strides, ptr):
    f = Z.zfp_field()
    f.type = ztype
    n = list(dims) + [0] * (4 - len(dims))
    s = list(strides) + [0] * (4 - len(dims))
    f.nx, f.ny, f.nz, f.nw = n
    f.sx, f.sy, f.sz, f.sw = s
    f.data = ptr
    return f


def write_prefix(st, start):
    if start:
        st.write_bits(PREFIX & ((1 << start) - 1), start)


def configs(ztype, dims):
    return [("rate8", O.fixed_rate(8, ztype, dims)), ("rate13", O.fixed_rate(13, ztype, dims)),
            ("prec16", O.fixed_precision(16)), ("acc1e-3", O.fixed_accuracy(1e-3)),
            ("rev", O.reversible())]


SHAPES = {1: (13,), 2: (9, 7), 3: (6, 5, 7)}


def source(ztype, dims, seed):
    dt = O.DTYPES[ztype]
    return smooth(tuple(reversed(dims)), dt, seed, scale=1e3 if np.dtype(dt).kind == "f" else 1.0)


def gpu_code_strided(arr, ztype, dims, strides, oc, start, device):
    """Compress `arr` from a strided copy, then decompress into a second strided
    buffer pre-filled with a pattern; returns (bytes, words, read, buffer, lo)."""
    dt = arr.dtype
    lo, hi = span(dims, strides)
    idx = addresses(dims, strides, lo)
    pattern = (np.arange(hi - lo + 1) % 251 + 7).astype(dt)
    src = pattern.copy()
    src[idx.reshape(-1)] = arr.reshape(-1)
    dst = pattern.copy()
    cap = O.lib().orc_maximum_size(oc, ztype, (O.ctypes.c_size_t * 3)(*(list(dims) + [0] * (3 - len(dims)))))
    nwords = (cap + 32) // 8
    esz = dt.itemsize
    if device:
        buf = torch.zeros(nwords, dtype=torch.int64, device="cuda")
        dsrc = torch.from_numpy(src).cuda()
        ddst = torch.from_numpy(dst).cuda()
        psrc, pdst = dsrc.data_ptr() - lo * esz, ddst.data_ptr() - lo * esz
    else:
        buf = np.zeros(nwords, dtype=np.uint64)
        psrc, pdst = src.ctypes.data - lo * esz, dst.ctypes.data - lo * esz
    st = Z.Stream(buf, zcfg(oc))
    write_prefix(st, start)
    fs = zfield(ztype, dims, strides, psrc)
    n = Z.lib().zfp_compress(st.zs, ctypes.byref(fs))
    st.rewind()
    st.rseek(start)
    fd = zfield(ztype, dims, strides, pdst)
    m = Z.lib().zfp_decompress(st.zs, ctypes.byref(fd))
    st.close()
    if device:
        torch.cuda.synchronize()
        words = buf.cpu().numpy().view(np.uint64)
        dst = ddst.cpu().numpy()
        assert np.array_equal(dsrc.cpu().numpy().view(np.uint8), src.view(np.uint8))
    else:
        words = buf
    return n, words, m, dst, pattern, idx


@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("ztype", TYPES)
def test_strided_layouts_match_oracle(dims, ztype):
    shape = SHAPES[dims][:dims]
    arr = source(ztype, shape, seed=dims * 13 + ztype)
    u = np.uint32 if arr.dtype.itemsize == 4 else np.uint64
    for lname, strides in layouts(shape):
        for cname, oc in configs(ztype, dims):
            for start, device in ((0, True), (13, False), (13, True)):
                ctx = f"{lname} {strides} {cname} start={start} device={device}"
                on, oerr, ow = O.compress(arr, oc, start_bits=start)
                assert oerr == 0
                n, words, m, dst, pattern, idx = gpu_code_strided(arr, ztype, shape, strides, oc,
                                                                  start, device)
                assert n == on, ctx
                assert np.array_equal(words[: n // 8], ow[: n // 8]), ctx
                ref, rn, rerr = O.decompress(ow, arr.shape, arr.dtype, oc, start_bits=start)
                assert rerr == 0 and m == rn, ctx
                want = pattern.copy()
                want[idx.reshape(-1)] = ref.reshape(-1)
                assert np.array_equal(dst.view(u), want.view(u)), ctx


# ------------------------------------------------------------ aliasing strides

def oracle_decode_into(words, ztype, dims, strides, buf, origin, cfg, start=0):
    dt = O.DTYPES[ztype]
    st = O.OracleStream(words=np.ascontiguousarray(words, dtype=np.uint64).copy())
    if start:
        O.lib().orc_rseek(ctypes.byref(st.s), start)
    n = list(dims) + [0] * (3 - len(dims))
    s = list(strides) + [0] * (3 - len(strides))
    f = O.Field(ztype, n[0], n[1], n[2], s[0], s[1], s[2],
                buf.ctypes.data + origin * np.dtype(dt).itemsize)
    err = ctypes.c_int(0)
    got = O.lib().orc_decompress(ctypes.byref(st.s), ctypes.byref(cfg), ctypes.byref(f),
                                 ctypes.byref(err))
    assert err.value == 0
    return got


ALIAS = [
    (O.DOUBLE, (8, 8), (1, 1)),
    (O.FLOAT, (64, 64), (1, 3)),
    (O.INT32, (37, 29), (2, -3)),
    (O.FLOAT, (24, 20, 16), (1, 2, 3)),
    (O.DOUBLE, (19, 17, 21), (-1, 5, -7)),
    (O.INT64, (100,), (1,)),  # not aliasing: the plain path, for contrast
]


@pytest.mark.parametrize("ztype,dims,strides", ALIAS)
def test_aliasing_strides_decode_as_the_serial_decode(ztype, dims, strides):
    """src/decompress.rs:380-400: 8 x 8 f64 at rate 16 into strides [1, 1]
    (15 elements) equals the serial decode, and so do larger fields with many
    blocks writing each element, for fixed- and variable-rate streams, host and
    device, from a start bit inside a word."""
    shape = tuple(reversed(dims))
    dt = O.DTYPES[ztype]
    arr = source(ztype, dims, seed=len(dims) + ztype)
    u = np.uint32 if np.dtype(dt).itemsize == 4 else np.uint64
    f0 = zfield(ztype, dims, strides, 4096)
    assert bool(Z.lib().zfp_hip_field_may_alias(ctypes.byref(f0))) == (strides != (1,))
    lo, hi = span(dims, strides)
    esz = np.dtype(dt).itemsize
    for oc in (O.fixed_rate(16, ztype, len(dims)), O.fixed_precision(14), O.reversible()):
        for start, device in ((0, False), (0, True), (13, True)):
            on, oerr, ow = O.compress(arr, oc, start_bits=start)
            assert oerr == 0
            pattern = (np.arange(hi - lo + 1) % 97 + 3).astype(dt)
            want = pattern.copy()
            rn = oracle_decode_into(ow, ztype, dims, strides, want, -lo, oc, start)
            nw = len(ow) + 4
            words = np.zeros(nw, dtype=np.uint64)
            words[: len(ow)] = ow
            if device:
                buf = torch.from_numpy(words.view(np.int64).copy()).cuda()
                dd = torch.from_numpy(pattern.copy()).cuda()
                ptr = dd.data_ptr() - lo * esz
            else:
                buf = words
                dd = pattern.copy()
                ptr = dd.ctypes.data - lo * esz
            st = Z.Stream(buf, zcfg(oc))
            st.rseek(start)
            fd = zfield(ztype, dims, strides, ptr)
            m = Z.lib().zfp_decompress(st.zs, ctypes.byref(fd))
            st.close()
            got = dd.cpu().numpy() if device else dd
            ctx = f"{oc.minbits},{oc.maxbits},{oc.maxprec},{oc.minexp} start={start} device={device}"
            assert m == rn, ctx
            assert np.array_equal(got.view(u), want.view(u)), ctx


def test_batched_decode_into_aliasing_strides_takes_the_ordered_path():
    """A batch hands a decode into aliasing strides to the single-array call:
    the same values as zfp_decompress, beside ordinary arrays in the batch."""
    ztype, dims, strides = O.FLOAT, (40, 36), (1, 2)
    oc = O.fixed_rate(12, ztype, 2)
    arr = source(ztype, dims, seed=3)
    other = source(ztype, (33, 31), seed=4)
    lo, hi = span(dims, strides)
    outs, streams, fields, keep = [], [], [], []
    for a, st_ in ((arr, strides), (other, None)):
        n, err, w = O.compress(a, oc)
        buf = torch.from_numpy(np.concatenate([w, np.zeros(4, np.uint64)]).view(np.int64).copy()).cuda()
        s = Z.Stream(buf, zcfg(oc))
        if st_ is None:
            out = torch.zeros(a.shape, dtype=torch.float32, device="cuda")
            f = Z.make_field(out)
        else:
            out = torch.full((hi - lo + 1,), 5.0, dtype=torch.float32, device="cuda")
            f = zfield(ztype, dims, strides, out.data_ptr() - lo * 4)
        keep.append((buf, s, f))
        outs.append(out)
        streams.append(s.zs)
        fields.append(ctypes.pointer(f))
    sizes = (ctypes.c_size_t * 2)()
    SP = ctypes.POINTER(Z.zfp_stream)
    done = Z.lib().zfp_hip_decompress_batch((SP * 2)(*streams),
                                            (ctypes.POINTER(Z.zfp_field) * 2)(*fields), 2, sizes, 0)
    assert done == 2
    want = np.full(hi - lo + 1, 5.0, dtype=np.float32)
    n, err, w = O.compress(arr, oc)
    oracle_decode_into(w, ztype, dims, strides, want, -lo, oc)
    assert np.array_equal(outs[0].cpu().numpy().view(np.uint32), want.view(np.uint32))
    n2, err2, w2 = O.compress(other, oc)
    ref, _, _ = O.decompress(w2, other.shape, other.dtype, oc)
    assert np.array_equal(outs[1].cpu().numpy().view(np.uint32), ref.view(np.uint32))
    for _, s, _ in keep:
        s.close()
