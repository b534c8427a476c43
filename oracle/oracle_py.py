"""ctypes binding of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the oracle is the parity checker and the timed CPU
baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this module; the shipped library never does.
"""
import ctypes
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

INT32, INT64, FLOAT, DOUBLE = 1, 2, 3, 4
ROUND_NEVER, ROUND_FIRST, ROUND_LAST = 0, 1, 2
DTYPES = {INT32: np.int32, INT64: np.int64, FLOAT: np.float32, DOUBLE: np.float64}
TYPE_OF = {np.dtype(v): k for k, v in DTYPES.items()}

ZFP_MIN_BITS, ZFP_MAX_BITS, ZFP_MAX_PREC, ZFP_MIN_EXP = 1, 16658, 64, -1074


class Stream(ctypes.Structure):
    _fields_ = [("words", ctypes.POINTER(ctypes.c_uint64)), ("nwords", ctypes.c_size_t),
                ("word_pos", ctypes.c_size_t), ("buffer", ctypes.c_uint64),
                ("bits", ctypes.c_uint32), ("overflowed", ctypes.c_int),
                ("overread", ctypes.c_int), ("seek_word", ctypes.c_size_t)]


class Config(ctypes.Structure):
    _fields_ = [("minbits", ctypes.c_uint32), ("maxbits", ctypes.c_uint32),
                ("maxprec", ctypes.c_uint32), ("minexp", ctypes.c_int32),
                ("rounding", ctypes.c_int32), ("tight", ctypes.c_int32)]


class Field(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("nx", ctypes.c_size_t), ("ny", ctypes.c_size_t),
                ("nz", ctypes.c_size_t), ("sx", ctypes.c_ssize_t), ("sy", ctypes.c_ssize_t),
                ("sz", ctypes.c_ssize_t), ("data", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.POINTER
        L.orc_stream_init.argtypes = [P(Stream), P(ctypes.c_uint64), ctypes.c_size_t]
        for name in ("orc_flush", "orc_align"):
            getattr(L, name).argtypes = [P(Stream)]
            getattr(L, name).restype = ctypes.c_uint32
        L.orc_rewind.argtypes = [P(Stream)]
        L.orc_rseek.argtypes = [P(Stream), ctypes.c_uint64]
        L.orc_wseek.argtypes = [P(Stream), ctypes.c_uint64]
        L.orc_wtell.argtypes = [P(Stream)]
        L.orc_wtell.restype = ctypes.c_uint64
        L.orc_rtell.argtypes = [P(Stream)]
        L.orc_rtell.restype = ctypes.c_uint64
        L.orc_write_bits.argtypes = [P(Stream), ctypes.c_uint64, ctypes.c_uint32]
        L.orc_write_bits.restype = ctypes.c_uint64
        L.orc_read_bits.argtypes = [P(Stream), ctypes.c_uint32]
        L.orc_read_bits.restype = ctypes.c_uint64
        L.orc_skip.argtypes = [P(Stream), ctypes.c_uint64]
        L.orc_pad.argtypes = [P(Stream), ctypes.c_uint64]
        L.orc_overread.argtypes = [P(Stream)]
        L.orc_encode_block.argtypes = [P(Stream), P(Config), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p, P(ctypes.c_ssize_t), P(ctypes.c_size_t)]
        L.orc_encode_block.restype = ctypes.c_uint32
        L.orc_decode_block.argtypes = L.orc_encode_block.argtypes
        L.orc_decode_block.restype = ctypes.c_uint32
        for name in ("orc_compress", "orc_decompress"):
            getattr(L, name).argtypes = [P(Stream), P(Config), P(Field), P(ctypes.c_int)]
            getattr(L, name).restype = ctypes.c_size_t
        for name in ("orc_compress_threads", "orc_decompress_threads"):
            getattr(L, name).argtypes = [P(Stream), P(Config), P(Field), ctypes.c_int,
                                         P(ctypes.c_int)]
            getattr(L, name).restype = ctypes.c_size_t
        L.orc_block_offsets.argtypes = [P(Stream), P(Config), P(Field), P(ctypes.c_uint64)]
        L.orc_rate_bits.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_rate_bits.restype = ctypes.c_uint32
        L.orc_block_bits.argtypes = [P(Config), ctypes.c_int, ctypes.c_int]
        L.orc_block_bits.restype = ctypes.c_uint32
        L.orc_maximum_size.argtypes = [P(Config), ctypes.c_int, P(ctypes.c_size_t)]
        L.orc_maximum_size.restype = ctypes.c_size_t
        L.orc_num_blocks.argtypes = [P(Field)]
        L.orc_num_blocks.restype = ctypes.c_size_t
        vp, ci, cu = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32
        for name, args, res in (
                ("orc_exponent_block_f32", [vp, ci], ci), ("orc_exponent_block_f64", [vp, ci], ci),
                ("orc_fwd_cast_f32", [vp, vp, ci, ci], None), ("orc_fwd_cast_f64", [vp, vp, ci, ci], None),
                ("orc_inv_cast_f32", [vp, vp, ci, ci], None), ("orc_inv_cast_f64", [vp, vp, ci, ci], None),
                ("orc_truncate_f64", [ctypes.c_double], ctypes.c_int64),
                ("orc_precision_f", [ci, cu, ctypes.c_int32, ci, ci], cu),
                ("orc_xform_i32", [vp, ci, ci], None), ("orc_xform_i64", [vp, ci, ci], None),
                ("orc_fwd_round_i32", [vp, ci, cu], None), ("orc_fwd_round_i64", [vp, ci, cu], None),
                ("orc_inv_round_u32", [vp, ci, cu, cu], None), ("orc_inv_round_u64", [vp, ci, cu, cu], None),
                ("orc_to_planes", [vp, ci, vp], None)):
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def config(minbits, maxbits, maxprec, minexp, rounding=ROUND_NEVER, tight=False):
    return Config(minbits, maxbits, maxprec, minexp, rounding, int(bool(tight)))


def fixed_rate(rate, ztype, dims, align=False, **kw):
    bits = lib().orc_rate_bits(rate, ztype, dims, int(align))
    return config(bits, bits, ZFP_MAX_PREC, ZFP_MIN_EXP, **kw)


def fixed_precision(prec, **kw):
    return config(ZFP_MIN_BITS, ZFP_MAX_BITS, min(prec, 64) if prec > 0 else 64, ZFP_MIN_EXP, **kw)


def fixed_accuracy(tol, **kw):
    import math
    emin = math.frexp(tol)[1] - 1 if tol > 0 else ZFP_MIN_EXP
    return config(ZFP_MIN_BITS, ZFP_MAX_BITS, ZFP_MAX_PREC, emin, **kw)


def reversible(**kw):
    return config(ZFP_MIN_BITS, ZFP_MAX_BITS, ZFP_MAX_PREC, ZFP_MIN_EXP - 1, **kw)


class OracleStream:
    """A zero-initialised word buffer with an oracle cursor."""

    def __init__(self, nbytes=None, words=None):
        if words is None:
            words = np.zeros((nbytes + 7) // 8, dtype=np.uint64)
        self.words = words
        self.s = Stream()
        lib().orc_stream_init(ctypes.byref(self.s), self._ptr(), len(self.words))

    def _ptr(self):
        return self.words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))

    def flush(self):
        return lib().orc_flush(ctypes.byref(self.s))

    def rewind(self):
        lib().orc_rewind(ctypes.byref(self.s))

    def wtell(self):
        return lib().orc_wtell(ctypes.byref(self.s))

    def rtell(self):
        return lib().orc_rtell(ctypes.byref(self.s))

    def committed(self):
        n = min(self.s.word_pos, len(self.words))
        return self.words[:n].copy()


def _field(arr, strides=None, shape=None):
    """Field over a numpy array; shape is (nx[, ny[, nz]]) in zfp order (x fastest)."""
    ztype = TYPE_OF[arr.dtype]
    if shape is None:
        shape = tuple(reversed(arr.shape))
    nx = shape[0]
    ny = shape[1] if len(shape) > 1 else 0
    nz = shape[2] if len(shape) > 2 else 0
    sx, sy, sz = strides if strides is not None else (0, 0, 0)
    return Field(ztype, nx, ny, nz, sx, sy, sz, arr.ctypes.data)


def compress(arr, cfg, capacity=None, threads=1, start_bits=0):
    """Compress a C-contiguous numpy array (zfp x = last numpy axis).

    Returns (bytes_or_0, err, words) where words is the committed stream."""
    ztype = TYPE_OF[arr.dtype]
    dims = arr.ndim
    if capacity is None:
        dv = [d for d in reversed(arr.shape)] + [0] * (3 - dims)
        capacity = lib().orc_maximum_size(ctypes.byref(cfg), ztype,
                                          (ctypes.c_size_t * 3)(*dv)) + 8 + start_bits // 8
    st = OracleStream(capacity)
    if start_bits:
        lib().orc_write_bits(ctypes.byref(st.s), 0, 0)
        for _ in range(start_bits // 64):
            lib().orc_write_bits(ctypes.byref(st.s), 0x0123456789abcdef, 64)
        if start_bits % 64:
            lib().orc_write_bits(ctypes.byref(st.s), 0x5a5a5a5a5a5a5a5a & ((1 << (start_bits % 64)) - 1),
                                 start_bits % 64)
    f = _field(arr)
    err = ctypes.c_int(0)
    if threads > 1:
        n = lib().orc_compress_threads(ctypes.byref(st.s), ctypes.byref(cfg), ctypes.byref(f),
                                       threads, ctypes.byref(err))
    else:
        n = lib().orc_compress(ctypes.byref(st.s), ctypes.byref(cfg), ctypes.byref(f),
                               ctypes.byref(err))
    return n, err.value, st.words[: (n // 8)].copy() if n else st.words.copy()


def decompress(words, shape, dtype, cfg, threads=1, start_bits=0):
    """Decompress into a new numpy array of the given numpy shape."""
    out = np.zeros(shape, dtype=dtype)
    st = OracleStream(words=np.ascontiguousarray(words, dtype=np.uint64))
    if start_bits:
        lib().orc_rseek(ctypes.byref(st.s), start_bits)
    f = _field(out)
    err = ctypes.c_int(0)
    if threads > 1:
        n = lib().orc_decompress_threads(ctypes.byref(st.s), ctypes.byref(cfg), ctypes.byref(f),
                                         threads, ctypes.byref(err))
    else:
        n = lib().orc_decompress(ctypes.byref(st.s), ctypes.byref(cfg), ctypes.byref(f),
                                 ctypes.byref(err))
    return out, n, err.value


def block_offsets(words, shape, dtype, cfg, start_bits=0):
    """Per-block start bits (plus the end) of a stream, by serial decoding."""
    out = np.zeros(shape, dtype=dtype)
    st = OracleStream(words=np.ascontiguousarray(words, dtype=np.uint64))
    if start_bits:
        lib().orc_rseek(ctypes.byref(st.s), start_bits)
    f = _field(out)
    nb = lib().orc_num_blocks(ctypes.byref(f))
    offs = np.zeros(nb + 1, dtype=np.uint64)
    lib().orc_block_offsets(ctypes.byref(st.s), ctypes.byref(cfg), ctypes.byref(f),
                            offs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return offs


def encode_block(st, cfg, ztype, dims, buf, origin, strides, lengths):
    """Encode one block from numpy buffer `buf` at element offset `origin`."""
    esz = buf.dtype.itemsize
    s = (ctypes.c_ssize_t * 4)(*(list(strides) + [0] * (4 - len(strides))))
    n = (ctypes.c_size_t * 4)(*(list(lengths) + [1] * (4 - len(lengths))))
    return lib().orc_encode_block(ctypes.byref(st.s), ctypes.byref(cfg), ztype, dims,
                                  buf.ctypes.data + origin * esz, s, n)


def decode_block(st, cfg, ztype, dims, buf, origin, strides, lengths):
    esz = buf.dtype.itemsize
    s = (ctypes.c_ssize_t * 4)(*(list(strides) + [0] * (4 - len(strides))))
    n = (ctypes.c_size_t * 4)(*(list(lengths) + [1] * (4 - len(lengths))))
    return lib().orc_decode_block(ctypes.byref(st.s), ctypes.byref(cfg), ztype, dims,
                                  buf.ctypes.data + origin * esz, s, n)


def _blocks(shape):
    """(origin, strides, lengths) of each block of a C-contiguous array of numpy
    shape `shape`, in the reference's order: x fastest (field_plan.rs:41-77),
    partial blocks at the high ends."""
    n = list(reversed(shape))
    nb = [(v + 3) // 4 for v in n]
    strides = [int(np.prod(n[:a], dtype=np.int64)) for a in range(len(n))]
    for idx in np.ndindex(*reversed(nb)):
        c = list(reversed(idx))
        origin = sum(4 * c[a] * strides[a] for a in range(len(n)))
        yield origin, strides, [min(4, n[a] - 4 * c[a]) for a in range(len(n))]


def compress_blocks(arr, cfg, nwords):
    """The reference's serial compress loop (src/compress.rs:22-115) over the
    block coder, for 1-D to 4-D arrays: (words after the flush, per-block start
    bits plus the end)."""
    st = OracleStream(words=np.zeros(nwords, dtype=np.uint64))
    flat = np.ascontiguousarray(arr).reshape(-1)
    offs = [0]
    for origin, strides, lens in _blocks(arr.shape):
        encode_block(st, cfg, TYPE_OF[arr.dtype], arr.ndim, flat, origin, strides, lens)
        offs.append(st.wtell())
    st.flush()
    return st.words.copy(), np.array(offs, dtype=np.uint64)


def decompress_blocks(words, shape, dtype, cfg):
    """The serial decompress loop (src/decompress.rs:27-147) over the block decoder."""
    out = np.zeros(int(np.prod(shape)), dtype=dtype)
    st = OracleStream(words=np.ascontiguousarray(words, dtype=np.uint64).copy())
    for origin, strides, lens in _blocks(shape):
        decode_block(st, cfg, TYPE_OF[np.dtype(dtype)], len(shape), out, origin, strides, lens)
    return out.reshape(shape)
