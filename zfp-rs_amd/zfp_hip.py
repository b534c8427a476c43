"""Python binding of libzfp_hip.so (the MI355X ZFP codec) over its C ABI.

Mirrors zfp-rs's Rust API for the compress/decompress path:
  ZfpConfig.fixed_rate / fixed_precision / fixed_accuracy / reversible / expert
  (src/config.rs:266-354), ZfpConfig.with_rounding + ZfpRounding
  (src/config.rs:115-149, 444-463), and compress / decompress
  (src/bitstream/ops.rs:98-136, 218-255) returning byte counts or raising the
  reference's error kinds (BufferTooSmall, Truncated, Field).

Arrays may be numpy arrays (host memory, staged through HBM) or torch tensors
on the GPU (coded in place).  Everything runs on the GPU: the library has no
CPU codec, and this module raises if the library is not built.
"""
import ctypes
import math
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libzfp_hip.so"

ZFP_MIN_BITS, ZFP_MAX_BITS, ZFP_MAX_PREC, ZFP_MIN_EXP = 1, 16658, 64, -1074
ZFP_HEADER_MAGIC, ZFP_HEADER_META, ZFP_HEADER_MODE, ZFP_HEADER_FULL = 1, 2, 4, 7
TYPE_INT32, TYPE_INT64, TYPE_FLOAT, TYPE_DOUBLE = 1, 2, 3, 4
EXEC_SERIAL, EXEC_OMP, EXEC_CUDA, EXEC_HIP = 0, 1, 2, 3
ROUND_FIRST, ROUND_NEVER, ROUND_LAST = -1, 0, 1

_NP_TYPE = {np.dtype(np.int32): TYPE_INT32, np.dtype(np.int64): TYPE_INT64,
            np.dtype(np.float32): TYPE_FLOAT, np.dtype(np.float64): TYPE_DOUBLE}


class ZfpError(Exception):
    pass


class BufferTooSmall(ZfpError):
    pass


class Truncated(ZfpError):
    pass


class FieldError(ZfpError):
    pass


class zfp_field(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("nx", ctypes.c_size_t), ("ny", ctypes.c_size_t),
                ("nz", ctypes.c_size_t), ("nw", ctypes.c_size_t), ("sx", ctypes.c_ssize_t),
                ("sy", ctypes.c_ssize_t), ("sz", ctypes.c_ssize_t), ("sw", ctypes.c_ssize_t),
                ("data", ctypes.c_void_p)]


class zfp_execution(ctypes.Structure):
    _fields_ = [("policy", ctypes.c_int), ("params", ctypes.c_void_p)]


class zfp_stream(ctypes.Structure):
    _fields_ = [("minbits", ctypes.c_uint), ("maxbits", ctypes.c_uint), ("maxprec", ctypes.c_uint),
                ("minexp", ctypes.c_int32), ("stream", ctypes.c_void_p), ("exec", zfp_execution)]


# name -> (restype, argtypes) of every C entry point include/zfp.h and
# include/zfp_hip.h declare
P = ctypes.POINTER
_VP, _SZ, _U, _U64, _I32, _D, _B = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint,
                                    ctypes.c_uint64, ctypes.c_int32, ctypes.c_double, ctypes.c_int32)
_FP, _SP = P(zfp_field), P(zfp_stream)
SIGNATURES = {
    "stream_open": (_VP, [_VP, _SZ]), "stream_close": (None, [_VP]), "stream_clone": (_VP, [_VP]),
    "stream_alignment": (_SZ, []), "stream_data": (_VP, [_VP]), "stream_size": (_SZ, [_VP]),
    "stream_capacity": (_SZ, [_VP]), "stream_stride_block": (_SZ, [_VP]),
    "stream_stride_delta": (ctypes.c_ssize_t, [_VP]), "stream_read_bit": (_U, [_VP]),
    "stream_write_bit": (_U, [_VP, _U]), "stream_read_bits": (_U64, [_VP, _SZ]),
    "stream_write_bits": (_U64, [_VP, _U64, _SZ]), "stream_rtell": (_U64, [_VP]),
    "stream_wtell": (_U64, [_VP]), "stream_rewind": (None, [_VP]), "stream_rseek": (None, [_VP, _U64]),
    "stream_wseek": (None, [_VP, _U64]), "stream_skip": (None, [_VP, _U64]),
    "stream_pad": (None, [_VP, _U64]), "stream_align": (_SZ, [_VP]), "stream_flush": (_SZ, [_VP]),
    "stream_copy": (None, [_VP, _VP, _U64]),
    "zfp_field_alloc": (_FP, []), "zfp_field_1d": (_FP, [_VP, ctypes.c_int, _SZ]),
    "zfp_field_2d": (_FP, [_VP, ctypes.c_int, _SZ, _SZ]),
    "zfp_field_3d": (_FP, [_VP, ctypes.c_int, _SZ, _SZ, _SZ]),
    "zfp_field_4d": (_FP, [_VP, ctypes.c_int, _SZ, _SZ, _SZ, _SZ]),
    "zfp_field_free": (None, [_FP]), "zfp_field_pointer": (_VP, [_FP]), "zfp_field_begin": (_VP, [_FP]),
    "zfp_field_type": (ctypes.c_int, [_FP]), "zfp_field_precision": (_U, [_FP]),
    "zfp_field_dimensionality": (_U, [_FP]), "zfp_field_size": (_SZ, [_FP, P(_SZ)]),
    "zfp_field_size_bytes": (_SZ, [_FP]), "zfp_field_blocks": (_SZ, [_FP]),
    "zfp_field_stride": (_B, [_FP, P(ctypes.c_ssize_t)]), "zfp_field_is_contiguous": (_B, [_FP]),
    "zfp_field_metadata": (_U64, [_FP]), "zfp_field_set_pointer": (None, [_FP, _VP]),
    "zfp_field_set_type": (ctypes.c_int, [_FP, ctypes.c_int]),
    "zfp_field_set_size_1d": (None, [_FP, _SZ]), "zfp_field_set_size_2d": (None, [_FP, _SZ, _SZ]),
    "zfp_field_set_size_3d": (None, [_FP, _SZ, _SZ, _SZ]),
    "zfp_field_set_size_4d": (None, [_FP, _SZ, _SZ, _SZ, _SZ]),
    "zfp_field_set_stride_1d": (None, [_FP, ctypes.c_ssize_t]),
    "zfp_field_set_stride_2d": (None, [_FP, ctypes.c_ssize_t, ctypes.c_ssize_t]),
    "zfp_field_set_stride_3d": (None, [_FP, ctypes.c_ssize_t, ctypes.c_ssize_t, ctypes.c_ssize_t]),
    "zfp_field_set_stride_4d": (None, [_FP] + [ctypes.c_ssize_t] * 4),
    "zfp_field_set_metadata": (_B, [_FP, _U64]),
    "zfp_type_size": (_SZ, [ctypes.c_int]), "zfp_stream_open": (_SP, [_VP]),
    "zfp_stream_close": (None, [_SP]), "zfp_stream_bit_stream": (_VP, [_SP]),
    "zfp_stream_compression_mode": (ctypes.c_int, [_SP]), "zfp_stream_rate": (_D, [_SP, _U]),
    "zfp_stream_precision": (_U, [_SP]), "zfp_stream_accuracy": (_D, [_SP]),
    "zfp_stream_mode": (_U64, [_SP]), "zfp_stream_params": (None, [_SP, P(_U), P(_U), P(_U), P(_I32)]),
    "zfp_stream_compressed_size": (_SZ, [_SP]),
    "zfp_block_maximum_size": (_SZ, [ctypes.c_int, _U, _B]),
    "zfp_stream_maximum_size": (_SZ, [_SP, _FP]), "zfp_write_header": (_SZ, [_SP, _FP, _U]),
    "zfp_read_header": (_SZ, [_SP, _FP, _U]), "zfp_compress": (_SZ, [_SP, _FP]),
    "zfp_decompress": (_SZ, [_SP, _FP]), "zfp_stream_rewind": (None, [_SP]),
    "zfp_stream_set_bit_stream": (None, [_SP, _VP]), "zfp_stream_set_reversible": (None, [_SP]),
    "zfp_stream_set_rate": (_D, [_SP, _D, ctypes.c_int, _U, _B]),
    "zfp_stream_set_precision": (_U, [_SP, _U]), "zfp_stream_set_accuracy": (_D, [_SP, _D]),
    "zfp_stream_set_mode": (ctypes.c_int, [_SP, _U64]),
    "zfp_stream_set_params": (_B, [_SP, _U, _U, _U, _I32]),
    "zfp_stream_execution": (ctypes.c_int, [_SP]), "zfp_stream_omp_threads": (_U, [_SP]),
    "zfp_stream_omp_chunk_size": (_U, [_SP]), "zfp_stream_set_execution": (_B, [_SP, ctypes.c_int]),
    "zfp_stream_set_omp_threads": (_B, [_SP, _U]), "zfp_stream_set_omp_chunk_size": (_B, [_SP, _U]),
    "zfp_stream_flush": (_SZ, [_SP]), "zfp_stream_align": (_SZ, [_SP]),
    "zfp_stream_set_rounding": (_B, [_SP, ctypes.c_int, _B]),
    "zfp_stream_rounding": (ctypes.c_int, [_SP, P(_B)]),
    "zfp_stream_set_hip_stream": (_B, [_SP, _VP]),
    "zfp_stream_index_export": (_SZ, [_VP, P(ctypes.c_uint32), _SZ, P(_U64)]),
    "zfp_stream_index_import": (_B, [_VP, P(ctypes.c_uint32), _SZ, _U64]),
    "zfp_stream_set_timing": (_B, [_SP, _B]),
    "zfp_stream_kernel_ms": (_D, [_SP, P(_D)]),
    "zfp_hip_device": (ctypes.c_int, []),
    "zfp_hip_fail_after": (ctypes.c_long, [ctypes.c_long]),
    "zfp_hip_last_error": (ctypes.c_char_p, []),
    "zfp_hip_field_may_alias": (_B, [_FP]),
    "zfp_hip_compress_noalloc": (_SZ, [_SP, _FP]),
    "zfp_hip_noalloc_max_blocks": (_U64, []),
    "zfp_hip_compress_batch": (_SZ, [P(_SP), P(_FP), _SZ, P(_SZ), _U]),
    "zfp_hip_decompress_batch": (_SZ, [P(_SP), P(_FP), _SZ, P(_SZ), _U]),
    "zfp_hip_batch_plan": (_VP, [ctypes.c_int, P(_SP), P(_FP), _SZ, _U]),
    "zfp_hip_batch_run": (_SZ, [_VP, P(_SZ)]),
    "zfp_hip_batch_count": (_SZ, [_VP]),
    "zfp_hip_batch_tiling": (_D, [_VP, P(_U), P(_U)]),
    "zfp_hip_batch_free": (None, [_VP]),
    "zfp_hip_batch_submit": (_VP, [ctypes.c_int, P(_SP), P(_FP), _SZ, _U, _VP]),
    "zfp_hip_batch_wait": (_SZ, [_VP, P(_SZ)]),
    "zfp_hip_batch_ready": (_B, [_VP]),
}
BATCH_DEVICE_FIELDS = 1
BATCH_NO_TUNE = 2
for _name in ("zfp_config_none", "zfp_config_rate", "zfp_config_precision", "zfp_config_accuracy",
              "zfp_config_reversible", "zfp_config_expert"):
    SIGNATURES[_name] = (None, None)  # struct returns: exported, not called from Python
DATA_SYMBOLS = ("stream_word_bits", "zfp_codec_version", "zfp_library_version")

_lib = None


def lib():
    """The loaded library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"libzfp_hip.so is not built at {LIB_PATH}; run __graft_entry__.build()")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            if args is not None:
                fn.restype = res
                fn.argtypes = args
        _lib = L
    return _lib


class ZfpRounding:
    """ZfpRounding (src/config.rs:115-149): NEVER, FIRST(tight), LAST(tight)."""

    def __init__(self, kind=ROUND_NEVER, tight_error=False):
        self.kind, self.tight_error = kind, bool(tight_error)

    @classmethod
    def never(cls):
        return cls(ROUND_NEVER, False)

    @classmethod
    def first(cls, tight_error=False):
        return cls(ROUND_FIRST, tight_error)

    @classmethod
    def last(cls, tight_error=False):
        return cls(ROUND_LAST, tight_error)

    def __repr__(self):
        name = {ROUND_NEVER: "Never", ROUND_FIRST: "First", ROUND_LAST: "Last"}[self.kind]
        return name if self.kind == ROUND_NEVER else f"{name}(tight_error={self.tight_error})"


class ZfpConfig:
    """The four expert parameters plus rounding (src/config.rs:163-174)."""

    def __init__(self, minbits, maxbits, maxprec, minexp, rounding=None):
        self.minbits, self.maxbits, self.maxprec, self.minexp = minbits, maxbits, maxprec, minexp
        self.rounding = rounding or ZfpRounding.never()

    @classmethod
    def fixed_rate(cls, rate, ztype, dims, align=False):
        """fixed_rate (config.rs:266-282, rate_bits :85-102)."""
        if not (rate >= 0):
            raise ValueError("invalid rate")
        bits = math.floor(4 ** dims * rate + 0.5)
        if not (0 <= bits <= 0xffffffff):
            raise ValueError("invalid rate")
        bits = max(bits, {TYPE_FLOAT: 9, TYPE_DOUBLE: 12}.get(ztype, 0))
        if align:
            bits = (bits + 63) // 64 * 64
        if not 1 <= bits <= ZFP_MAX_BITS:
            raise ValueError("invalid rate")
        return cls(bits, bits, ZFP_MAX_PREC, ZFP_MIN_EXP)

    @classmethod
    def fixed_precision(cls, precision):
        p = min(precision, ZFP_MAX_PREC) if precision > 0 else ZFP_MAX_PREC
        return cls(ZFP_MIN_BITS, ZFP_MAX_BITS, p, ZFP_MIN_EXP)

    @classmethod
    def fixed_accuracy(cls, tolerance):
        if tolerance > 0:
            e = 0 if math.isinf(tolerance) else math.frexp(tolerance)[1]
            emin = e - 1
        else:
            emin = ZFP_MIN_EXP
        return cls(ZFP_MIN_BITS, ZFP_MAX_BITS, ZFP_MAX_PREC, emin)

    @classmethod
    def reversible(cls):
        return cls(ZFP_MIN_BITS, ZFP_MAX_BITS, ZFP_MAX_PREC, ZFP_MIN_EXP - 1)

    @classmethod
    def expert(cls, minbits, maxbits, maxprec, minexp):
        if not (minbits <= maxbits and 0 < maxprec <= 64):
            raise ValueError("invalid parameters")
        return cls(minbits, maxbits, maxprec, minexp)

    def with_rounding(self, rounding):
        return ZfpConfig(self.minbits, self.maxbits, self.maxprec, self.minexp, rounding)

    def __repr__(self):
        return (f"ZfpConfig({self.minbits}, {self.maxbits}, {self.maxprec}, {self.minexp}, "
                f"{self.rounding!r})")


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _describe(arr):
    """(ztype, ptr, shape, strides) of a C-contiguous numpy array or torch tensor."""
    if _is_torch(arr):
        import torch
        tmap = {torch.int32: TYPE_INT32, torch.int64: TYPE_INT64, torch.float32: TYPE_FLOAT,
                torch.float64: TYPE_DOUBLE}
        return tmap[arr.dtype], arr.data_ptr(), tuple(arr.shape), tuple(arr.stride())
    ztype = _NP_TYPE[arr.dtype]
    isz = arr.dtype.itemsize
    return ztype, arr.ctypes.data, arr.shape, tuple(s // isz for s in arr.strides)


def make_field(arr):
    """A zfp_field over an array; zfp's x is the array's last axis."""
    ztype, ptr, shape, strides = _describe(arr)
    if not 1 <= len(shape) <= 4:
        raise FieldError("1-D to 4-D arrays only")
    f = zfp_field()
    f.type = ztype
    n = list(reversed(shape)) + [0] * (4 - len(shape))
    s = list(reversed(strides)) + [0] * (4 - len(shape))
    f.nx, f.ny, f.nz, f.nw = n
    f.sx, f.sy, f.sz, f.sw = s
    f.data = ptr
    return f


class Stream:
    """A zfp_stream over a bit stream on a word buffer (host numpy or device torch)."""

    def __init__(self, buffer, config=None):
        self._buffer = buffer
        if _is_torch(buffer):
            ptr, nbytes = buffer.data_ptr(), buffer.numel() * buffer.element_size()
        else:
            ptr, nbytes = buffer.ctypes.data, buffer.nbytes
        L = lib()
        self.bs = L.stream_open(ptr, nbytes)
        if not self.bs:
            raise ZfpError("stream_open failed")
        self.zs = L.zfp_stream_open(self.bs)
        L.zfp_stream_set_execution(self.zs, EXEC_HIP)
        if config is not None:
            self.set_config(config)

    def set_config(self, c):
        L = lib()
        if not L.zfp_stream_set_params(self.zs, c.minbits, c.maxbits, c.maxprec, c.minexp):
            raise ValueError("invalid parameters")
        L.zfp_stream_set_rounding(self.zs, c.rounding.kind, int(c.rounding.tight_error))

    def set_hip_stream(self, handle):
        lib().zfp_stream_set_hip_stream(self.zs, handle)

    def set_timing(self, on=True):
        lib().zfp_stream_set_timing(self.zs, int(on))

    def kernel_ms(self):
        """(block-coding kernel ms, all-kernels ms) of the last call, HIP-event timed."""
        call = ctypes.c_double(0)
        k = lib().zfp_stream_kernel_ms(self.zs, ctypes.byref(call))
        return k, call.value

    def compress(self, arr):
        """Bytes of stream after the flush (ops.rs:218-224), or BufferTooSmall."""
        f = make_field(arr)
        n = lib().zfp_compress(self.zs, ctypes.byref(f))
        if n == 0:
            raise BufferTooSmall("compress failed (buffer too small or invalid field; "
                                 f"last library failure: {last_error()})")
        return n

    def decompress(self, arr):
        """Bytes read through the aligned end (ops.rs:98-104), or Truncated."""
        f = make_field(arr)
        n = lib().zfp_decompress(self.zs, ctypes.byref(f))
        if n == 0:
            raise Truncated("decompress failed (truncated stream or invalid field; "
                            f"last library failure: {last_error()})")
        return n

    def rewind(self):
        lib().zfp_stream_rewind(self.zs)

    def wtell(self):
        return lib().stream_wtell(self.bs)

    def rtell(self):
        return lib().stream_rtell(self.bs)

    def write_bits(self, value, n):
        return lib().stream_write_bits(self.bs, value, n)

    def read_bits(self, n):
        return lib().stream_read_bits(self.bs, n)

    def rseek(self, off):
        lib().stream_rseek(self.bs, off)

    def index(self):
        """(start_bit, per-block bit lengths) of the last variable-rate compression, or None."""
        L = lib()
        start = ctypes.c_uint64(0)
        n = L.zfp_stream_index_export(self.bs, None, 0, ctypes.byref(start))
        if n == 0:
            return None
        lens = np.zeros(n, dtype=np.uint32)
        L.zfp_stream_index_export(self.bs, lens.ctypes.data_as(P(ctypes.c_uint32)), n,
                                  ctypes.byref(start))
        return start.value, lens

    def set_index(self, start_bit, lengths):
        lens = np.ascontiguousarray(lengths, dtype=np.uint32)
        return bool(lib().zfp_stream_index_import(self.bs, lens.ctypes.data_as(P(ctypes.c_uint32)),
                                                  len(lens), start_bit))

    def close(self):
        if getattr(self, "zs", None):
            lib().zfp_stream_close(self.zs)
            self.zs = None
        if getattr(self, "bs", None):
            lib().stream_close(self.bs)
            self.bs = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _batch(fn, streams, arrays, flags):
    n = len(streams)
    if len(arrays) != n:
        raise ValueError("one array per stream")
    fields = [make_field(a) for a in arrays]
    zs = (_SP * n)(*[s.zs for s in streams])
    fps = (_FP * n)(*[ctypes.pointer(f) for f in fields])
    sizes = (ctypes.c_size_t * n)()
    done = fn(zs, fps, n, sizes, flags)
    return done, list(sizes)


class Batch:
    """Prebuilt argument tables for repeated batched calls over the same
    streams and arrays (a serving loop's steady state)."""

    def __init__(self, streams, arrays, outputs=None):
        n = len(streams)
        self.n = n
        self._keep = (streams, arrays, outputs)
        self.zs = (_SP * n)(*[s.zs for s in streams])
        self._fin = [make_field(a) for a in arrays]
        self.fin = (_FP * n)(*[ctypes.pointer(f) for f in self._fin])
        outs = outputs if outputs is not None else arrays
        self._fout = [make_field(a) for a in outs]
        self.fout = (_FP * n)(*[ctypes.pointer(f) for f in self._fout])
        self.sizes = (ctypes.c_size_t * n)()

    def compress(self, flags=BATCH_DEVICE_FIELDS):
        return lib().zfp_hip_compress_batch(self.zs, self.fin, self.n, self.sizes, flags)

    def decompress(self, flags=BATCH_DEVICE_FIELDS):
        return lib().zfp_hip_decompress_batch(self.zs, self.fout, self.n, self.sizes, flags)


class BatchPlan:
    """zfp_hip_batch_plan: a batch planned (and its tiling tuned) once, then
    replayed from a captured HIP graph by run().  `encode` selects compress
    (arrays -> streams) or decompress (streams -> arrays); the streams' cursors
    at construction are where every run starts."""

    def __init__(self, encode, streams, arrays, flags=BATCH_DEVICE_FIELDS):
        n = len(streams)
        if len(arrays) != n:
            raise ValueError("one array per stream")
        self.n = n
        self._keep = (streams, arrays)
        self._fields = [make_field(a) for a in arrays]
        zs = (_SP * n)(*[s.zs for s in streams])
        fps = (_FP * n)(*[ctypes.pointer(f) for f in self._fields])
        self.sizes = (ctypes.c_size_t * n)()
        self.handle = lib().zfp_hip_batch_plan(1 if encode else 0, zs, fps, n, flags)
        if not self.handle:
            raise ValueError("batch not plannable: every array must be device-resident, "
                             "one per stream, with at least one block")

    def run(self):
        """Replay; returns (arrays coded, per-array byte counts)."""
        done = lib().zfp_hip_batch_run(self.handle, self.sizes)
        return done, list(self.sizes)

    def tiling(self):
        """(blocks per workgroup, LDS slice or 0, tuned ms or -1)."""
        t, s = ctypes.c_uint(0), ctypes.c_uint(0)
        ms = lib().zfp_hip_batch_tiling(self.handle, ctypes.byref(t), ctypes.byref(s))
        return t.value, s.value, ms

    def close(self):
        if self.handle:
            lib().zfp_hip_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def last_error():
    """zfp_hip_last_error: why this thread's last failed call failed, or ''."""
    return lib().zfp_hip_last_error().decode()


def compress_batch(streams, arrays, flags=0):
    """zfp_hip_compress_batch: arrays[i] into streams[i] (Stream objects), with
    one group of launches per class of arrays.  Returns (arrays coded, the
    per-array byte counts zfp_compress would return, 0 for a failure)."""
    return _batch(lib().zfp_hip_compress_batch, streams, arrays, flags)


def decompress_batch(streams, arrays, flags=0):
    """zfp_hip_decompress_batch: the inverse of compress_batch."""
    return _batch(lib().zfp_hip_decompress_batch, streams, arrays, flags)


def maximum_size(config, arr):
    """zfp_stream_maximum_size for this config and array (config.rs:642-658)."""
    s = Stream(np.zeros(1, dtype=np.uint64), config)
    try:
        f = make_field(arr)
        return lib().zfp_stream_maximum_size(s.zs, ctypes.byref(f))
    finally:
        s.close()


def compress(arr, config, buffer=None):
    """Compress into a new (or given) word buffer; returns (buffer, nbytes)."""
    if buffer is None:
        buffer = np.zeros((maximum_size(config, arr) + 7) // 8, dtype=np.uint64)
    s = Stream(buffer, config)
    try:
        n = s.compress(arr)
    finally:
        s.close()
    return buffer, n


def decompress(buffer, config, out):
    """Decompress a stream that starts at bit 0 of `buffer` into `out`; returns bytes read."""
    s = Stream(buffer, config)
    try:
        return s.decompress(out)
    finally:
        s.close()
