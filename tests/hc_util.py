"""ctypes binding of tests/native/libhostcheck.so: the GPU block codec built for the CPU."""
import ctypes
import pathlib

import numpy as np

LIB = pathlib.Path(__file__).resolve().parent / "native" / "libhostcheck.so"


class HcCfg(ctypes.Structure):
    _fields_ = [("minbits", ctypes.c_uint32), ("maxbits", ctypes.c_uint32),
                ("maxprec", ctypes.c_uint32), ("minexp", ctypes.c_int32),
                ("rounding", ctypes.c_int32), ("tight", ctypes.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(str(LIB))
        P = ctypes.POINTER
        L.hc_compress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                  P(ctypes.c_int64), P(ctypes.c_int64), P(HcCfg),
                                  P(ctypes.c_uint64), ctypes.c_uint64, ctypes.c_uint64,
                                  ctypes.c_uint32, P(ctypes.c_uint32), ctypes.c_uint32, ctypes.c_uint32]
        L.hc_compress.restype = ctypes.c_uint64
        L.hc_divide.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64]
        L.hc_divide.restype = None
        L.hc_decompress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                    P(ctypes.c_int64), P(ctypes.c_int64), P(HcCfg),
                                    P(ctypes.c_uint64), ctypes.c_uint64, ctypes.c_uint64,
                                    ctypes.c_uint32, P(ctypes.c_uint32), ctypes.c_uint32]
        L.hc_decompress.restype = ctypes.c_uint64
        L.hc_alias_scatter.argtypes = [ctypes.c_void_p, ctypes.c_void_p, P(ctypes.c_int64),
                                       P(ctypes.c_int64), ctypes.c_int64, ctypes.c_int, ctypes.c_uint32]
        L.hc_alias_scatter.restype = ctypes.c_int
        L.hc_compress_1p.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                     P(ctypes.c_int64), P(ctypes.c_int64), P(HcCfg),
                                     P(ctypes.c_uint64), ctypes.c_uint64, ctypes.c_uint64,
                                     P(ctypes.c_uint32), ctypes.c_uint32, ctypes.c_uint32]
        L.hc_compress_1p.restype = ctypes.c_uint64
        L.hc_walk.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(HcCfg), P(ctypes.c_uint64),
                              ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, P(ctypes.c_uint64)]
        L.hc_walk.restype = ctypes.c_uint64
        _lib = L
    return _lib


def _dims(arr):
    n = list(reversed(arr.shape)) + [1] * (4 - arr.ndim)
    s = [1, n[0], n[0] * n[1], n[0] * n[1] * n[2]]
    return (ctypes.c_int64 * 4)(*n), (ctypes.c_int64 * 4)(*s)


def cfg_of(oc):
    return HcCfg(oc.minbits, oc.maxbits, oc.maxprec, oc.minexp, oc.rounding, oc.tight)


def compress(arr, ztype, oc, words, start_bit, fixed_bits, rev, lds_T=0, lds_cap=0):
    """lds_T > 0 runs the LDS path with workgroups of lds_T blocks; lds_cap > 0
    sets the words a variable-rate workgroup may stage (else the default)."""
    n, s = _dims(arr)
    nb = 1
    for v in list(reversed(arr.shape)):
        nb *= (v + 3) // 4
    lens = np.zeros(nb, dtype=np.uint32)
    c = cfg_of(oc)
    end = lib().hc_compress(ztype, arr.ndim, int(rev), arr.ctypes.data, n, s, ctypes.byref(c),
                            words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(words),
                            start_bit, fixed_bits,
                            lens.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), lds_T, lds_cap)
    return end, lens


def decompress(shape, dtype, ztype, oc, words, start_bit, fixed_bits, lens, rev, lds_T=0):
    """Returns (values, loaded) where loaded is one past the highest word read
    (LDS path only; 0 otherwise)."""
    out = np.zeros(shape, dtype=dtype)
    n, s = _dims(out)
    c = cfg_of(oc)
    lp = lens.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) if lens is not None else None
    loaded = lib().hc_decompress(ztype, out.ndim, int(rev), out.ctypes.data, n, s, ctypes.byref(c),
                                 words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(words),
                                 start_bit, fixed_bits, lp, lds_T)
    return out, loaded


def divide(n, d):
    """n // d (uint32 array) through the kernels' invariant-divisor division."""
    import numpy as np
    n = np.ascontiguousarray(n, dtype=np.uint32)
    q = np.empty_like(n)
    lib().hc_divide(n.ctypes.data, int(d), q.ctypes.data, n.size)
    return q


def walk(ztype, dims, oc, words, start_bit, count, rev):
    """The kernels' length reader: (offsets of `count` blocks, end bit)."""
    offs = np.zeros(count + 1, dtype=np.uint64)
    c = cfg_of(oc)
    end = lib().hc_walk(ztype, dims, int(rev), ctypes.byref(c),
                        words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(words), start_bit,
                        count, offs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    offs[count] = end
    return offs


def compress_1p(arr, ztype, oc, words, start_bit, rev, T=0, cap=0):
    """The one-pass variable-rate compress (T, cap: 0 for the plan's)."""
    n, s = _dims(arr)
    nb = 1
    for v in list(reversed(arr.shape)):
        nb *= (v + 3) // 4
    lens = np.zeros(nb, dtype=np.uint32)
    c = cfg_of(oc)
    end = lib().hc_compress_1p(ztype, arr.ndim, int(rev), arr.ctypes.data, n, s, ctypes.byref(c),
                               words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(words),
                               start_bit, lens.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), T, cap)
    return end, lens
