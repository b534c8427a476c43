"""Helpers restating the reference's C-port test utilities.

- LCG generators: tests/c/encode_block.rs:35-95 (upstream zfp/tests/utils/rand{32,64}.c)
- Jenkins one-at-a-time hashes: tests/c/checksums.rs:14-151
- key computation: tests/c/checksums.rs:189-252
"""
import json
import math
import pathlib

import numpy as np

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
M32 = 0xffffffff
M64 = 0xffffffffffffffff


class Rand32:
    def __init__(self):
        self.x = 5

    def next_unsigned(self):
        self.x = (0x5deece66d * self.x + 0xb) % (1 << 48)
        return (self.x >> 16) & 0x7fffffff

    def next_signed_int(self):
        return self.next_unsigned() - 0x40000000

    def next_signed_f32(self):
        u = (self.next_unsigned() >> 7) & 0x00ffffff
        s = u - 0x00800000
        return np.float32(math.ldexp(float(np.float32(s)), -12))


class Rand64:
    def __init__(self):
        self.x = 5

    def next_unsigned(self):
        self.x = (2862933555777941757 * self.x + 3037000493) & M64
        return self.x & 0x7fffffffffffffff

    def next_signed_int(self):
        return self.next_unsigned() - (1 << 62)

    def next_signed_f64(self):
        u = (self.next_unsigned() >> 11) & 0x001fffffffffffff
        s = u - 0x0010000000000000
        return np.float64(math.ldexp(float(s), -26))


def _hv(val, h):
    h = (h + val) & M32
    h = (h + (h << 10)) & M32
    h ^= h >> 6
    return h


def _hf(h):
    h = (h + (h << 3)) & M32
    h ^= h >> 11
    h = (h + (h << 15)) & M32
    return h


def hash_bitstream(words):
    h1 = h2 = 0
    for w in np.asarray(words, dtype=np.uint64).tolist():
        h1 = _hv(w & M32, h1)
        h2 = _hv((w >> 32) & M32, h2)
    return _hf(h1) + (_hf(h2) << 32)


def hash_array32(vals):
    h = 0
    for v in np.asarray(vals).view(np.uint32).tolist():
        h = _hv(v, h)
    return _hf(h)


def hash_array64(vals):
    h1 = h2 = 0
    for v in np.asarray(vals).view(np.uint64).tolist():
        h1 = _hv(v & M32, h1)
        h2 = _hv((v >> 32) & M32, h2)
    return _hf(h1) + (_hf(h2) << 32)


def compute_key(tt, subject, n, mode, misc):
    """tt: 0 full block, 1 partial, 2 array; subject: 0 input, 1 stream, 2 decoded."""
    k1 = ((((tt << 2) + subject) << 3) + mode) << 4
    k1 += misc
    dims = 4 if n[3] else 3 if n[2] else 2 if n[1] else 1
    if dims == 1:
        k2 = n[0] - 1
    elif dims == 2:
        k2 = ((n[0] - 1) << 24) + (n[1] - 1)
    elif dims == 3:
        k2 = ((((n[0] - 1) << 16) + (n[1] - 1)) << 16) + (n[2] - 1)
    else:
        k2 = ((((((n[0] - 1) << 12) + (n[1] - 1)) << 12) + (n[2] - 1)) << 12) + (n[3] - 1)
    return k1, k2


_TABLES = None


def checksum(dims, tname, key1, key2):
    global _TABLES
    if _TABLES is None:
        _TABLES = json.loads((GOLDEN / "checksums.json").read_text())
    for k1, k2, v in _TABLES[f"{dims}d_{tname}"]:
        if k1 == key1 and k2 == key2:
            return v
    return None


TNAMES = {1: "int32", 2: "int64", 3: "float", 4: "double"}
SPECIAL_F32 = [0x00000000, 0x80000000, 0x00000001, 0x80000001, 0x7f7fffff, 0xff7fffff,
               0x7f800000, 0xff800000, 0x7fc00000, 0x7fa00000]
SPECIAL_F64 = [0x0, 0x8000000000000000, 0x1, 0x8000000000000001, 0x7fefffffffffffff,
               0xffefffffffffffff, 0x7ff0000000000000, 0xfff0000000000000,
               0x7ff8000000000000, 0x7ff4000000000000]


def lcg_block(ztype, n):
    """The C-port suite's seeded random block of n values (encode_block.rs:152-155)."""
    if ztype == 1:
        r = Rand32(); return np.array([r.next_signed_int() for _ in range(n)], dtype=np.int32)
    if ztype == 2:
        r = Rand64(); return np.array([r.next_signed_int() for _ in range(n)], dtype=np.int64)
    if ztype == 3:
        r = Rand32(); return np.array([r.next_signed_f32() for _ in range(n)], dtype=np.float32)
    r = Rand64(); return np.array([r.next_signed_f64() for _ in range(n)], dtype=np.float64)


def strided_array(ztype, dims, dummy=99):
    """make_strided_array of encode_block_strided.rs (SX=2, SY=24, SZ=192,
    SW=2304; the 4-D layout at :685-712)."""
    dtype = {1: np.int32, 2: np.int64, 3: np.float32, 4: np.float64}[ztype]
    SX, SY, SZ, SW = 2, 24, 192, 2304
    cx = 4 * SX
    cy = SY // SX if dims >= 2 else 1
    cz = SZ // SY if dims >= 3 else 1
    cw = SW // SZ if dims >= 4 else 1
    data = np.full(cx * cy * cz * cw, dummy, dtype=dtype)
    vals = iter(lcg_block(ztype, 4 ** dims))
    for l in range(cw):
        for k in range(cz):
            for j in range(cy):
                for i in range(cx):
                    if (i % 2 == 0 and (dims < 2 or j % (cy // 4) == 0) and (dims < 3 or k % (cz // 4) == 0)
                            and (dims < 4 or l % (cw // 4) == 0)):
                        data[cx * cy * cz * l + cx * cy * k + cx * j + i] = next(vals)
    strides = [SX, SY, SZ, SW][:dims]
    return data, strides


def hash_strided(arr, dims, strides):
    """hashStridedArray over the 4^d strided positions (checksums.rs:79-151)."""
    idx = []
    for w in range(4 if dims >= 4 else 1):
        for z in range(4 if dims >= 3 else 1):
            for y in range(4 if dims >= 2 else 1):
                for x in range(4):
                    off = x * strides[0]
                    if dims >= 2: off += y * strides[1]
                    if dims >= 3: off += z * strides[2]
                    if dims >= 4: off += w * strides[3]
                    idx.append(off)
    vals = arr[np.array(idx)]
    return hash_array32(vals) if arr.dtype.itemsize == 4 else hash_array64(vals)
