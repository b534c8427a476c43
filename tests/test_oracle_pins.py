"""Pins of the oracle beyond the block checksums: the reference's own
known answers and the closed forms its tests check.

Each test names the reference test it restates (file:line).  Where the
reference compares against C libzfp or a helper of its own, the expected value
here is the documented answer or an independent closed form (frexp, exact
integer scaling, the transform matrices, naive bit extraction), never the
oracle's own formula retyped.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_py as O

L = O.lib


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def compress_bytes(arr, cfg, start_bits=0):
    n, err, words = O.compress(np.ascontiguousarray(arr), cfg, start_bits=start_bits)
    assert err == 0
    return n, words


def round_trip(arr, cfg):
    n, words = compress_bytes(arr, cfg)
    out, m, err = O.decompress(words, arr.shape, arr.dtype, cfg)
    assert err == 0
    return words[: n // 8].tobytes(), out


def doc_values():
    """tests/proptest/differences.rs:14-18: (0.37 * i).sin() * 100 for i in 0..16."""
    return np.array([math.sin(0.37 * float(i)) * 100.0 for i in range(16)], dtype=np.float64)


# ------------------------------------------------------------ documented sizes

@pytest.mark.parametrize("cfg,nbytes", [
    ((1, 11, 64, -1074), 8), ((12, 12, 64, -1074), 8),
    ((1, 18, 64, -1075), 16), ((19, 19, 64, -1075), 16)])
def test_max_bits_below_the_header_sizes(cfg, nbytes):
    """docs/differences-from-c.md "max_bits below the block header" (zfp-rs
    column); tests/proptest/differences.rs:45-60."""
    n, _ = compress_bytes(doc_values(), O.config(*cfg))
    assert n == nbytes


def test_maximum_size_holds_blocks_longer_than_max_bits():
    """differences.rs:147-157: maximum_size of expert(1, 11, 64, -1074) for 16
    f64 values is 32 bytes, and the stream is 8."""
    cfg = O.config(1, 11, 64, -1074)
    dims = (ctypes.c_size_t * 3)(16, 0, 0)
    assert L().orc_maximum_size(ctypes.byref(cfg), O.DOUBLE, dims) == 32
    assert compress_bytes(doc_values(), cfg)[0] == 8


@pytest.mark.parametrize("minbits,nbytes", [(100, 104), (1, 96)])
def test_reversible_zero_blocks_are_padded(minbits, nbytes):
    """differences.rs:160-208: with the first four values zero, the reversible
    stream is 104 bytes for min_bits 100 (96 for 1) and decodes losslessly."""
    data = doc_values()
    data[:4] = 0.0
    cfg = O.config(minbits, 16658, 64, -1075)
    b, out = round_trip(data, cfg)
    assert len(b) == nbytes
    assert np.array_equal(out.view(np.uint64), data.view(np.uint64))


@pytest.mark.parametrize("kind,cfg,nbytes,err", [
    ("f32", "prec16", 32, "3.1e-35"), ("f32", "rate8", 16, "6.1e-33"),
    ("f64", "prec16", 40, "1.1e-302")])
def test_tiny_blocks_keep_their_values(kind, cfg, nbytes, err):
    """differences.rs:210-235 (zfp-rs column): stream sizes and the largest
    error, to two significant figures, of blocks below C's cast threshold."""
    if kind == "f32":
        data = np.array([(np.float32(i) - np.float32(7.5)) * np.float32(1.0e-31) for i in range(16)],
                        dtype=np.float32)
    else:
        data = doc_values() * 1.0e-300
    ztype = O.FLOAT if kind == "f32" else O.DOUBLE
    c = O.fixed_precision(16) if cfg == "prec16" else O.fixed_rate(8, ztype, 1)
    b, out = round_trip(data, c)
    assert len(b) == nbytes
    e = max(abs(float(a) - float(o)) for a, o in zip(data, out))
    assert f"{e:.1e}" == err


def test_basic_example():
    """examples/basic.rs:1-9: [0, 1, 2, 3, 4] f64 at fixed rate 8 compresses to
    8 bytes and decompresses to the same values (reading 8 bytes)."""
    data = np.arange(5, dtype=np.float64)
    cfg = O.fixed_rate(8, O.DOUBLE, 1)
    n, words = compress_bytes(data, cfg)
    assert n == 8
    out, m, err = O.decompress(words, data.shape, data.dtype, cfg)
    assert err == 0 and m == 8
    assert np.array_equal(out, data)


def test_extreme_expert_exponents_do_not_wrap_precision():
    """tests/numerical_boundaries.rs:18-28: min_exp = i32::MIN codes as -1075;
    i32::MAX gives an 8-byte stream that decodes to zeros."""
    data = np.full(4, 1e-100, dtype=np.float64)
    lo = round_trip(data, O.config(1, 16658, 64, -2 ** 31))
    dflt = round_trip(data, O.config(1, 16658, 64, -1075))
    assert lo[0] == dflt[0] and np.array_equal(lo[1], dflt[1])
    hi = round_trip(data, O.config(1, 16658, 64, 2 ** 31 - 1))
    assert len(hi[0]) == 8 and not hi[1].any()


# ------------------------------------------------------------ rounding relations

def _block64(rng):
    return rng.uniform(-1.0e6, 1.0e6, size=(4, 4, 4)).astype(np.float64)


@pytest.mark.parametrize("seed", range(8))
def test_round_last_matches_its_precision_peer(seed):
    """tests/proptest/rounding.rs:59-77: ROUND_LAST is decode-only, so its stream
    equals Never's; with tight error, Never's at twice the tolerance."""
    rng = np.random.default_rng(seed)
    data = _block64(rng)
    e = int(rng.integers(-40, 20))
    base = O.fixed_accuracy(math.ldexp(1.0, e))
    for tight in (False, True):
        ref = O.fixed_accuracy(math.ldexp(1.0, e + 1)) if tight else base
        want, _ = round_trip(data, ref)
        got, _ = round_trip(data, O.config(base.minbits, base.maxbits, base.maxprec, base.minexp,
                                            O.ROUND_LAST, tight))
        assert got == want


@pytest.mark.parametrize("seed", range(8))
def test_fixed_accuracy_holds_for_every_rounding(seed):
    """rounding.rs:79-98: every rounding stays within the tolerance."""
    rng = np.random.default_rng(100 + seed)
    data = _block64(rng)
    e = int(rng.integers(-20, 11))
    tol = math.ldexp(1.0, e)
    for rnd, tight in ((O.ROUND_NEVER, False), (O.ROUND_FIRST, False), (O.ROUND_FIRST, True),
                       (O.ROUND_LAST, False), (O.ROUND_LAST, True)):
        c = O.fixed_accuracy(tol)
        _, out = round_trip(data, O.config(c.minbits, c.maxbits, c.maxprec, c.minexp, rnd, tight))
        assert np.max(np.abs(out - data)) <= tol


@pytest.mark.parametrize("seed", range(8))
def test_round_first_and_tight_error_change_the_stream(seed):
    """rounding.rs:100-124: First and First+tight each change the stream."""
    rng = np.random.default_rng(200 + seed)
    data = _block64(rng)
    e = int(rng.integers(-20, 11))
    c = O.fixed_accuracy(math.ldexp(1.0, e))
    never, _ = round_trip(data, c)
    first, _ = round_trip(data, O.config(c.minbits, c.maxbits, c.maxprec, c.minexp, O.ROUND_FIRST, False))
    tight, _ = round_trip(data, O.config(c.minbits, c.maxbits, c.maxprec, c.minexp, O.ROUND_FIRST, True))
    assert first != never and tight != first


@pytest.mark.parametrize("seed", range(6))
def test_reversible_is_lossless_under_every_rounding(seed):
    """rounding.rs:126-160: reversible streams and values ignore rounding."""
    rng = np.random.default_rng(300 + seed)
    if seed % 2:
        data = rng.integers(-1000, 1000, size=(4, 4, 4)).astype(np.float64)
    else:
        data = rng.standard_normal((4, 4, 4)) * 10.0 ** rng.integers(-30, 30)
    want, _ = round_trip(data, O.reversible())
    for rnd, tight in ((O.ROUND_FIRST, False), (O.ROUND_FIRST, True), (O.ROUND_LAST, False),
                       (O.ROUND_LAST, True)):
        got, out = round_trip(data, O.config(1, 16658, 64, -1075, rnd, tight))
        assert got == want
        assert np.array_equal(out.view(np.uint64), data.view(np.uint64))


# ------------------------------------------------------------ rounding vectors

def test_fwd_round_vectors():
    """src/codec/encode/core.rs:695-720."""
    b = np.array([1, -2, 3, -4], dtype=np.int32)
    L().orc_fwd_round_i32(ptr(b), 4, 32)
    assert b.tolist() == [1, -2, 3, -4]
    b = np.array([1, -2, 3, -4], dtype=np.int64)
    L().orc_fwd_round_i64(ptr(b), 4, 64)
    assert b.tolist() == [1, -2, 3, -4]
    odd = np.zeros(2, dtype=np.int32)
    L().orc_fwd_round_i32(ptr(odd), 2, 3)
    assert odd.tolist() == [0x2aaaaaaa >> 3] * 2
    even = np.zeros(2, dtype=np.int32)
    L().orc_fwd_round_i32(ptr(even), 2, 4)
    assert even.tolist() == [-(0x2aaaaaaa >> 4)] * 2
    w = np.array([2 ** 31 - 1], dtype=np.int32)
    L().orc_fwd_round_i32(ptr(w), 1, 1)
    assert int(w[0]) == ((2 ** 31 - 1 + (0x2aaaaaaa >> 1) + 2 ** 31) % 2 ** 32) - 2 ** 31


def test_inv_round_vectors():
    """src/codec/decode/core.rs:363-404."""
    d = np.array([1, 2, 3], dtype=np.uint32)
    L().orc_inv_round_u32(ptr(d), 3, 1, 31)
    assert d.tolist() == [1, 2, 3]
    d = np.array([1, 2, 3], dtype=np.uint64)
    L().orc_inv_round_u64(ptr(d), 3, 1, 63)
    assert d.tolist() == [1, 2, 3]
    d = np.zeros(4, dtype=np.uint32)
    L().orc_inv_round_u32(ptr(d), 4, 2, 4)
    hi, lo = 0x2aaaaaaa >> 4, 0x55555555 >> 4
    assert d.tolist() == [hi, hi, lo, lo]
    lo8, hi8 = 0x55555555 >> 8, 0x2aaaaaaa >> 8
    for m, want in ((0, [lo8] * 3), (3, [hi8] * 3), (99, [hi8] * 3)):
        d = np.zeros(3, dtype=np.uint32)
        L().orc_inv_round_u32(ptr(d), 3, m, 8)
        assert d.tolist() == want
    d = np.array([2 ** 32 - 1], dtype=np.uint32)
    L().orc_inv_round_u32(ptr(d), 1, 0, 8)
    assert int(d[0]) == (2 ** 32 - 1 + lo8) % 2 ** 32


def test_precision_and_truncate_vectors():
    """encode/core.rs:520-538 (truncate_f64) and :722-730 (precision_f)."""
    assert L().orc_precision_f(-120, 64, -149, 1, 0) == 33
    assert L().orc_precision_f(-120, 64, -149, 1, 1) == 32
    assert L().orc_precision_f(0, 12, -149, 3, 0) == 12
    assert L().orc_precision_f(0, 12, -149, 3, 1) == 12
    two63 = 9223372036854775808.0
    imin = -2 ** 63
    for v, want in ((0.0, 0), (-0.0, 0), (1.9, 1), (-1.9, -1), (-two63, imin), (two63, imin),
                    (1.7976931348623157e308, imin), (math.inf, imin), (-math.inf, imin),
                    (math.nan, imin), (4611686018427387904.5, 4611686018427387904)):
        assert L().orc_truncate_f64(v) == want, v


# ------------------------------------------------------------ exponent and casts

def _xorshift(seed):
    x = seed
    m = (1 << 64) - 1

    def nxt():
        nonlocal x
        x ^= (x << 13) & m
        x ^= x >> 7
        x ^= (x << 17) & m
        return x
    return nxt


def _frexp_exponent(vals, floor, empty):
    """The reference's helper (encode/core.rs:540-562): frexp of the largest
    |x| with NaNs skipped, at least `floor`, or `empty` for a zero block."""
    mx = 0.0
    for v in vals:
        a = abs(float(v))
        if mx < a:  # NaN compares false
            mx = a
    return max(math.frexp(mx)[1], floor) if mx > 0.0 else empty


def _f32(bits):
    return np.array([bits], dtype=np.uint32).view(np.float32)[0]


def _f64(bits):
    return np.array([bits], dtype=np.uint64).view(np.float64)[0]


def test_exponent_block_matches_frexp():
    """encode/core.rs:565-603, same special values and generator."""
    f32max = np.finfo(np.float32).max
    tiny = np.finfo(np.float32).tiny
    specials = [0.0, -0.0, tiny, -tiny / np.float32(3.0), _f32(1), f32max, -f32max,
                np.inf, -np.inf, np.nan, -np.nan, _f32(0x7fffffff), 1.0, 0.75, -3.5e-20]
    specials = [np.float32(v) for v in specials]
    ex32 = L().orc_exponent_block_f32
    ex64 = L().orc_exponent_block_f64
    for a in specials:
        for b in specials:
            blk = np.array([a, b, np.float32(0.5) * a, -b], dtype=np.float32)
            assert ex32(ptr(blk), 4) == _frexp_exponent(blk, -126, -127)
            blk = np.array([a, b, 1e-310, -float(b)], dtype=np.float64)
            assert ex64(ptr(blk), 4) == _frexp_exponent(blk, -1022, -1023)
    assert ex32(ptr(np.array([np.inf, 1, 0, -2], dtype=np.float32)), 4) == 0
    assert ex64(ptr(np.array([-np.inf, 1, 0, -2], dtype=np.float64)), 4) == 0
    nxt = _xorshift(0x2545F4914F6CDD1D)
    for _ in range(10000):
        blk = np.array([nxt() & 0xffffffff for _ in range(4)], dtype=np.uint32).view(np.float32)
        assert ex32(ptr(blk), 4) == _frexp_exponent(blk, -126, -127)
        blk = np.array([nxt() for _ in range(4)], dtype=np.uint64).view(np.float64)
        assert ex64(ptr(blk), 4) == _frexp_exponent(blk, -1022, -1023)


def test_fwd_cast_scales_exactly():
    """encode/core.rs:605-660: fwd_cast is trunc(x * 2^(30 - emax)) (f64:
    2^(62 - emax)) computed exactly, tiny blocks included."""
    nxt = _xorshift(0x9E3779B97F4A7C15)
    tiny = [0, 0]
    ex32, ex64 = L().orc_exponent_block_f32, L().orc_exponent_block_f64
    for top in range(0, 0xff, 3):
        for _ in range(60):
            vals = []
            for _ in range(4):
                r = nxt()
                if r >> 61:
                    e = (r >> 16) % (top + 1)
                    vals.append(((nxt() & 0xffffffff) & 0x807fffff) | (e << 23))
                else:
                    vals.append(0)
            blk = np.array(vals, dtype=np.uint32).view(np.float32)
            emax = ex32(ptr(blk), 4)
            if emax == -127:
                continue
            tiny[0] += emax < -97
            got = np.zeros(4, dtype=np.int32)
            L().orc_fwd_cast_f32(ptr(got), ptr(blk), 4, emax)
            want = [int(math.ldexp(float(f), 30 - emax)) for f in blk]
            assert got.tolist() == want
    for top in range(0, 0x7ff, 7):
        for _ in range(20):
            vals = []
            for _ in range(4):
                r = nxt()
                if r >> 61:
                    e = (r >> 16) % (top + 1)
                    vals.append((nxt() & 0x800fffffffffffff) | (e << 52))
                else:
                    vals.append(0)
            blk = np.array(vals, dtype=np.uint64).view(np.float64)
            emax = ex64(ptr(blk), 4)
            if emax == -1023:
                continue
            tiny[1] += emax < -961
            got = np.zeros(4, dtype=np.int64)
            L().orc_fwd_cast_f64(ptr(got), ptr(blk), 4, emax)
            want = [int(math.ldexp(float(f), 62 - emax)) for f in blk]
            assert got.tolist() == want
    assert tiny[0] > 200 and tiny[1] > 200, tiny


def test_fwd_cast_of_nan_in_a_tiny_block():
    """encode/core.rs:672-683."""
    blk = np.array([3.0e-31, np.nan, 0.0, -3.0e-31], dtype=np.float32)
    emax = L().orc_exponent_block_f32(ptr(blk), 4)
    got = np.zeros(4, dtype=np.int32)
    L().orc_fwd_cast_f32(ptr(got), ptr(blk), 4, emax)
    assert got[1] == -2 ** 31 and got[2] == 0 and got[0] == -got[3] and got[0] > 0


@pytest.mark.parametrize("seed", range(4))
def test_inv_cast_is_a_rounded_power_of_two_scaling(seed):
    """decode/core.rs:162-175: x = ldexp(1, emax - 30) * float(i), the int to
    float conversion rounding to nearest even, the scale possibly subnormal."""
    rng = np.random.default_rng(seed)
    i32 = rng.integers(-2 ** 31, 2 ** 31, size=64, dtype=np.int64).astype(np.int32)
    i64 = rng.integers(-2 ** 63, 2 ** 63, size=64, dtype=np.int64)
    for emax in (-160, -149, -126, -30, 0, 30, 127):
        got = np.zeros(64, dtype=np.float32)
        L().orc_inv_cast_f32(ptr(i32), ptr(got), 64, emax)
        sc = np.float32(math.ldexp(1.0, emax - 30)) if emax - 30 >= -149 else np.float32(0.0)
        want = (sc * i32.astype(np.float32)).astype(np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), emax
    for emax in (-1100, -1074, -1022, 0, 1023):
        got = np.zeros(64, dtype=np.float64)
        L().orc_inv_cast_f64(ptr(i64), ptr(got), 64, emax)
        sc = math.ldexp(1.0, emax - 62)
        want = sc * i64.astype(np.float64)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), emax


# ------------------------------------------------------------ transforms

FWD = np.array([[4, 4, 4, 4], [5, 1, -1, -5], [-4, 4, 4, -4], [-2, 6, -6, 2]]) / 16.0
INV = np.array([[4, 6, -4, -1], [4, 2, 4, 5], [4, -2, 4, -5], [4, -6, -4, 1]]) / 4.0


def _apply(mat, v, dims):
    """The separable transform: mat along x, y, z of a block (x fastest)."""
    a = v.reshape((4,) * dims).astype(np.float64)
    for axis in range(dims):
        a = np.moveaxis(np.tensordot(mat, np.moveaxis(a, dims - 1 - axis, 0), axes=1), 0,
                        dims - 1 - axis)
    return a.reshape(-1)


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_transforms_follow_their_matrices(dims):
    """transform.rs:63-72 and :87-96: the lifts are the documented matrices up
    to the rounding of their halvings, along x, then y, then z."""
    assert np.allclose(INV @ FWD, np.eye(4))
    rng = np.random.default_rng(dims)
    n = 4 ** dims
    bound = {1: 2.0, 2: 4.0, 3: 8.0}[dims]
    for _ in range(200):
        v = rng.integers(-2 ** 20, 2 ** 20, size=n)
        for dt, fn in ((np.int32, L().orc_xform_i32), (np.int64, L().orc_xform_i64)):
            p = v.astype(dt)
            fn(ptr(p), n, 0)
            assert np.max(np.abs(p - _apply(FWD, v, dims))) <= bound
            q = v.astype(dt)
            fn(ptr(q), n, 1)
            assert np.max(np.abs(q - _apply(INV, v, dims))) <= 8 * bound


@pytest.mark.parametrize("n", [4, 16, 64])
def test_reversible_transforms_invert(n):
    """transform.rs:308-315, plus full-range (wrapping) blocks."""
    blk = np.array([(i - 20) ** 3 for i in range(n)], dtype=np.int32)
    p = blk.copy()
    L().orc_xform_i32(ptr(p), n, 2)
    assert not np.array_equal(p, blk)
    L().orc_xform_i32(ptr(p), n, 3)
    assert np.array_equal(p, blk)
    rng = np.random.default_rng(n)
    for dt, fn in ((np.int32, L().orc_xform_i32), (np.int64, L().orc_xform_i64)):
        info = np.iinfo(dt)
        v = rng.integers(info.min, info.max, size=n, dtype=dt, endpoint=True)
        p = v.copy()
        fn(ptr(p), n, 2)
        fn(ptr(p), n, 3)
        assert np.array_equal(p, v)


def test_planes_match_naive_extraction():
    """bitplane.rs:859-903: plane k holds bit k of every coefficient."""
    rng = np.random.default_rng(9)
    for n in (4, 16, 64):
        u = rng.integers(0, 2 ** 64, size=n, dtype=np.uint64)
        planes = np.zeros(64, dtype=np.uint64)
        L().orc_to_planes(ptr(u), n, ptr(planes))
        for k in range(64):
            want = sum(((int(u[i]) >> k) & 1) << i for i in range(n))
            assert int(planes[k]) == want


# ------------------------------------------------------------ bit stream vectors

WORD1 = (1 << 64) - 1
WORD2 = 0x5555555555555555


class S:
    """An oracle stream of three words (src/bitstream.rs:41-47)."""

    def __init__(self, words=None):
        self.o = O.OracleStream(24) if words is None else O.OracleStream(words=words)
        self.p = ctypes.byref(self.o.s)

    def wb(self, v, n):
        return L().orc_write_bits(self.p, v, n)

    def rb(self, n):
        return L().orc_read_bits(self.p, n)

    @property
    def bits(self):
        return self.o.s.bits

    @property
    def buf(self):
        return self.o.s.buffer

    @property
    def pos(self):
        return self.o.s.word_pos


def test_bitstream_vectors():
    """src/bitstream.rs:373-880, the cases with exact words and cursors."""
    s = S()  # pad completes the word (:373-383)
    s.wb(0xfff, 12)
    L().orc_pad(s.p, 52)
    assert s.pos == 1 and int(s.o.words[0]) == 0xfff
    s = S()  # write_bit at the buffer's last place (:420-430)
    s.wb(0, 63)
    s.wb(1, 1)
    assert s.pos == 1 and int(s.o.words[0]) == 1 << 63 and s.buf == 0
    s = S()  # bits written from the LSB; the rest returned (:468-477)
    assert s.wb(WORD1, 3) == WORD1 >> 3 and s.bits == 3 and s.buf == 7
    s = S()  # filling the buffer exactly (:479-493)
    s.wb(WORD1, 5)
    assert s.wb(WORD2 & 0x07ffffffffffffff, 59) == 0
    L().orc_rewind(s.p)
    assert s.rb(64) == 0x1f + 0xaaaaaaaaaaaaaaa0
    s = S()  # overflow into a new buffer (:495-511)
    s.wb(WORD1, 5)
    w = (WORD2 + 0x8000000000000000) & WORD1
    rem = s.wb(w, 63)
    assert s.bits == 4 and s.buf == (w >> 59) & 0xf and rem == w >> 63
    s = S()  # reads across two words (:566-599)
    s.wb(0, 48)
    s.wb(0xffff, 16)
    s.wb(WORD2, 64)
    L().orc_rseek(s.p, 48)
    assert s.bits == 16 and s.buf == 0xffff
    got = s.rb(61)
    assert s.bits == 19 and got == 0xffff | ((WORD2 & ((1 << 45) - 1)) << 16)
    assert s.buf == WORD2 >> 45
    s = S()  # seeks (:631-682)
    s.wb(WORD1, 64)
    s.wb(WORD2, 64)
    L().orc_wseek(s.p, 69)
    assert s.pos == 1 and s.bits == 5 and s.buf == WORD2 & 0x1f
    L().orc_rseek(s.p, 69)
    assert s.pos == 2 and s.bits == 59 and s.buf == WORD2 >> 5
    L().orc_rseek(s.p, 64)
    assert s.pos == 1 and s.bits == 0 and s.buf == 0
    L().orc_rewind(s.p)  # skips (:703-743)
    s.rb(3)
    L().orc_skip(s.p, 69)
    assert s.pos == 2 and s.bits == 64 - 72 % 64 and s.buf == WORD2 >> (72 % 64)
    s = S()  # past the end (:745-757)
    off = 5 * 64 + 3
    L().orc_rseek(s.p, off)
    assert L().orc_rtell(s.p) == off and s.rb(10) == 0
    L().orc_skip(s.p, 100)
    assert L().orc_rtell(s.p) == off + 110
    s = S()  # overread (:759-787)
    end = 3 * 64
    L().orc_rseek(s.p, end)
    L().orc_skip(s.p, 64)
    assert not L().orc_overread(s.p)
    L().orc_rseek(s.p, end + 3)
    assert L().orc_overread(s.p)
    L().orc_rseek(s.p, 0)
    assert L().orc_overread(s.p)
    s = S()
    L().orc_rseek(s.p, end - 1)
    assert s.rb(1) == 0 and not L().orc_overread(s.p)
    assert s.rb(1) == 0 and L().orc_overread(s.p)
    s = S()  # writes past the end are dropped (:807-819)
    L().orc_wseek(s.p, off)
    s.wb(0x3ff, 10)
    L().orc_flush(s.p)
    assert s.o.s.overflowed and L().orc_wtell(s.p) == 6 * 64 and not s.o.words.any()
    s = S()  # align keeps the word (:821-837)
    s.wb(WORD1, 64)
    s.wb(WORD2, 64)
    L().orc_rewind(s.p)
    s.rb(3)
    L().orc_align(s.p)
    assert s.pos == 1 and s.bits == 0 and s.buf == 0
    s = S()  # flush pads the word (:854-871)
    s.wb(WORD1, 3)
    L().orc_flush(s.p)
    assert s.pos == 1 and int(s.o.words[0]) == 7
    s = S()  # pad past the end all at once (:964-972)
    s.wb(3, 2)
    L().orc_pad(s.p, (1 << 64) - 1)
    assert s.o.s.overflowed and int(s.o.words[0]) == 3 and not s.o.words[1:].any()


@pytest.mark.parametrize("start", [0, 5, 64 + 63, 128, 3 * 64 + 1, 7 * 64 + 9])
def test_pad_equals_writing_zeros(start):
    """src/bitstream.rs:942-962."""
    for n in (0, 1, 63, 64, 65, 3 * 64, 6 * 64 + 5, 11 * 64):
        a, b = S(), S()
        for s in (a, b):
            L().orc_wseek(s.p, start)
            s.wb(0x2a5, 10)
        L().orc_pad(a.p, n)
        r = n
        while r > 0:
            k = min(r, 64)
            b.wb(0, k)
            r -= k
        for f in ("word_pos", "buffer", "bits", "overflowed"):
            assert getattr(a.o.s, f) == getattr(b.o.s, f), (start, n, f)
        assert np.array_equal(a.o.words, b.o.words)


# ------------------------------------------------------------ field-level invariants

@pytest.mark.parametrize("ztype", [O.INT32, O.INT64, O.FLOAT, O.DOUBLE])
def test_threaded_equals_serial_and_fixed_rate_length(ztype):
    """tests/pipeline.rs and src/decompress.rs:379-478: the chunked parallel
    compressor and the seek-parallel decompressor give the serial bytes and
    values; a fixed-rate stream is ceil((start + blocks * maxbits) / 64)
    words (compress.rs:44-53)."""
    rng = np.random.default_rng(ztype)
    dt = O.DTYPES[ztype]
    shape = (13, 22, 9)
    if np.dtype(dt).kind == "f":
        arr = np.cumsum(rng.standard_normal(shape), axis=2).astype(dt)
    else:
        arr = np.cumsum(rng.integers(-50, 50, size=shape), axis=2).astype(dt)
    blocks = 4 * 6 * 3
    for cfg in (O.fixed_rate(7, ztype, 3), O.fixed_precision(19), O.fixed_accuracy(1e-3),
                O.reversible()):
        for start in (0, 29):
            n1, e1, w1 = O.compress(arr, cfg, start_bits=start)
            n4, e4, w4 = O.compress(arr, cfg, threads=4, start_bits=start)
            assert e1 == e4 == 0 and n1 == n4 and np.array_equal(w1, w4)
            if cfg.minbits == cfg.maxbits:
                assert n1 == (start + blocks * cfg.maxbits + 63) // 64 * 8
            d1, m1, _ = O.decompress(w1, arr.shape, arr.dtype, cfg, start_bits=start)
            d4, m4, _ = O.decompress(w1, arr.shape, arr.dtype, cfg, threads=4, start_bits=start)
            assert m1 == m4 and np.array_equal(d1, d4)
            if cfg.minexp < -1074:
                assert np.array_equal(d1.view(np.uint8), arr.view(np.uint8))
