"""Closed-form pins of the oracle's fixed-precision, fixed-accuracy and
rounding paths, derived by hand from the block grammar and checked without the
oracle's own formulas.

The blocks are chosen so that their streams can be written down directly:

- A constant integer block lifts to a lone DC coefficient (the forward lift of
  (c, c, c, c) is (c, 0, 0, 0): src/codec/transform.rs:74-86, the matrix of
  :67-72), so each bit plane holds one bit `a` at position 0.  In the
  embedded coder (src/codec/bitplane.rs:539-610) a plane of such a block is:
  before any coefficient is significant, "0" (one negative group test) when
  a = 0 and "1","1","0" (positive test, the one-bit at position 0, negative
  test) when a = 1; afterwards, the verbatim bit a and one negative test "0".
- Under ROUND_FIRST every coefficient of that block moves by the same bias
  (src/codec/encode/core.rs:53-74: +-((NBMASK >> 2) >> maxprec), added when
  maxprec is odd), so each plane holds a bit `a` at position 0 and one bit `r`
  shared by every other position.  Such a plane codes as "1","0","1","1"*(N-3),
  "1" or "1","1"*(N-1)+"1" when r = 1 before significance, and afterwards as
  a then "1"*(2N-3) (r = 1, group tests walking every position) or a, "0".
- Under ROUND_LAST (src/codec/bitplane.rs:661-667, src/codec/decode/core.rs:62-82)
  the decoder adds (NBMASK >> 2) >> prec to the first m coefficients and
  (NBMASK >> 1) >> prec to the rest.  An unconstrained decode (fixed precision
  p) ends with m = 0 and prec = p + 1; a decode cut by its budget ends with the
  m and prec of the plane it stopped in, which the streams below are built to
  control (hand traces in each test).

The decoded integers are mapped back to coefficients with the forward
transform, which inverts the inverse lift exactly on integers (checked here
too), so every coefficient, bias included, is compared.  Negabinary is
computed with the textbook base(-2) digit recurrence, not the mask identity
the codec uses.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_py as O

L = O.lib

NB32, NB64 = 0xAAAAAAAA, 0xAAAAAAAAAAAAAAAA


# ------------------------------------------------------------ independent helpers

def negabinary(v, nbits):
    """Digits of v in base -2 (least significant first) as an nbits-bit integer."""
    u, i = 0, 0
    while v != 0:
        r = v & 1
        v = (v - r) // -2
        u |= r << i
        i += 1
    assert i <= nbits, "value needs more base(-2) digits than the coefficient holds"
    return u


def from_negabinary(u, nbits):
    """Sum of bit_i * (-2)^i, wrapped to a signed nbits-bit integer."""
    v = sum(((u >> i) & 1) * (-2) ** i for i in range(nbits))
    v %= 1 << nbits
    return v - (1 << nbits) if v >> (nbits - 1) else v


def fwd_lift_line(x, y, z, w):
    """The forward lift as src/codec/transform.rs:74-86 lists its steps (Python ints)."""
    x = (x + w) >> 1; w -= x
    z = (z + y) >> 1; y -= z
    x = (x + z) >> 1; z -= x
    w = (w + y) >> 1; y -= w
    w += y >> 1; y -= w >> 1
    return x, y, z, w


def inv_lift_line(x, y, z, w):
    """The inverse lift as src/codec/transform.rs:98-110 lists its steps."""
    y += w >> 1; w -= y >> 1
    y += w; w -= y - w
    z += x; x -= z - x
    y += z; z -= y - z
    w += x; x -= w - x
    return x, y, z, w


def words_of(bits):
    """Pack a list of 0/1 (first written first) into 64-bit words, LSB first."""
    nw = max(1, (len(bits) + 63) // 64)
    w = [0] * nw
    for i, b in enumerate(bits):
        w[i // 64] |= b << (i % 64)
    return np.array(w, dtype=np.uint64)


def s(text):
    return [int(ch) for ch in text]


def dc_plane_codes(u0, intprec, kmin):
    """Stream of a block whose only nonzero coefficient is u0 (negabinary), planes
    intprec-1 down to kmin."""
    out, sig = [], False
    for k in range(intprec - 1, kmin - 1, -1):
        a = (u0 >> k) & 1
        if not sig:
            out += s("110") if a else s("0")
            sig = bool(a)
        else:
            out += [a, 0]
    return out


def two_row_plane_codes(u0, ur, intprec, kmin, size):
    """Stream of a block whose coefficient 0 is u0 and whose other coefficients all
    equal ur (negabinary)."""
    out, n = [], 0
    for k in range(intprec - 1, kmin - 1, -1):
        a, r = (u0 >> k) & 1, (ur >> k) & 1
        if n == 0:
            if a and r:
                out += s("11") + s("11") * (size - 2) + s("1"); n = size
            elif a:
                out += s("110"); n = 1
            elif r:
                out += s("101") + s("11") * (size - 3) + s("1"); n = size
            else:
                out += s("0")
        elif n == 1:
            out += [a]
            if r:
                out += s("11") * (size - 2) + s("1"); n = size
            else:
                out += s("0")
        else:
            out += [a] + [r] * (size - 1)
    return out


def peel(block, dims, nbits):
    """Coefficients (natural order) of a decoded integer block: the forward
    transform applied to it."""
    b = np.array(block, dtype=np.int32 if nbits == 32 else np.int64).reshape(-1).copy()
    fn = L().orc_xform_i32 if nbits == 32 else L().orc_xform_i64
    fn(b.ctypes.data_as(ctypes.c_void_p), b.size, 0)
    return [int(v) for v in b]


def peel_is_exact(coeffs, dims, nbits):
    """The peel is exact while the inverse transform does not wrap: each axis's
    lift grows the largest magnitude less than fourfold (the inverse matrix's
    rows sum to at most 15/4 in magnitude, and no lifting step goes further)."""
    return max(abs(c) for c in coeffs) * 4 ** dims < 2 ** (nbits - 1)


def decode(words, shape, dtype, cfg):
    out, n, err = O.decompress(words, shape, dtype, cfg)
    assert err == 0
    return out, n


def cfg_with(base, rounding=O.ROUND_NEVER, tight=False):
    return O.config(base.minbits, base.maxbits, base.maxprec, base.minexp, rounding, tight)


INT_CASES = [(np.int32, 32), (np.int64, 64)]


def test_forward_lift_inverts_the_inverse_lift():
    """The peel used below is exact: fwd(inv(y)) == y for every integer line
    (inv(fwd(x)) is not, the forward lift drops low bits)."""
    rng = np.random.default_rng(5)
    for _ in range(20000):
        y = tuple(int(v) for v in rng.integers(-2 ** 40, 2 ** 40, size=4))
        assert fwd_lift_line(*inv_lift_line(*y)) == y
    assert fwd_lift_line(7, 7, 7, 7) == (7, 0, 0, 0)
    assert inv_lift_line(9, 0, 0, 0) == (9, 9, 9, 9)
    # and the oracle's transform agrees with the listed steps on a line
    for _ in range(200):
        v = [int(x) for x in rng.integers(-2 ** 29, 2 ** 29, size=4)]
        b = np.array(v, dtype=np.int32)
        L().orc_xform_i32(b.ctypes.data_as(ctypes.c_void_p), 4, 1)
        assert [int(x) for x in b] == list(inv_lift_line(*v))


# ------------------------------------------------------------ fixed precision, DC-only blocks

@pytest.mark.parametrize("dtype,P", INT_CASES)
@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("p", [1, 2, 5, 12, 20, 31, 33, 47, 64])
@pytest.mark.parametrize("c", [0, 1, -1, 0x12345, -0x2345678])
def test_fixed_precision_dc_block_stream(dtype, P, dims, p, c):
    """A constant integer block at fixed precision p codes planes P-1 .. P-p
    (kmin = P - p, src/codec/bitplane.rs:519) with the DC-only plane codes,
    pads to one bit (minbits) and flushes; decoding gives the constant block of
    the DC's top p negabinary digits (inv_lift(v, 0, 0, 0) = (v, v, v, v))."""
    block = np.full((4,) * dims, c, dtype=dtype)
    cfg = O.fixed_precision(p)
    n, err, words = O.compress(block, cfg)
    assert err == 0
    kmin = max(0, P - p)
    u0 = negabinary(c, P)
    bits = dc_plane_codes(u0, P, kmin) or [0]
    want = words_of(bits)
    assert n == len(want) * 8
    assert np.array_equal(words[: len(want)], want)
    out, m = decode(words, block.shape, dtype, cfg)
    assert m == n
    v = from_negabinary(u0 & ~((1 << kmin) - 1), P)
    assert np.all(out == v)


@pytest.mark.parametrize("dtype,P", INT_CASES)
@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("p", [1, 4, 12, 20, 29, 30, 31, 40, 61, 62, 63])
@pytest.mark.parametrize("c", [0, 5, -0x2345678])
def test_round_last_unconstrained_bias(dtype, P, dims, p, c):
    """ROUND_LAST after an unconstrained decode: m = 0 and prec = P - kmin + 1
    (C's decode_ints_prec exits with k = kmin - 1), i.e. p + 1 for p <= P, so
    every coefficient gains (NBMASK >> 1) >> prec, unless prec >= P - 1.  The
    zero block is rounded too: its inverse transform is not skipped."""
    block = np.full((4,) * dims, c, dtype=dtype)
    base = O.fixed_precision(p)
    n, err, words = O.compress(block, base)
    assert err == 0
    out, m = decode(words, block.shape, dtype, cfg_with(base, O.ROUND_LAST))
    assert m == n
    kmin = max(0, P - p)
    prec = P - kmin + 1
    mask = (1 << P) - 1
    u = [negabinary(c, P) & ~((1 << kmin) - 1)] + [0] * (4 ** dims - 1)
    nb = NB32 if P == 32 else NB64
    if prec < P - 1:
        lo = (nb >> 1) >> prec
        u = [(x + lo) & mask for x in u]
    want = [from_negabinary(x, P) for x in u]
    if not peel_is_exact(want, dims, P):
        pytest.skip("the bias is large enough for the inverse transform to wrap")
    coeffs = peel(out, dims, P)
    # every coefficient but the DC is the same, so the sequency order does not matter
    assert coeffs[0] == want[0]
    assert coeffs[1:] == want[1:]


@pytest.mark.parametrize("dtype,P", INT_CASES)
@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("p", [1, 2, 3, 8, 13, 20, 29, 31, 32, 40, 63])
@pytest.mark.parametrize("c", [0, 3, -0x123456])
def test_round_first_constant_block_stream(dtype, P, dims, p, c):
    """ROUND_FIRST: the transform runs first, then every coefficient moves by
    the bias (+ when p is odd, - when even), then the order and negabinary
    (src/codec/encode/integer.rs:50-69).  The stream is the two-row plane code
    of (c +- b, +-b, ..., +-b); decoding it with Never gives those coefficients
    cut to p planes."""
    size = 4 ** dims
    block = np.full((4,) * dims, c, dtype=dtype)
    base = O.fixed_precision(p)
    n, err, words = O.compress(block, cfg_with(base, O.ROUND_FIRST))
    assert err == 0
    kmin = max(0, P - p)
    nb = NB32 if P == 32 else NB64
    b = ((nb >> 2) >> p) if p < P else 0
    sgn = 1 if p % 2 else -1
    u0 = negabinary(c + sgn * b, P)
    ur = negabinary(sgn * b, P)
    bits = two_row_plane_codes(u0, ur, P, kmin, size) or [0]
    want = words_of(bits)
    assert n == len(want) * 8
    assert np.array_equal(words[: len(want)], want)
    out, m = decode(words, block.shape, dtype, base)
    assert m == n
    top = ~((1 << kmin) - 1)
    want = [from_negabinary(u0 & top, P)] + [from_negabinary(ur & top, P)] * (size - 1)
    if peel_is_exact(want, dims, P):
        assert peel(out, dims, P) == want


def _block_with_coefficients(dc, ac, dims, dtype):
    """An integer block whose forward transform is (dc, ac, ac, ..., ac): the
    inverse transform of that vector (exact, since fwd(inv(y)) == y)."""
    b = np.array([dc] + [ac] * (4 ** dims - 1), dtype=dtype)
    fn = L().orc_xform_i32 if dtype == np.int32 else L().orc_xform_i64
    fn(b.ctypes.data_as(ctypes.c_void_p), b.size, 1)
    return b.reshape((4,) * dims)


@pytest.mark.parametrize("dtype,P", INT_CASES)
@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("p", [11, 14, 20, 27])
@pytest.mark.parametrize("edge", ["carry", "brink"])
def test_round_first_biases_the_transformed_coefficients(dtype, P, dims, p, edge):
    """ROUND_FIRST adds the bias to the transform's output, not to its input.

    A block whose coefficients are (d, q, ..., q) codes as the two-row planes of
    nb(d +- b) and nb(q +- b).  The values k negabinary digits can hold run from
    the all-odd-digits pattern (the least) to the all-even-digits one (the
    greatest), so a q whose low kmin digits are the greatest pattern carries
    into plane kmin when b > 0 is added, and one whose digits are the least
    pattern borrows from it when b is subtracted: the bias then shows in the
    coded planes of every AC coefficient, which biasing the values before the
    transform would leave alone.  The "brink" q sits exactly b inside that end,
    so the bias just fails to carry and a bias of any other size would."""
    size = 4 ** dims
    kmin = P - p
    nb = NB32 if P == 32 else NB64
    b = (nb >> 2) >> p
    sgn = 1 if p % 2 else -1
    low = (0x5555555555555555 if sgn > 0 else 0xAAAAAAAAAAAAAAAA) & ((1 << kmin) - 1)
    d = from_negabinary(1 << kmin, P)
    q = d + from_negabinary(low, P)
    top = ~((1 << kmin) - 1)
    if edge == "brink":
        q -= sgn * b
        assert not (negabinary(q + sgn * b, P) ^ negabinary(q, P)) & top
        assert (negabinary(q + sgn * (b + 1), P) ^ negabinary(q, P)) & top
        assert (negabinary(q + 2 * sgn * b, P) ^ negabinary(q, P)) & top
    else:
        assert (negabinary(q + sgn * b, P) ^ negabinary(q, P)) & top
    block = _block_with_coefficients(d, q, dims, dtype)
    assert max(abs(int(v)) for v in block.reshape(-1)) < 2 ** (P - 2)
    assert peel(block, dims, P) == [d] + [q] * (size - 1)
    base = O.fixed_precision(p)
    n, err, words = O.compress(block, cfg_with(base, O.ROUND_FIRST))
    assert err == 0
    bits = two_row_plane_codes(negabinary(d + sgn * b, P), negabinary(q + sgn * b, P), P, kmin,
                               size)
    want = words_of(bits)
    assert n == len(want) * 8
    assert np.array_equal(words[: len(want)], want)
    if edge == "carry":
        # without the bias the AC planes differ
        n0, _, w0 = O.compress(block, base)
        assert w0[: n0 // 8].tobytes() != words[: n // 8].tobytes()


def test_round_first_fixed_rate_leaves_full_precision_streams_alone():
    """Fixed-rate configs have maxprec 64, so ROUND_FIRST never biases them
    (fwd_round only acts below the integer width): the stream equals Never's."""
    rng = np.random.default_rng(9)
    data = rng.standard_normal((8, 8, 8)).astype(np.float32)
    for rate in (4, 8, 16):
        base = O.fixed_rate(rate, O.FLOAT, 3)
        n0, _, w0 = O.compress(data, base)
        for tight in (False, True):
            n1, _, w1 = O.compress(data, cfg_with(base, O.ROUND_FIRST, tight))
            assert n1 == n0 and np.array_equal(w0[: n0 // 8], w1[: n1 // 8])


# ------------------------------------------------------------ ROUND_LAST under a binding budget

def _expert(bits, P):
    return O.config(bits, bits, 64, -1074)


@pytest.mark.parametrize("dtype,P", INT_CASES)
def test_round_last_budget_cut_states(dtype, P):
    """Hand traces of decode_planes (src/codec/bitplane.rs:675-770) on 1-D
    integer blocks (PERM_1 is the identity), budget = the block's bits, each
    stream after z = 6 zeros (six empty planes, one negative test each, which
    add 6 to prec):

    S1 = 110 10 00 1, 8 bits: plane P-7 takes 110 (coefficient 0 significant),
      the next two planes take a verbatim bit and a negative test, the fourth
      only its verbatim bit before the budget ends.  m = 1, prec = 10.
    S2 = 11110 ab 1, 8 bits: the first coded plane makes coefficients 0 and 1
      significant; the next takes two verbatim bits and a positive test whose search the
      budget cuts, so the one-bit is implied at position 2.  m = 3, prec = 8.
    S3 = 11110 ab, 7 bits: the budget ends after the verbatim bits.  m = 2,
      prec = 8.
    Z = twenty zeros (no prefix): a run of empty planes ends the budget with no
      coefficient significant.  m = 4 (the whole block), prec = 20.
    Each coefficient is checked with and without ROUND_LAST."""
    nb = NB32 if P == 32 else NB64
    mask = (1 << P) - 1
    z = 6                      # leading empty planes keep the values from wrapping
    t = P - 1 - z
    Z = [0] * z
    cases = []
    # (bits, maxbits, u (negabinary per coefficient), m, prec)
    cases.append((Z + s("11010001"), z + 8, [(1 << t) | (1 << (t - 1)) | (1 << (t - 3)), 0, 0, 0],
                  1, z + 4))
    for a, b in ((0, 0), (1, 0), (0, 1), (1, 1)):
        u = [(1 << t) | (a << (t - 1)), (1 << t) | (b << (t - 1)), 1 << (t - 1), 0]
        cases.append((Z + s("11110") + [a, b] + [1], z + 8, u, 3, z + 2))
        u3 = [(1 << t) | (a << (t - 1)), (1 << t) | (b << (t - 1)), 0, 0]
        cases.append((Z + s("11110") + [a, b], z + 7, u3, 2, z + 2))
    cases.append(([0] * 20, 20, [0, 0, 0, 0], 4, 20))
    for bits, maxbits, u, m, prec in cases:
        words = words_of(bits)
        cfg = _expert(maxbits, P)
        out, _ = decode(words, (4,), dtype, cfg)
        # Never: the coefficients exactly as coded (the zero block skips the transform)
        assert peel(out, 1, P) == [from_negabinary(x, P) for x in u], bits
        out, _ = decode(words, (4,), dtype, cfg_with(cfg, O.ROUND_LAST))
        hi, lo = (nb >> 2) >> prec, (nb >> 1) >> prec
        want = [(x + (hi if i < m else lo)) & mask for i, x in enumerate(u)]
        assert peel(out, 1, P) == [from_negabinary(x, P) for x in want], bits


@pytest.mark.parametrize("dtype,P", INT_CASES)
@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("maxbits", [5, 20, 31])
def test_round_last_zero_block_under_a_fixed_rate_budget(dtype, P, dims, maxbits):
    """A zero block coded at a rate below one bit per plane leaves `maxbits`
    empty planes; the decoder stops there with m = size and prec = maxbits, so
    every coefficient gains (NBMASK >> 2) >> maxbits (none when maxbits >= P - 1)."""
    size = 4 ** dims
    block = np.zeros((4,) * dims, dtype=dtype)
    cfg = _expert(maxbits, P)
    n, err, words = O.compress(block, cfg)
    assert err == 0 and n == 8 and int(words[0]) == 0
    out, _ = decode(words, block.shape, dtype, cfg_with(cfg, O.ROUND_LAST))
    nb = NB32 if P == 32 else NB64
    hi = (nb >> 2) >> maxbits if maxbits < P - 1 else 0
    assert peel(out, dims, P) == [from_negabinary(hi, P)] * size


# ------------------------------------------------------------ float blocks: header + DC planes

def _float_dc_case(dtype, c):
    """(EBITS, EBIAS, P, emax, integer DC) of a constant float block c != 0: the
    exponent is frexp's, the cast scales by 2^(P-2-emax) exactly."""
    if dtype == np.float32:
        ebits, ebias, P = 8, 127, 32
    else:
        ebits, ebias, P = 11, 1023, 64
    emax = math.frexp(c)[1]
    dc = int(c * 2.0 ** (P - 2 - emax))
    assert dc == c * 2.0 ** (P - 2 - emax)
    return ebits, ebias, P, emax, dc


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("c", [1.0, -3.0, 0.15625, 1.5e3])
@pytest.mark.parametrize("p", [1, 7, 16, 32, 50])
def test_fixed_precision_float_dc_block(dtype, dims, c, p):
    """Constant float block at fixed precision p (src/codec/encode/float.rs:22-113):
    a one-bit, the biased exponent in EBITS bits, then the DC plane codes of
    min(p, P) planes (precision_f is clamped by maxprec when minexp is the
    default -1074)."""
    ebits, ebias, P, emax, dc = _float_dc_case(dtype, c)
    block = np.full((4,) * dims, c, dtype=dtype)
    cfg = O.fixed_precision(p)
    n, err, words = O.compress(block, cfg)
    assert err == 0
    e = emax + ebias
    header = [1] + [(e >> i) & 1 for i in range(ebits)]
    prec = min(p, P)
    bits = header + dc_plane_codes(negabinary(dc, P), P, P - prec)
    want = words_of(bits)
    assert n == len(want) * 8
    assert np.array_equal(words[: len(want)], want)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("c", [1.0, -3.0, 1.5e3])
@pytest.mark.parametrize("tol_exp", [-30, -12, -4, 0, 3, 10, 12])
def test_fixed_accuracy_float_dc_block(dtype, dims, c, tol_exp):
    """Fixed accuracy 2^t: minexp = t (frexp(2^t).e - 1) and the block codes
    emax - minexp + 2*dims + 2 planes (clamped to [0, 64] and to P by the
    planes there are); with none, the block is a single zero bit.  Pins kmin
    for the accuracy mode."""
    ebits, ebias, P, emax, dc = _float_dc_case(dtype, c)
    block = np.full((4,) * dims, c, dtype=dtype)
    cfg = O.fixed_accuracy(math.ldexp(1.0, tol_exp))
    assert cfg.minexp == tol_exp
    n, err, words = O.compress(block, cfg)
    assert err == 0
    prec = max(0, min(64, emax - tol_exp + 2 * dims + 2))
    if prec == 0:
        bits = [0]
    else:
        e = emax + ebias
        bits = [1] + [(e >> i) & 1 for i in range(ebits)]
        bits += dc_plane_codes(negabinary(dc, P), P, max(0, P - prec))
    want = words_of(bits)
    assert n == len(want) * 8
    assert np.array_equal(words[: len(want)], want)
    out, _ = decode(words, block.shape, dtype, cfg)
    assert np.max(np.abs(out.astype(np.float64) - c)) <= math.ldexp(1.0, tol_exp)
