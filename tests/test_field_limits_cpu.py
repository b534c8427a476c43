"""Host-side checks of field planning (no GPU): aliasing strides, shapes whose
block count overflows, and fields past the codec's block limit.

- strides_may_alias: the reference's own vectors (src/field_plan.rs:292-326)
  through zfp_hip_field_may_alias.
- The alias scatter (zfpx_alias.h, built for the CPU in tests/native): blocks
  decoded into a dense copy, then stored by ticket, equal the oracle's serial
  decode into the aliasing strides, which is what the reference's parallel
  decoder falls back to (src/decompress.rs:157-163, 380-400).
- A 3-D field of 2^24 per axis overflows the block count: the reference
  reports num_blocks as usize::MAX and zfp_compress fails (ShapeTooLarge,
  src/field.rs:84-106, 559-569) without writing.
- A 1-D f32 field of 2^34 + 4 values is legal: the field queries report it,
  while this library's compress and decompress refuse it (its kernels index
  at most 2^32 - 1 blocks) and leave the stream untouched.
"""
import ctypes

import numpy as np
import pytest

import hc_util as HC
import oracle_py as O
import zfp_hip as Z

L = Z.lib

SIZE_MAX = 2 ** 64 - 1


def field(ztype, dims, strides, data=None):
    n = list(dims) + [0] * (4 - len(dims))
    f = L().zfp_field_4d(data, ztype, *n) if len(dims) == 4 else None
    if f is None:
        mk = {1: L().zfp_field_1d, 2: L().zfp_field_2d, 3: L().zfp_field_3d}[len(dims)]
        f = mk(data, ztype, *dims)
    s = list(strides) + [0] * (4 - len(strides))
    L().zfp_field_set_stride_4d(f, *s)
    return f


# ------------------------------------------------------------ strides_may_alias

@pytest.mark.parametrize("dims,strides,want", [
    ((8, 8), (1, 8), False),
    ((8, 8), (-1, 8), False),
    ((8, 8), (8, 1), False),
    ((8,), (1,), False),
    ((8, 8), (1, 1), True),     # both axes step by one element
    ((8, 8), (1, 4), True),     # a stride inside the span of a smaller one
    ((8, 8), (3, 0), True),     # zero selects the natural stride, 8 < 3*7 + 1
    ((8, 8, 4), (1, 8, 64), False),
    ((8, 8, 4), (1, 8, 63), True),
    ((8, 8, 4), (-64, 1, -8), False),  # sorted by |stride|: 1, 8, 64 each clear
    ((8, 8, 4), (-64, 1, -7), True),   # 7 < 1*7 + 1
    ((8, 8, 4), (-1, -8, -64), False),
    ((1, 8), (5, 5), False),    # an axis of one value never decides
    ((8, 1, 8), (1, 0, 8), False),
])
def test_strides_may_alias_matches_the_reference(dims, strides, want):
    f = field(Z.TYPE_FLOAT, dims, strides, ctypes.c_void_p(4096))
    try:
        assert bool(L().zfp_hip_field_may_alias(f)) == want
    finally:
        L().zfp_field_free(f)


def test_explicit_zero_stride_is_natural_not_aliasing():
    """A zero stride selects the natural stride (field.rs:576-585), so it cannot
    repeat an element; a 1-D field never aliases."""
    f = field(Z.TYPE_DOUBLE, (8, 8), (1, 0), ctypes.c_void_p(4096))
    assert not L().zfp_hip_field_may_alias(f)
    L().zfp_field_free(f)


# ------------------------------------------------------------ the alias scatter

def oracle_decode_into(words, ztype, dims, strides, buf, origin, cfg):
    """Oracle serial decode into `buf` with element [0,0,0] at index `origin`."""
    dt = O.DTYPES[ztype]
    st = O.OracleStream(words=np.ascontiguousarray(words, dtype=np.uint64).copy())
    n = list(dims) + [0] * (3 - len(dims))
    s = list(strides) + [0] * (3 - len(strides))
    f = O.Field(ztype, n[0], n[1], n[2], s[0], s[1], s[2],
                buf.ctypes.data + origin * np.dtype(dt).itemsize)
    err = ctypes.c_int(0)
    O.lib().orc_decompress(ctypes.byref(st.s), ctypes.byref(cfg), ctypes.byref(f), ctypes.byref(err))
    assert err.value == 0


def span(dims, strides):
    lo = sum(s * (n - 1) for n, s in zip(dims, strides) if s < 0)
    hi = sum(s * (n - 1) for n, s in zip(dims, strides) if s > 0)
    return lo, hi


ALIAS_CASES = [
    # (ztype, dims (x first), strides)
    (O.DOUBLE, (8, 8), (1, 1)),          # src/decompress.rs:366-400
    (O.FLOAT, (8, 8), (1, 4)),
    (O.INT32, (9, 7), (2, 3)),
    (O.INT64, (13,), (1,)),               # no alias: the scatter is a plain copy
    (O.FLOAT, (6, 5, 7), (1, 2, 3)),
    (O.DOUBLE, (5, 9, 6), (-1, 3, -4)),
    (O.FLOAT, (10, 10), (-2, -3)),
    (O.INT32, (4, 4, 4), (16, 4, 1)),      # permuted, no alias
    (O.DOUBLE, (7, 6, 5), (5, -7, 11)),
]


@pytest.mark.parametrize("ztype,dims,strides", ALIAS_CASES)
def test_alias_scatter_equals_the_serial_decode(ztype, dims, strides):
    dt = O.DTYPES[ztype]
    rng = np.random.default_rng(sum(dims) * 7 + ztype)
    shape = tuple(reversed(dims))
    src = (rng.standard_normal(shape) * 100).astype(dt) if np.dtype(dt).kind == "f" \
        else rng.integers(-1000, 1000, size=shape).astype(dt)
    for cfg in (O.fixed_rate(16, ztype, len(dims)), O.fixed_precision(12), O.reversible()):
        n, err, words = O.compress(src, cfg)
        assert err == 0
        dense, _, derr = O.decompress(words, shape, dt, cfg)
        assert derr == 0
        lo, hi = span(dims, strides)
        init = (np.arange(hi - lo + 1) * 3 + 1).astype(dt)   # untouched elements keep these
        want = init.copy()
        oracle_decode_into(words, ztype, dims, strides, want, -lo, cfg)
        got = init.copy()
        nd = (ctypes.c_int64 * 4)(*(list(dims) + [1] * (4 - len(dims))))
        sd = (ctypes.c_int64 * 4)(*(list(strides) + [0] * (4 - len(strides))))
        rc = HC.lib().hc_alias_scatter(dense.ctypes.data, got.ctypes.data, nd, sd, lo, len(dims),
                                       np.dtype(dt).itemsize)
        assert rc == 0
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), cfg


# ------------------------------------------------------------ block-count overflow and limits

def test_block_count_overflow_is_refused_without_writing():
    f = field(Z.TYPE_FLOAT, (2 ** 24, 2 ** 24, 2 ** 24), (1, 1, 1), ctypes.c_void_p(4096))
    try:
        assert L().zfp_field_blocks(f) == SIZE_MAX
        assert L().zfp_field_size(f, None) == SIZE_MAX
        assert L().zfp_field_dimensionality(f) == 3
        buf = np.zeros(64, dtype=np.uint64)
        s = Z.Stream(buf, Z.ZfpConfig.fixed_rate(8, Z.TYPE_FLOAT, 3))
        assert L().zfp_compress(s.zs, f) == 0
        assert s.wtell() == 0 and not buf.any()
        assert L().zfp_decompress(s.zs, f) == 0
        assert s.rtell() == 0
        s.close()
    finally:
        L().zfp_field_free(f)


def test_fields_past_the_codec_block_limit():
    n = 2 ** 34 + 4
    f = field(Z.TYPE_FLOAT, (n,), (0,), ctypes.c_void_p(4096))
    try:
        assert L().zfp_field_size(f, None) == n
        assert L().zfp_field_blocks(f) == 2 ** 32 + 1
        assert L().zfp_field_size_bytes(f) == 4 * n
        assert L().zfp_field_dimensionality(f) == 1
        buf = np.zeros(64, dtype=np.uint64)
        s = Z.Stream(buf, Z.ZfpConfig.fixed_rate(8, Z.TYPE_FLOAT, 1))
        assert L().zfp_write_header(s.zs, f, 0x7) > 0   # headers do not code blocks
        pos = s.wtell()
        assert L().zfp_compress(s.zs, f) == 0
        assert s.wtell() == pos
        s.close()
    finally:
        L().zfp_field_free(f)
    # one block under the limit still plans (nothing is coded here)
    f = field(Z.TYPE_FLOAT, (4 * (2 ** 32 - 1),), (0,), ctypes.c_void_p(4096))
    assert L().zfp_field_blocks(f) == 2 ** 32 - 1
    L().zfp_field_free(f)
