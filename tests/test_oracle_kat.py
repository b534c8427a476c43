"""Pin the oracle against the reference's own block-level known answers.

The checksums come from the reference's C-port suite (tests/c/checksums.rs,
tables extracted to tests/golden/checksums.json by tests/golden/make_fixtures.py).
Each test restates one of the reference's tests:
  0x0   seeded input          tests/c/encode_block.rs  when_seeded_random_data_generated_expect_checksum_matches
  0xa0  full-block stream     tests/c/encode_block.rs  given_block_when_encode_block_expect_bitstream_checksum_matches
  0x120 full-block decode     tests/c/decode_block.rs  given_block_when_decode_block_expect_array_checksum_matches
  0x2a0 partial strided enc   tests/c/encode_block_strided.rs  ..._partial_block_strided_expect_bitstream_checksum_matches
  0x320 partial strided dec   tests/c/decode_block_strided.rs  ..._partial_block_strided_expect_array_checksum_matches
  0xd1+ reversible specials   tests/c/encode_block.rs  given_block_when_encode_special_blocks_expect_bitstream_checksum_matches
"""
import numpy as np
import pytest

import oracle_py as O
from kat_util import (SPECIAL_F32, SPECIAL_F64, TNAMES, checksum, compute_key, hash_array32,
                      hash_array64, hash_bitstream, hash_strided, lcg_block, strided_array)

CASES = [(d, t) for d in (1, 2, 3, 4) for t in (1, 2, 3, 4)]
MODE_FIXED_RATE, MODE_REVERSIBLE = 2, 5


def dim_lens(d):
    return [4] * d + [0] * (4 - d)


def rate_cfg(d):
    n = 4 ** d
    return O.config(n * 19, n * 19, 64, -1074)


def hash_vals(v):
    return hash_array32(v) if v.dtype.itemsize == 4 else hash_array64(v)


@pytest.mark.parametrize("dims,ztype", CASES)
def test_seeded_input_matches(dims, ztype):
    blk = lcg_block(ztype, 4 ** dims)
    k1, k2 = compute_key(0, 0, dim_lens(dims), 0, 0)
    assert hash_vals(blk) == checksum(dims, TNAMES[ztype], k1, k2)


def encode_full(dims, ztype):
    blk = lcg_block(ztype, 4 ** dims)
    st = O.OracleStream(65536)
    bits = O.encode_block(st, rate_cfg(dims), ztype, dims, blk, 0, [1, 4, 16, 64][:dims], [4] * dims)
    assert bits == st.wtell()
    st.flush()
    return blk, st


@pytest.mark.parametrize("dims,ztype", CASES)
def test_full_block_stream_matches(dims, ztype):
    _, st = encode_full(dims, ztype)
    k1, k2 = compute_key(0, 1, dim_lens(dims), MODE_FIXED_RATE, 0)
    assert hash_bitstream(st.committed()) == checksum(dims, TNAMES[ztype], k1, k2)


@pytest.mark.parametrize("dims,ztype", CASES)
def test_full_block_decode_matches(dims, ztype):
    blk, st = encode_full(dims, ztype)
    st.rewind()
    out = np.zeros_like(blk)
    bits = O.decode_block(st, rate_cfg(dims), ztype, dims, out, 0, [1, 4, 16, 64][:dims], [4] * dims)
    assert bits == 4 ** dims * 19
    k1, k2 = compute_key(0, 2, dim_lens(dims), MODE_FIXED_RATE, 0)
    assert hash_vals(out) == checksum(dims, TNAMES[ztype], k1, k2)


PARTIAL = [1, 2, 3, 4]


@pytest.mark.parametrize("dims,ztype", CASES)
def test_partial_strided_stream_matches(dims, ztype):
    data, strides = strided_array(ztype, dims)
    st = O.OracleStream(65536)
    O.encode_block(st, rate_cfg(dims), ztype, dims, data, 0, strides, PARTIAL[:dims])
    st.flush()
    k1, k2 = compute_key(1, 1, dim_lens(dims), MODE_FIXED_RATE, 0)
    assert hash_bitstream(st.committed()) == checksum(dims, TNAMES[ztype], k1, k2)


@pytest.mark.parametrize("dims,ztype", CASES)
def test_partial_strided_decode_matches(dims, ztype):
    data, strides = strided_array(ztype, dims)
    st = O.OracleStream(65536)
    O.encode_block(st, rate_cfg(dims), ztype, dims, data, 0, strides, PARTIAL[:dims])
    st.flush()
    st.rewind()
    out = np.zeros_like(data)
    O.decode_block(st, rate_cfg(dims), ztype, dims, out, 0, strides, PARTIAL[:dims])
    k1, k2 = compute_key(1, 2, dim_lens(dims), MODE_FIXED_RATE, 0)
    assert hash_strided(out, dims, strides) == checksum(dims, TNAMES[ztype], k1, k2)


@pytest.mark.parametrize("dims,ztype", [(d, t) for d in (1, 2, 3, 4) for t in (3, 4)])
def test_reversible_special_blocks_match(dims, ztype):
    specials = SPECIAL_F32 if ztype == 3 else SPECIAL_F64
    dtype = np.float32 if ztype == 3 else np.float64
    utype = np.uint32 if ztype == 3 else np.uint64
    checked = 0
    for i, bits in enumerate(specials):
        k1, k2 = compute_key(0, 1, dim_lens(dims), MODE_REVERSIBLE, i + 1)
        want = checksum(dims, TNAMES[ztype], k1, k2)
        if want is None:
            continue
        raw = np.zeros(4 ** dims, dtype=utype)
        raw[::4] = bits
        blk = raw.view(dtype)
        st = O.OracleStream(65536)
        O.encode_block(st, O.reversible(), ztype, dims, blk, 0, [1, 4, 16, 64][:dims], [4] * dims)
        st.flush()
        assert hash_bitstream(st.committed()) == want, f"special {i}"
        # and it decodes bit for bit
        st.rewind()
        out = np.zeros_like(blk)
        O.decode_block(st, O.reversible(), ztype, dims, out, 0, [1, 4, 16, 64][:dims], [4] * dims)
        assert np.array_equal(out.view(utype), raw), f"special {i} round trip"
        checked += 1
    assert checked >= 6
