"""Host-side checks of the C ABI (no GPU needed): the library loads and exports
every symbol include/*.h declares, the bit-stream cursor behaves as the
reference's (against the oracle's restatement of src/bitstream/core.rs), and
the configuration entry points give the reference's values
(src/config.rs tests, zfp-rs-ffi/src/stream.rs).
"""
import ctypes
import pathlib
import re

import numpy as np
import pytest

import oracle_py as O
import zfp_hip as Z

ROOT = pathlib.Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, re.M):
            name = m.group(1)
            if name not in ("if", "while", "return", "sizeof") and not name.startswith("ZFP_"):
                names.add(name)
    return names


def test_every_declared_symbol_is_exported():
    L = Z.lib()
    names = declared_functions()
    assert "zfp_compress" in names and "zfp_decompress" in names and len(names) > 90
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    for sym in Z.DATA_SYMBOLS:
        assert hasattr(L, sym)
    assert ctypes.c_uint32.in_dll(L, "zfp_codec_version").value == 5
    assert ctypes.c_uint32.in_dll(L, "zfp_library_version").value == 10001
    assert ctypes.c_size_t.in_dll(L, "stream_word_bits").value == 64
    # every binding signature names a real export
    for n in Z.SIGNATURES:
        assert hasattr(L, n), n


def _oracle_ops():
    L = O.lib()
    L.orc_write_bit.argtypes = [ctypes.POINTER(O.Stream), ctypes.c_uint32]
    L.orc_write_bit.restype = ctypes.c_uint32
    L.orc_read_bit.argtypes = [ctypes.POINTER(O.Stream)]
    L.orc_read_bit.restype = ctypes.c_uint32
    return L


@pytest.mark.parametrize("seed", range(6))
def test_bitstream_ops_match_oracle(seed):
    """Random mixes of bit I/O, seeks, pads and flushes, past the end included."""
    rng = np.random.default_rng(seed)
    OL = _oracle_ops()
    ZL = Z.lib()
    nwords = int(rng.integers(1, 6))
    zb = np.zeros(nwords, dtype=np.uint64)
    ob = np.zeros(nwords, dtype=np.uint64)
    bs = ZL.stream_open(zb.ctypes.data, zb.nbytes)
    os_ = O.OracleStream(words=ob)
    so = ctypes.byref(os_.s)
    for step in range(400):
        op = int(rng.integers(0, 10))
        if op == 0:
            n = int(rng.integers(0, 70))
            v = int(rng.integers(0, 2 ** 63)) * 2 + int(rng.integers(0, 2))
            assert ZL.stream_write_bits(bs, v, n) == OL.orc_write_bits(so, v, min(n, 64)), step
        elif op == 1:
            b = int(rng.integers(0, 2))
            ZL.stream_write_bit(bs, b)
            OL.orc_write_bit(so, b)
        elif op == 2:
            n = int(rng.integers(0, 70))
            assert ZL.stream_read_bits(bs, n) == OL.orc_read_bits(so, min(n, 64)), step
        elif op == 3:
            assert ZL.stream_read_bit(bs) == OL.orc_read_bit(so)
        elif op == 4:
            off = int(rng.integers(0, 64 * (nwords + 2)))
            ZL.stream_rseek(bs, off)
            OL.orc_rseek(so, off)
        elif op == 5:
            off = int(rng.integers(0, 64 * (nwords + 2)))
            ZL.stream_wseek(bs, off)
            OL.orc_wseek(so, off)
        elif op == 6:
            n = int(rng.integers(0, 200))
            ZL.stream_pad(bs, n)
            OL.orc_pad(so, n)
        elif op == 7:
            assert ZL.stream_flush(bs) == OL.orc_flush(so)
        elif op == 8:
            assert ZL.stream_align(bs) == OL.orc_align(so)
        else:
            n = int(rng.integers(0, 150))
            ZL.stream_skip(bs, n)
            OL.orc_skip(so, n)
        assert ZL.stream_wtell(bs) == OL.orc_wtell(so), step
        assert ZL.stream_rtell(bs) == OL.orc_rtell(so), step
        assert np.array_equal(zb, ob), step
        assert ZL.stream_size(bs) == os_.s.word_pos * 8
    ZL.stream_close(bs)


def _stream():
    ZL = Z.lib()
    buf = np.zeros(64, dtype=np.uint64)
    bs = ZL.stream_open(buf.ctypes.data, buf.nbytes)
    return ZL, bs, ZL.zfp_stream_open(bs), buf


def _params(ZL, zs):
    a, b, c, d = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint(), ctypes.c_int32()
    ZL.zfp_stream_params(zs, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d))
    return a.value, b.value, c.value, d.value


def test_config_entry_points_match_reference_values():
    ZL, bs, zs, _ = _stream()
    # config.rs:770-808
    assert ZL.zfp_stream_set_rate(zs, 8.0, Z.TYPE_DOUBLE, 3, 1) == 8.0
    assert _params(ZL, zs)[1] == 512
    ZL.zfp_stream_set_rate(zs, 5.0, Z.TYPE_DOUBLE, 3, 1)
    assert _params(ZL, zs)[:2] == (320, 320)
    ZL.zfp_stream_set_rate(zs, 0.1, Z.TYPE_FLOAT, 1, 0)
    assert _params(ZL, zs)[:2] == (9, 9)
    ZL.zfp_stream_set_rate(zs, 0.1, Z.TYPE_DOUBLE, 1, 0)
    assert _params(ZL, zs)[:2] == (12, 12)
    assert ZL.zfp_stream_set_rate(zs, float("nan"), Z.TYPE_DOUBLE, 1, 0) == 0.0
    # modes and mode words (config.rs:469-600)
    assert ZL.zfp_stream_set_precision(zs, 0) == 64
    assert ZL.zfp_stream_compression_mode(zs) == 1  # expert (the defaults)
    ZL.zfp_stream_set_precision(zs, 20)
    assert ZL.zfp_stream_compression_mode(zs) == 3
    assert ZL.zfp_stream_mode(zs) == 2048 + 19
    assert ZL.zfp_stream_set_accuracy(zs, 1e-6) == 2.0 ** -20
    assert _params(ZL, zs)[3] == -20
    assert ZL.zfp_stream_compression_mode(zs) == 4
    assert ZL.zfp_stream_set_accuracy(zs, float("inf")) == 0.5
    ZL.zfp_stream_set_reversible(zs)
    assert ZL.zfp_stream_compression_mode(zs) == 5 and ZL.zfp_stream_mode(zs) == 2176
    ZL.zfp_stream_set_rate(zs, 8.0, Z.TYPE_FLOAT, 3, 0)
    assert ZL.zfp_stream_mode(zs) == 511 and ZL.zfp_stream_rate(zs, 3) == 8.0
    # expert words round-trip through set_mode
    assert ZL.zfp_stream_set_params(zs, 100, 600, 20, -30)
    m = ZL.zfp_stream_mode(zs)
    assert m > 0xffe
    ZL.zfp_stream_set_precision(zs, 7)
    assert ZL.zfp_stream_set_mode(zs, m) == 1
    assert _params(ZL, zs) == (100, 600, 20, -30)
    assert not ZL.zfp_stream_set_params(zs, 5, 4, 20, 0)
    # block maximum sizes (zfp-rs-ffi/src/stream.rs:178-206)
    assert ZL.zfp_block_maximum_size(Z.TYPE_FLOAT, 3, 0) == 2120
    assert ZL.zfp_block_maximum_size(Z.TYPE_DOUBLE, 3, 1) == 4178
    assert ZL.zfp_block_maximum_size(0, 3, 1) == 0
    # execution policies: cuda is refused, hip accepted, omp settings kept
    assert not ZL.zfp_stream_set_execution(zs, Z.EXEC_CUDA)
    assert ZL.zfp_stream_set_execution(zs, Z.EXEC_HIP) and ZL.zfp_stream_execution(zs) == Z.EXEC_HIP
    assert ZL.zfp_stream_set_omp_threads(zs, 3) and ZL.zfp_stream_omp_threads(zs) == 3
    assert ZL.zfp_stream_set_omp_chunk_size(zs, 7) and ZL.zfp_stream_omp_chunk_size(zs) == 7
    assert ZL.zfp_stream_set_rounding(zs, Z.ROUND_LAST, 1)
    tight = ctypes.c_int32()
    assert ZL.zfp_stream_rounding(zs, ctypes.byref(tight)) == Z.ROUND_LAST and tight.value == 1
    assert not ZL.zfp_stream_set_rounding(zs, 7, 0)
    ZL.zfp_stream_close(zs)
    ZL.stream_close(bs)


def test_maximum_size_matches_oracle():
    ZL, bs, zs, _ = _stream()
    for ztype, dt in ((1, np.int32), (2, np.int64), (3, np.float32), (4, np.float64)):
        for shape in ((37,), (9, 13), (5, 6, 7)):
            arr = np.zeros(shape, dtype=dt)
            f = Z.make_field(arr)
            for oc in (O.fixed_rate(8, ztype, len(shape)), O.fixed_precision(20), O.reversible(),
                       O.config(1, 11, 64, -1074)):
                ZL.zfp_stream_set_params(zs, oc.minbits, oc.maxbits, oc.maxprec, oc.minexp)
                dv = (ctypes.c_size_t * 3)(*(list(reversed(shape)) + [0] * (3 - len(shape))))
                assert ZL.zfp_stream_maximum_size(zs, ctypes.byref(f)) == \
                    O.lib().orc_maximum_size(oc, ztype, dv)
    ZL.zfp_stream_close(zs)
    ZL.stream_close(bs)


def test_header_round_trip_and_metadata():
    ZL, bs, zs, buf = _stream()
    arr = np.zeros((5, 6, 7), dtype=np.float64)
    f = Z.make_field(arr)
    meta = ZL.zfp_field_metadata(ctypes.byref(f))
    # field.rs:678-710: dims (x first) in 16-bit lanes, then dimensionality-1, then type
    want = ((((6 << 16) | (5 << 32) | 4) << 2 | 2) << 2) | 3
    # x = 7 -> 6, y = 6 -> 5, z = 5 -> 4
    want = ((((7 - 1) | ((6 - 1) << 16) | ((5 - 1) << 32)) << 2 | 2) << 2) | 3
    assert meta == want
    ZL.zfp_stream_set_accuracy(zs, 1e-3)
    assert ZL.zfp_write_header(zs, ctypes.byref(f), Z.ZFP_HEADER_FULL) == 96
    ZL.stream_flush(bs)
    ZL.zfp_stream_rewind(zs)
    g = Z.zfp_field()
    ZL.zfp_stream_set_precision(zs, 3)
    assert ZL.zfp_read_header(zs, ctypes.byref(g), Z.ZFP_HEADER_FULL) == 96
    assert (g.type, g.nx, g.ny, g.nz, g.nw) == (4, 7, 6, 5, 0)
    assert _params(ZL, zs)[3] == -10
    # field accessors
    assert ZL.zfp_field_dimensionality(ctypes.byref(f)) == 3
    assert ZL.zfp_field_blocks(ctypes.byref(f)) == 2 * 2 * 2
    dims = (ctypes.c_size_t * 4)()
    assert ZL.zfp_field_size(ctypes.byref(f), dims) == 210 and list(dims) == [7, 6, 5, 0]
    assert ZL.zfp_field_is_contiguous(ctypes.byref(f)) == 1
    ZL.zfp_stream_close(zs)
    ZL.stream_close(bs)


def test_stream_pad_and_copy_vectors():
    """src/bitstream.rs:964-972 (a huge pad returns at once, dropping the words
    past the end) and :1025-1050 (copy across words), through the C ABI."""
    ZL = Z.lib()
    word1, word2 = (1 << 64) - 1, 0x5555555555555555
    buf = np.zeros(3, dtype=np.uint64)
    bs = ZL.stream_open(buf.ctypes.data, buf.nbytes)
    ZL.stream_write_bits(bs, 3, 2)
    ZL.stream_pad(bs, (1 << 64) - 1)
    assert int(buf[0]) == 3 and not buf[1:].any()
    ZL.stream_close(bs)
    sbuf = np.zeros(3, dtype=np.uint64)
    dbuf = np.zeros(3, dtype=np.uint64)
    src = ZL.stream_open(sbuf.ctypes.data, sbuf.nbytes)
    dst = ZL.stream_open(dbuf.ctypes.data, dbuf.nbytes)
    ZL.stream_write_bits(src, word1, 64)
    ZL.stream_write_bits(src, word2, 64)
    ZL.stream_flush(src)
    ZL.stream_rseek(src, 58)
    ZL.stream_wseek(dst, 5)
    ZL.stream_copy(dst, src, 68)
    w2_bits = 5 + (64 - 58)
    assert int(dbuf[0]) == (((word1 >> 58) << 5) + (word2 << w2_bits)) & word1
    assert ZL.stream_wtell(dst) == 5 + 68
    ZL.stream_flush(dst)
    assert int(dbuf[1]) == (word2 >> w2_bits) & ((1 << ((5 + 68) % 64)) - 1)
    ZL.stream_close(src)
    ZL.stream_close(dst)
