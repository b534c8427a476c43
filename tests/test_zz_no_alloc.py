"""zfp_hip_compress_noalloc against zfp_compress, through the C ABI.

Two calls in a row, the second with a smaller field, must give the same stream
words, sizes and cursors as zfp_compress (the static scratch keeps the first
call's contents); streams decode to the oracle's values; inputs the entry point
does not take (too many blocks, host data) return 0 and leave the stream as it
was; a stream's offset index is refreshed where it has room and marked stale
where it has not.
"""
import ctypes

import numpy as np
import pytest

import oracle_py as O
import zfp_hip as Z
from synth import grf

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

L = Z.lib


def compress_both(arr, cfg, words_cap):
    """(zfp_compress words/size/wtell, noalloc words/size/wtell) on fresh
    device streams, 13 bits into a word."""
    out = []
    for fn in (L().zfp_compress, L().zfp_hip_compress_noalloc):
        buf = torch.zeros(words_cap, dtype=torch.int64, device="cuda")
        st = Z.Stream(buf, cfg)
        st.write_bits(0x1abc, 13)
        dev = torch.from_numpy(np.ascontiguousarray(arr)).cuda()
        f = Z.make_field(dev)
        n = fn(st.zs, ctypes.byref(f))
        pos = st.wtell()
        st.close()
        torch.cuda.synchronize()
        out.append((buf.cpu().numpy().view(np.uint64).copy(), n, pos))
    return out


CASES = [
    ("precision 16", lambda: Z.ZfpConfig.fixed_precision(16), O.fixed_precision(16)),
    ("accuracy 1e-3", lambda: Z.ZfpConfig.fixed_accuracy(1e-3), O.fixed_accuracy(1e-3)),
    ("reversible", lambda: Z.ZfpConfig.reversible(), O.reversible()),
    ("rate 8", lambda: Z.ZfpConfig.fixed_rate(8, Z.TYPE_FLOAT, 3), O.fixed_rate(8, O.FLOAT, 3)),
]


@pytest.mark.parametrize("name,zc,oc", CASES)
def test_noalloc_equals_compress_on_two_calls(name, zc, oc):
    first = grf((96, 96, 96), seed=21)           # 13,824 blocks
    second = grf((37, 41, 29), seed=22)          # fewer blocks, partial ones
    for arr in (first, second):
        cap = (Z.maximum_size(zc(), arr) + 64) // 8
        (w0, n0, p0), (w1, n1, p1) = compress_both(arr, zc(), cap)
        assert n1 == n0 and p1 == p0, name
        assert n0 > 0
        assert np.array_equal(w1[: n0 // 8], w0[: n0 // 8]), name
        # the stream is the oracle's
        on, oerr, ow = O.compress(arr, oc, start_bits=13)
        assert oerr == 0 and on == n0
        assert np.array_equal(w1[1: n0 // 8], ow[1: n0 // 8])


def test_noalloc_refusals_leave_the_stream():
    cfg = Z.ZfpConfig.fixed_precision(12)
    maxb = L().zfp_hip_noalloc_max_blocks()
    assert maxb == 1048576
    big = torch.zeros(4 * (maxb + 1), dtype=torch.float32, device="cuda")
    buf = torch.zeros(1024, dtype=torch.int64, device="cuda")
    st = Z.Stream(buf, cfg)
    st.write_bits(5, 3)
    f = Z.make_field(big)
    assert L().zfp_hip_compress_noalloc(st.zs, ctypes.byref(f)) == 0
    assert st.wtell() == 3
    host = np.ones((8, 8), dtype=np.float32)
    fh = Z.make_field(host)
    assert L().zfp_hip_compress_noalloc(st.zs, ctypes.byref(fh)) == 0
    assert st.wtell() == 3
    hbuf = np.zeros(64, dtype=np.uint64)
    hs = Z.Stream(hbuf, cfg)
    fd = Z.make_field(torch.ones((8, 8), dtype=torch.float32, device="cuda"))
    assert L().zfp_hip_compress_noalloc(hs.zs, ctypes.byref(fd)) == 0
    hs.close()
    # the 3 bits written are still in the cursor's buffer: nothing reached memory
    L().zfp_stream_flush(st.zs)
    st.close()
    torch.cuda.synchronize()
    assert int(buf.cpu()[0]) == 5 and not buf.cpu()[1:].any()
    # exactly the limit is taken
    edge = torch.zeros(4 * maxb, dtype=torch.float32, device="cuda")
    buf2 = torch.zeros((maxb * 4 * 12 + 8192) // 64 + 16, dtype=torch.int64, device="cuda")
    st2 = Z.Stream(buf2, cfg)
    fe = Z.make_field(edge)
    assert L().zfp_hip_compress_noalloc(st2.zs, ctypes.byref(fe)) > 0
    st2.close()


def test_noalloc_offset_index_is_refreshed_or_marked_stale():
    """A stream whose index buffer has room gets the new lengths (its decode
    reads through them); one without room keeps no stale index: the decode
    walks the stream.  Both decode to the oracle's values."""
    cfg = Z.ZfpConfig.fixed_precision(18)
    a = grf((32, 32, 32), seed=31)       # 512 blocks
    b = grf((32, 32, 32), seed=32)       # 512 blocks
    small = grf((16, 16, 16), seed=33)   # 64 blocks: its index buffer has no room for b's
    for first in (a, small):
        cap = (Z.maximum_size(cfg, b) + 64) // 8
        buf = torch.zeros(cap, dtype=torch.int64, device="cuda")
        st = Z.Stream(buf, cfg)
        st.compress(torch.from_numpy(first).cuda())   # leaves an index of first's blocks
        st.rewind()
        f = Z.make_field(torch.from_numpy(b).cuda())
        n = L().zfp_hip_compress_noalloc(st.zs, ctypes.byref(f))
        assert n > 0
        count = L().zfp_stream_index_export(st.bs, None, 0, None)
        assert count == (512 if first is a else 0)
        st.rewind()
        out = torch.zeros((32, 32, 32), dtype=torch.float32, device="cuda")
        st.decompress(out)
        st.close()
        torch.cuda.synchronize()
        on, _, ow = O.compress(b, O.fixed_precision(18))
        ref, _, _ = O.decompress(ow, b.shape, b.dtype, O.fixed_precision(18))
        assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32))
