"""GPU: injected reservation failures (zfp_hip_fail_after) leave the caller's
stream and field as they were, and the next call codes normally.

zfp-rs injects allocation failures into its parallel paths and checks that a
failed attempt writes nothing before its serial fallback gives the right
answer (src/compress.rs:164-194, src/decompress/pipeline.rs:76-94 and its test
at :295-335).  This library has no CPU fallback: a call that meets a failure
returns 0.  Each case arms the k-th reservation (workspace, staging, pinned
host memory) to fail, for every k the call reaches, and checks that the
failed call left the stream's words, its cursor and the output field
untouched; then an unarmed call gives the oracle's bytes and values."""

import ctypes

import numpy as np
import pytest

import oracle_py as O
import zfp_hip as Z

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from synth import grf  # noqa: E402

MAX_POINTS = 16


def zcfg(oc):
    return Z.ZfpConfig(oc.minbits, oc.maxbits, oc.maxprec, oc.minexp, Z.ZfpRounding.never())


def words_of(buf):
    return (buf.cpu().numpy() if torch.is_tensor(buf) else buf).copy()


def armed(k, call):
    """Run `call` with the k-th reservation armed to fail; (result, whether
    the failure was met)."""
    L = Z.lib()
    L.zfp_hip_fail_after(k)
    try:
        r = call()
    finally:
        left = L.zfp_hip_fail_after(-1)
    return r, left == -1


CASES = [
    ("rate 8", lambda: O.fixed_rate(8, O.FLOAT, 3)),
    ("accuracy 1e-4", lambda: O.fixed_accuracy(1e-4)),
]


@pytest.mark.parametrize("name,mk", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("device", [False, True], ids=["host", "device"])
def test_failed_compress_leaves_the_stream(name, mk, device):
    oc = mk()
    arr = grf((20, 24, 28), seed=31, scale=10.0)
    on, oerr, ow = O.compress(arr, oc)
    assert oerr == 0
    nw = len(ow) + 16
    L = Z.lib()
    met_any = False
    for k in range(MAX_POINTS):
        if device:
            buf = torch.full((nw,), 0x5a5a5a5a, dtype=torch.int64, device="cuda")
            x = torch.from_numpy(arr).cuda()
        else:
            buf = np.full(nw, 0x5a5a5a5a, dtype=np.uint64)
            x = arr
        s = Z.Stream(buf, zcfg(oc))
        before, pos = words_of(buf), s.wtell()
        f = Z.make_field(x)
        n, met = armed(k, lambda: L.zfp_compress(s.zs, ctypes.byref(f)))
        if not met:
            # the call reached fewer reservations: it ran unarmed
            assert n == on, (name, k)
            s.close()
            break
        met_any = True
        assert n == 0, (name, k)
        assert "injected reservation failure" in Z.last_error(), (name, k)
        assert s.wtell() == pos, (name, k)
        assert np.array_equal(words_of(buf), before), (name, k)
        # the next call on the same stream codes normally
        s.rewind()
        assert s.compress(x) == on, (name, k)
        got = words_of(buf).view(np.uint64)[: on // 8]
        assert np.array_equal(got, ow[: on // 8]), (name, k)
        s.close()
    # a fixed-rate compress of a device field into device words reserves
    # nothing (no lengths, offsets or staging), so it cannot meet a failure
    assert met_any == (not (device and name.startswith("rate"))), name


@pytest.mark.parametrize("indexed", [True, False], ids=["index", "no-index"])
@pytest.mark.parametrize("device", [False, True], ids=["host", "device"])
def test_failed_decompress_leaves_the_field(indexed, device):
    oc = O.fixed_accuracy(1e-4)
    arr = grf((20, 24, 28), seed=32, scale=10.0)
    on, oerr, ow = O.compress(arr, oc)
    ref, _, _ = O.decompress(ow, arr.shape, arr.dtype, oc)
    L = Z.lib()
    met_any = False
    for k in range(MAX_POINTS):
        words = np.concatenate([ow, np.zeros(8, dtype=np.uint64)])
        buf = torch.from_numpy(words.view(np.int64)).cuda() if device else words
        s = Z.Stream(buf, zcfg(oc))
        if indexed:
            s.rewind()
            assert s.compress(torch.from_numpy(arr).cuda() if device else arr) == on
            s.rewind()
        if device:
            out = torch.full(arr.shape, -7.0, dtype=torch.float32, device="cuda")
        else:
            out = np.full(arr.shape, -7.0, dtype=np.float32)
        pos = s.rtell()
        f = Z.make_field(out)
        n, met = armed(k, lambda: L.zfp_decompress(s.zs, ctypes.byref(f)))
        if not met:
            assert n == on, k
            s.close()
            break
        met_any = True
        assert n == 0, k
        assert s.rtell() == pos, k
        host = out.cpu().numpy() if device else out
        assert (host == -7.0).all(), k
        assert s.decompress(out) == on, k
        host = out.cpu().numpy() if device else out
        assert np.array_equal(host.view(np.uint32), ref.view(np.uint32)), k
        s.close()
    assert met_any


def test_failed_batch_leaves_the_streams():
    """A batch whose reservations fail codes none of its arrays and leaves
    every stream's words and cursor; the next batch codes them all."""
    oc = O.fixed_precision(16)
    arrs = [grf((8 + 4 * i, 12, 9), seed=40 + i) for i in range(5)]
    refs = [O.compress(a, oc) for a in arrs]
    bufs = [torch.full((len(r[2]) + 8,), 0x77, dtype=torch.int64, device="cuda") for r in refs]
    xs = [torch.from_numpy(a).cuda() for a in arrs]
    streams = [Z.Stream(b, zcfg(oc)) for b in bufs]
    before = [words_of(b) for b in bufs]
    L = Z.lib()
    (count, sizes), met = armed(0, lambda: Z.compress_batch(streams, xs, Z.BATCH_DEVICE_FIELDS))
    assert met and count == 0 and not any(sizes)
    for s, b, w in zip(streams, bufs, before):
        assert s.wtell() == 0 and np.array_equal(words_of(b), w)
    count, sizes = Z.compress_batch(streams, xs, Z.BATCH_DEVICE_FIELDS)
    assert count == len(arrs)
    for (on, _, ow), b, n in zip(refs, bufs, sizes):
        assert n == on and np.array_equal(words_of(b).view(np.uint64)[: on // 8], ow[: on // 8])
    for s in streams:
        s.close()
    assert L.zfp_hip_fail_after(-1) == -1
