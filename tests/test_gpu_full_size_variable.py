"""GPU: the variable-rate workloads at their full sizes -- 384^3 f64 (the k^-3
field scaled into [-1e3, 1e3]) at fixed accuracy 1e-6 and fixed precision 32,
8192^2 f32 at fixed precision 20, 256^3 i32 random walk reversible -- checked
through properties that do not need the oracle to code the whole array: the
offset index adds up to the stream, 512 sampled blocks are bit-exact against
the oracle (each block's bits at its indexed offset in the GPU stream equal the
oracle's code of that block, and its decoded values the oracle's), the same
stream decoded without its index (block lengths read on the host) equals the
indexed decode, and the whole round trip holds its bound (exact for
reversible, within the tolerance for fixed accuracy)."""

import numpy as np
import pytest

import oracle_py as O
import zfp_hip as Z

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from synth import grf  # noqa: E402

SAMPLES = 512


def zcfg(oc):
    return Z.ZfpConfig(oc.minbits, oc.maxbits, oc.maxprec, oc.minexp, Z.ZfpRounding.never())


def bits_of(words, start, n):
    """n bits of the LSB-first word stream from bit `start`, as an int."""
    w0, w1 = start // 64, (start + n + 63) // 64
    v = int.from_bytes(words[w0:w1].tobytes(), "little")
    return (v >> (start - 64 * w0)) & ((1 << n) - 1)


def block_view(arr, b):
    """Block b of a 2-D or 3-D numpy array (x = last axis, x fastest)."""
    nb = [(n + 3) // 4 for n in reversed(arr.shape)]
    c = []
    for m in nb:
        c.append(b % m)
        b //= m
    if arr.ndim == 2:
        return arr[4 * c[1]:4 * c[1] + 4, 4 * c[0]:4 * c[0] + 4]
    return arr[4 * c[2]:4 * c[2] + 4, 4 * c[1]:4 * c[1] + 4, 4 * c[0]:4 * c[0] + 4]


def cases():
    v1 = (grf((384, 384, 384), seed=11, dtype=np.float64, device="cuda") * 300.0).clamp_(-1e3, 1e3)
    yield "384^3 f64 accuracy 1e-6", v1, O.fixed_accuracy(1e-6), 1e-6
    yield "384^3 f64 precision 32", v1, O.fixed_precision(32), None
    del v1
    v2 = grf((8192, 8192), seed=12, dtype=np.float32, device="cuda")
    yield "8192^2 f32 precision 20", v2, O.fixed_precision(20), None
    del v2
    g = torch.Generator(device="cuda")
    g.manual_seed(13)
    v3 = torch.randint(-3, 4, (256 ** 3,), generator=g, device="cuda", dtype=torch.int32)
    v3 = torch.cumsum(v3, 0, dtype=torch.int32).reshape(256, 256, 256)
    yield "256^3 i32 reversible", v3, O.reversible(), 0


def test_full_size_variable_rate_sampled_blocks_match_oracle():
    rng = np.random.default_rng(5)
    for name, x, oc, tol in cases():
        cfg = zcfg(oc)
        cap = Z.maximum_size(cfg, x)
        buf = torch.zeros((cap + 7) // 8, dtype=torch.int64, device="cuda")
        s = Z.Stream(buf, cfg)
        n = s.compress(x)
        start, lens = s.index()
        nblocks = lens.size
        offs = np.concatenate([[0], np.cumsum(lens.astype(np.uint64))])
        assert start == 0 and (int(offs[-1]) + 63) // 64 * 8 == n, name
        words = buf[: n // 8].cpu().numpy().view(np.uint64)
        host = x.cpu().numpy()
        blocks = np.sort(rng.choice(nblocks, size=SAMPLES, replace=False))
        # the sampled blocks side by side along x, coded by the oracle
        packed = np.concatenate([block_view(host, int(b)) for b in blocks], axis=-1)
        on, oerr, ow = O.compress(np.ascontiguousarray(packed), oc)
        assert oerr == 0, name
        pos = 0
        for b in blocks:
            ln = int(lens[b])
            assert bits_of(words, int(offs[b]), ln) == bits_of(ow, pos, ln), (name, int(b))
            pos += ln
        assert on == (pos + 63) // 64 * 8, name  # the oracle's blocks are exactly as long
        # decode through the index; sampled values equal the oracle's decode
        y = torch.empty_like(x)
        s.rewind()
        assert s.decompress(y) == n, name
        ref, _, rerr = O.decompress(ow, packed.shape, packed.dtype, oc)
        assert rerr == 0, name
        got = y.cpu().numpy()
        u = np.uint32 if got.dtype.itemsize == 4 else np.uint64
        for k, b in enumerate(blocks):
            a = block_view(got, int(b))
            r = ref[..., 4 * k:4 * k + 4]
            assert np.array_equal(np.ascontiguousarray(a).view(u), np.ascontiguousarray(r).view(u)), (name, int(b))
        # the same words without their offset index: the host reads the block
        # lengths and the GPU decodes run by run; every value equals the
        # indexed decode (and so the sampled blocks equal the oracle's)
        s2 = Z.Stream(buf, cfg)
        y2 = torch.empty_like(x)
        assert s2.decompress(y2) == n, name
        iv = torch.int32 if x.element_size() == 4 else torch.int64
        assert torch.equal(y2.view(iv), y.view(iv)), name
        s2.close()
        del y2
        if tol == 0:
            assert torch.equal(x, y), name
        elif tol is not None:
            assert (x - y).abs().max().item() <= tol, name
        s.close()
        del buf, y, host, got
