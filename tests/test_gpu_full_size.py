"""GPU: the large fixed-rate workload at its full size -- the 512^3 f32 seeded
k^-3 field with 5 % zero blocks and 1 % subnormals, rates 4, 8 and 16, every
rounding -- checked through properties that do not need the oracle to code the
whole field: the exact stream size, 2048 sampled blocks bit-exact against the
oracle (fixed-rate blocks are independent and word-aligned at these rates, so
each block's words in the GPU stream equal the oracle's stream of that block
alone, and its decoded values equal the oracle's), and identical bytes from a
second compression."""

import numpy as np
import pytest

import oracle_py as O
import zfp_hip as Z

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from synth import grf  # noqa: E402

N = 512
NB = N // 4
SAMPLES = 2048

ZROUND = {O.ROUND_NEVER: Z.ROUND_NEVER, O.ROUND_FIRST: Z.ROUND_FIRST, O.ROUND_LAST: Z.ROUND_LAST}


@pytest.fixture(scope="module")
def field():
    x = grf((N, N, N), seed=2024, zero_blocks=0.05, subnormals=0.01, device="cuda")
    torch.cuda.synchronize()
    host = x.cpu().numpy()
    rng = np.random.default_rng(77)
    blocks = np.sort(rng.choice(NB ** 3, size=SAMPLES, replace=False))
    bz, rem = np.divmod(blocks, NB * NB)
    by, bx = np.divmod(rem, NB)
    # the sampled blocks side by side along x: block k at x = 4k
    packed = np.empty((4, 4, 4 * SAMPLES), dtype=np.float32)
    for k, (z, y, xx) in enumerate(zip(bz, by, bx)):
        packed[:, :, 4 * k:4 * k + 4] = host[4 * z:4 * z + 4, 4 * y:4 * y + 4, 4 * xx:4 * xx + 4]
    yield x, host, blocks, packed
    del x


@pytest.mark.parametrize("rate", [4, 8, 16])
def test_full_size_fixed_rate_sampled_blocks_match_oracle(field, rate):
    x, host, blocks, packed = field
    out = torch.empty_like(x)
    base = O.fixed_rate(rate, O.FLOAT, 3)
    bits = base.maxbits
    W = bits // 64
    assert bits % 64 == 0
    nblocks = NB ** 3
    for rnd, tight in ((O.ROUND_NEVER, False), (O.ROUND_FIRST, False), (O.ROUND_LAST, False),
                       (O.ROUND_FIRST, True)):
        oc = O.config(base.minbits, base.maxbits, base.maxprec, base.minexp, rnd, tight)
        cfg = Z.ZfpConfig(oc.minbits, oc.maxbits, oc.maxprec, oc.minexp,
                          Z.ZfpRounding(ZROUND[rnd], tight))
        ctx = f"rate {rate} rounding {rnd}/{tight}"
        buf = torch.zeros(nblocks * W + 1, dtype=torch.int64, device="cuda")
        s = Z.Stream(buf, cfg)
        n = s.compress(x)
        assert n == nblocks * bits // 8, ctx
        words = buf[: nblocks * W].view(nblocks, W)[torch.from_numpy(blocks).cuda()].cpu().numpy()
        on, oerr, ow = O.compress(packed, oc)
        assert oerr == 0 and on == SAMPLES * bits // 8, ctx
        assert np.array_equal(words.view(np.uint64).ravel(), ow[: SAMPLES * W]), ctx
        # decode: the sampled blocks' values equal the oracle's decode of them
        s.rewind()
        assert s.decompress(out) == n, ctx
        ref, rn, rerr = O.decompress(ow, packed.shape, np.float32, oc)
        assert rerr == 0, ctx
        got = out.cpu().numpy()
        bz, rem = np.divmod(blocks, NB * NB)
        by, bx = np.divmod(rem, NB)
        for k in range(0, SAMPLES, 7):
            z, y, xx = bz[k], by[k], bx[k]
            a = got[4 * z:4 * z + 4, 4 * y:4 * y + 4, 4 * xx:4 * xx + 4]
            assert np.array_equal(a.view(np.uint32), ref[:, :, 4 * k:4 * k + 4].view(np.uint32)), (ctx, k)
        if rnd == O.ROUND_NEVER:
            # a second compression gives the same bytes
            again = torch.zeros_like(buf)
            s2 = Z.Stream(again, cfg)
            assert s2.compress(x) == n
            assert torch.equal(again, buf), ctx
            s2.close()
        s.close()
        del buf
