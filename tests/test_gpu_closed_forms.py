"""The GPU library against the hand-written streams of the closed-form pins:
fixed-precision and fixed-accuracy DC-only blocks, ROUND_FIRST two-row blocks
and ROUND_LAST decodes of budget-cut streams, through the C ABI.

The expected bits come from the block grammar (the derivations are in
tests/test_oracle_closed_forms.py), not from the oracle's output.
"""
import math

import numpy as np
import pytest

import oracle_py as O
import zfp_hip as Z
from test_oracle_closed_forms import (NB32, NB64, _block_with_coefficients, dc_plane_codes,
                                      from_negabinary, negabinary, peel, peel_is_exact, s,
                                      two_row_plane_codes, words_of)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)


def gpu_compress(block, cfg):
    nwords = (Z.maximum_size(cfg, block) + 15) // 8
    buf = torch.zeros(nwords, dtype=torch.int64, device="cuda")
    st = Z.Stream(buf, cfg)
    n = st.compress(torch.from_numpy(np.ascontiguousarray(block)).cuda())
    st.close()
    return n, buf.cpu().numpy().view(np.uint64)


def gpu_decompress(words, shape, dtype, cfg):
    buf = torch.from_numpy(np.ascontiguousarray(words).view(np.int64).copy()).cuda()
    out = torch.zeros(shape, dtype=getattr(torch, np.dtype(dtype).name), device="cuda")
    st = Z.Stream(buf, cfg)
    st.decompress(out)
    st.close()
    return out.cpu().numpy()


INT_CASES = [(np.int32, 32), (np.int64, 64)]


@pytest.mark.parametrize("dtype,P", INT_CASES)
@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("p", [1, 5, 20, 31, 47])
def test_gpu_fixed_precision_dc_block(dtype, P, dims, p):
    c = -0x2345678
    block = np.full((4,) * dims, c, dtype=dtype)
    kmin = max(0, P - p)
    u0 = negabinary(c, P)
    want = words_of(dc_plane_codes(u0, P, kmin) or [0])
    n, words = gpu_compress(block, Z.ZfpConfig.fixed_precision(p))
    assert n == len(want) * 8
    assert np.array_equal(words[: len(want)], want)
    out = gpu_decompress(want, block.shape, dtype, Z.ZfpConfig.fixed_precision(p))
    assert np.all(out == from_negabinary(u0 & ~((1 << kmin) - 1), P))
    # ROUND_LAST: m = 0, prec = P - kmin + 1
    cfg = Z.ZfpConfig.fixed_precision(p).with_rounding(Z.ZfpRounding.last())
    out = gpu_decompress(want, block.shape, dtype, cfg)
    prec = P - kmin + 1
    nb = NB32 if P == 32 else NB64
    lo = (nb >> 1) >> prec if prec < P - 1 else 0
    mask = (1 << P) - 1
    coeffs = [from_negabinary(((u0 & ~((1 << kmin) - 1)) + lo) & mask, P)] + \
             [from_negabinary(lo, P)] * (4 ** dims - 1)
    if peel_is_exact(coeffs, dims, P):
        assert peel(out, dims, P) == coeffs


@pytest.mark.parametrize("dtype,P", INT_CASES)
@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("p", [11, 20, 27])
def test_gpu_round_first_two_row_block(dtype, P, dims, p):
    size = 4 ** dims
    kmin = P - p
    nb = NB32 if P == 32 else NB64
    b = (nb >> 2) >> p
    sgn = 1 if p % 2 else -1
    low = (0x5555555555555555 if sgn > 0 else 0xAAAAAAAAAAAAAAAA) & ((1 << kmin) - 1)
    d = from_negabinary(1 << kmin, P)
    for q in (d + from_negabinary(low, P), d + from_negabinary(low, P) - sgn * b):
        block = _block_with_coefficients(d, q, dims, dtype)
        want = words_of(two_row_plane_codes(negabinary(d + sgn * b, P), negabinary(q + sgn * b, P),
                                            P, kmin, size))
        cfg = Z.ZfpConfig.fixed_precision(p).with_rounding(Z.ZfpRounding.first())
        n, words = gpu_compress(block, cfg)
        assert n == len(want) * 8
        assert np.array_equal(words[: len(want)], want)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("tol_exp", [-30, -4, 3, 10])
def test_gpu_fixed_accuracy_float_dc_block(dtype, dims, tol_exp):
    c = -3.0
    ebits, ebias, P = (8, 127, 32) if dtype == np.float32 else (11, 1023, 64)
    emax = math.frexp(c)[1]
    dc = int(c * 2.0 ** (P - 2 - emax))
    prec = max(0, min(64, emax - tol_exp + 2 * dims + 2))
    if prec == 0:
        bits = [0]
    else:
        e = emax + ebias
        bits = [1] + [(e >> i) & 1 for i in range(ebits)]
        bits += dc_plane_codes(negabinary(dc, P), P, max(0, P - prec))
    want = words_of(bits)
    block = np.full((4,) * dims, c, dtype=dtype)
    n, words = gpu_compress(block, Z.ZfpConfig.fixed_accuracy(math.ldexp(1.0, tol_exp)))
    assert n == len(want) * 8
    assert np.array_equal(words[: len(want)], want)


@pytest.mark.parametrize("dtype,P", INT_CASES)
def test_gpu_round_last_budget_cut_states(dtype, P):
    """The hand-traced streams of test_round_last_budget_cut_states, decoded
    on the GPU: (bits, maxbits, coded coefficients, m, prec)."""
    nb = NB32 if P == 32 else NB64
    mask = (1 << P) - 1
    z = 6
    t = P - 1 - z
    Zs = [0] * z
    cases = [(Zs + s("11010001"), z + 8, [(1 << t) | (1 << (t - 1)) | (1 << (t - 3)), 0, 0, 0], 1, z + 4)]
    for a, b in ((0, 0), (1, 0), (0, 1), (1, 1)):
        cases.append((Zs + s("11110") + [a, b] + [1], z + 8,
                      [(1 << t) | (a << (t - 1)), (1 << t) | (b << (t - 1)), 1 << (t - 1), 0], 3, z + 2))
        cases.append((Zs + s("11110") + [a, b], z + 7,
                      [(1 << t) | (a << (t - 1)), (1 << t) | (b << (t - 1)), 0, 0], 2, z + 2))
    cases.append(([0] * 20, 20, [0, 0, 0, 0], 4, 20))
    for bits, maxbits, u, m, prec in cases:
        words = words_of(bits)
        cfg = Z.ZfpConfig.expert(maxbits, maxbits, 64, -1074)
        out = gpu_decompress(words, (4,), dtype, cfg)
        assert peel(out, 1, P) == [from_negabinary(x, P) for x in u], bits
        out = gpu_decompress(words, (4,), dtype, cfg.with_rounding(Z.ZfpRounding.last()))
        hi, lo = (nb >> 2) >> prec, (nb >> 1) >> prec
        want = [(x + (hi if i < m else lo)) & mask for i, x in enumerate(u)]
        assert peel(out, 1, P) == [from_negabinary(x, P) for x in want], bits
