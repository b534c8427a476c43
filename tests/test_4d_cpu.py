"""4-D fields: the oracle's block loop pinned against its field driver, then the
GPU block codec's source (built for the CPU) against the oracle on 4-D fields,
stream words and decoded values bit for bit.

The oracle's 4-D block coder itself is pinned by the reference's 4-D checksum
known answers (test_oracle_kat.py); 4-D fields are the reference's serial
compress loop over that coder (src/compress.rs:22-115,
src/codec/encode/dim4.rs, src/codec/decode/dim4.rs)."""
import numpy as np
import pytest

import hc_util as H
import oracle_py as O
from fields import random_walk, smooth, specials

TYPES = [O.INT32, O.INT64, O.FLOAT, O.DOUBLE]


def configs(ztype):
    out = [("rate%g" % r, O.fixed_rate(r, ztype, 4)) for r in (1, 4, 16)]
    out += [("prec%d" % p, O.fixed_precision(p)) for p in (6, 24, 64)]
    if ztype in (O.FLOAT, O.DOUBLE):
        out += [("acc%g" % t, O.fixed_accuracy(t)) for t in (1e-2, 1e-7)]
    out += [("rev", O.reversible()), ("exp_mixed", O.config(300, 2000, 20, -30)),
            ("exp_small", O.config(1, 40, 64, -1074))]
    for rnd, tight in ((O.ROUND_FIRST, False), (O.ROUND_LAST, True)):
        b = O.fixed_precision(12)
        out.append((f"prec12_r{rnd}", O.config(b.minbits, b.maxbits, b.maxprec, b.minexp, rnd, tight)))
    return out


def data(ztype, shape, kind, seed):
    dt = O.DTYPES[ztype]
    if kind == "specials":
        return specials(shape, dt, seed)
    if kind == "walk" or np.dtype(dt).kind != "f":
        return random_walk(shape, dt, seed)
    return smooth(shape, dt, seed, scale=1e3)


@pytest.mark.parametrize("shape", [(5, 6, 7), (8, 4, 9)])
def test_block_loop_equals_field_driver(shape):
    """The Python block loop the 4-D checks use gives orc_compress's stream on
    3-D fields, so it orders and pads blocks as the field driver does."""
    for ztype in (O.FLOAT, O.INT64):
        arr = data(ztype, shape, "smooth", 3)
        for _, cfg in configs(ztype)[:8]:
            n, err, ow = O.compress(arr, cfg)
            assert err == 0
            words, offs = O.compress_blocks(arr, cfg, n // 8 + 8)
            assert np.array_equal(words[: n // 8], ow[: n // 8]) and not words[n // 8:].any()
            assert np.array_equal(offs, O.block_offsets(ow, arr.shape, arr.dtype, cfg))
            ref, _, _ = O.decompress(ow, arr.shape, arr.dtype, cfg)
            got = O.decompress_blocks(words, arr.shape, arr.dtype, cfg)
            u = np.uint32 if arr.dtype.itemsize == 4 else np.uint64
            assert np.array_equal(got.view(u), ref.view(u))


@pytest.mark.parametrize("ztype", TYPES)
def test_4d_device_codec_matches_oracle(ztype):
    shape = (3, 5, 4, 6)  # w, z, y, x: 1 x 2 x 1 x 2 blocks, partial in w, z and x
    kinds = ("smooth", "specials") if ztype in (O.FLOAT, O.DOUBLE) else ("smooth", "walk")
    for kind in kinds:
        arr = data(ztype, shape, kind, seed=11 * ztype + len(kind))
        u = np.uint32 if arr.dtype.itemsize == 4 else np.uint64
        for name, cfg in configs(ztype):
            ctx = f"{name} {kind}"
            cap = 4 * 4 * 16658 // 64 + 64
            ow, offs = O.compress_blocks(arr, cfg, cap)
            nwords = (int(offs[-1]) + 63) // 64
            ref = O.decompress_blocks(ow, shape, arr.dtype, cfg)
            fixed = cfg.maxbits if (cfg.minbits == cfg.maxbits and
                                    O.lib().orc_block_bits(cfg, ztype, 4) == cfg.maxbits) else 0
            rev = cfg.minexp < -1074
            words = np.zeros(nwords + 4, dtype=np.uint64)
            end, lens = H.compress(arr, ztype, cfg, words, 0, fixed, rev)
            assert end == offs[-1], ctx
            assert np.array_equal(words[:nwords], ow[:nwords]) and not words[nwords:].any(), ctx
            got, _ = H.decompress(shape, arr.dtype, ztype, cfg, words, 0, fixed,
                                  None if fixed else lens, rev)
            assert np.array_equal(got.view(u), ref.view(u)), ctx


def sparse(shape, dt, seed):
    """Mostly zero, with a few spikes: long zero scans across plane words."""
    rng = np.random.default_rng(seed)
    a = np.zeros(shape, dtype=dt)
    flat = a.reshape(-1)
    idx = rng.choice(flat.size, size=max(1, flat.size // 40), replace=False)
    if np.dtype(dt).kind == "f":
        flat[idx] = rng.standard_normal(idx.size) * 10.0 ** rng.integers(-3, 4, idx.size)
    else:
        info = np.iinfo(dt)
        flat[idx] = rng.integers(info.min // 4, info.max // 4, idx.size, dtype=dt)
    return a


@pytest.mark.parametrize("ztype", TYPES)
def test_4d_coder_many_blocks_and_budgets(ztype):
    """24 blocks (2 x 2 x 2 x 3, partial in x) of smooth, random-walk and
    sparse data under budgets that end planes at many places (rates that are
    not whole planes, low precisions, ROUND_LAST under a binding rate)."""
    shape = (8, 8, 8, 11)
    dt = O.DTYPES[ztype]
    cfgs = [("rate%g" % r, O.fixed_rate(r, ztype, 4)) for r in (0.5, 2.5, 5, 8, 12.25, 20, 31)]
    cfgs += [("prec%d" % p, O.fixed_precision(p)) for p in (3, 17, 40)]
    cfgs += [("rev", O.reversible())]
    b = O.fixed_rate(6, ztype, 4)
    cfgs.append(("rate6_last", O.config(b.minbits, b.maxbits, b.maxprec, b.minexp, O.ROUND_LAST, True)))
    if ztype in (O.FLOAT, O.DOUBLE):
        cfgs.append(("acc1e-3", O.fixed_accuracy(1e-3)))
    for kind in ("smooth", "walk", "sparse"):
        arr = sparse(shape, dt, 5 + ztype) if kind == "sparse" else data(ztype, shape, kind, seed=3 * ztype + 1)
        u = np.uint32 if arr.dtype.itemsize == 4 else np.uint64
        for name, cfg in cfgs:
            ctx = f"{name} {kind}"
            cap = 24 * 16658 // 64 + 64
            ow, offs = O.compress_blocks(arr, cfg, cap)
            nwords = (int(offs[-1]) + 63) // 64
            ref = O.decompress_blocks(ow, shape, arr.dtype, cfg)
            fixed = cfg.maxbits if (cfg.minbits == cfg.maxbits and
                                    O.lib().orc_block_bits(cfg, ztype, 4) == cfg.maxbits) else 0
            rev = cfg.minexp < -1074
            words = np.zeros(nwords + 4, dtype=np.uint64)
            end, lens = H.compress(arr, ztype, cfg, words, 0, fixed, rev)
            assert end == offs[-1], ctx
            assert np.array_equal(words[:nwords], ow[:nwords]) and not words[nwords:].any(), ctx
            got, _ = H.decompress(shape, arr.dtype, ztype, cfg, words, 0, fixed,
                                  None if fixed else lens, rev)
            assert np.array_equal(got.view(u), ref.view(u)), ctx
            if fixed:
                # the fixed-rate compress and decode through LDS (k_encode_fx,
                # k_decode_fx; 4-D blocks take the whole-block layouts only,
                # and fall back to the global kernels otherwise)
                for T in (64, 7):
                    w2 = np.zeros(nwords + 4, dtype=np.uint64)
                    H.compress(arr, ztype, cfg, w2, 0, fixed, rev, lds_T=T)
                    assert np.array_equal(w2, words), f"{ctx} T={T}"
                    got, _ = H.decompress(shape, arr.dtype, ztype, cfg, words, 0, fixed, None, rev, lds_T=T)
                    assert np.array_equal(got.view(u), ref.view(u)), f"{ctx} T={T}"
