"""The GPU block codec's own source (zfp-rs_amd/csrc/zfpx_*.h), built for the CPU,
against the oracle: stream words and decoded values bit for bit.

This runs without a GPU and checks the exact per-block code and block placement
(fixed-rate seek, variable-rate length pass + offsets, atomic merge of shared
words, a start offset inside a word) that the HIP kernels execute.
"""
import itertools

import numpy as np
import pytest

import hc_util as H
import oracle_py as O
from fields import random_walk, smooth, specials

PREFIX_WORD = 0x0123456789abcdef
PREFIX_PART = 0x5a5a5a5a5a5a5a5a

SHAPES = {1: [(37,), (64,)], 2: [(9, 13), (16, 16)], 3: [(5, 6, 7), (8, 8, 8)]}
TYPES = [O.INT32, O.INT64, O.FLOAT, O.DOUBLE]
ROUNDINGS = [(O.ROUND_NEVER, False), (O.ROUND_FIRST, False), (O.ROUND_FIRST, True),
             (O.ROUND_LAST, False), (O.ROUND_LAST, True)]


def is_fixed(cfg, ztype, dims):
    return (cfg.minbits == cfg.maxbits and
            O.lib().orc_block_bits(cfg, ztype, dims) == cfg.maxbits)


def configs(ztype, dims):
    out = []
    for rate in (1, 3, 8, 16, 27):
        out.append(("rate%g" % rate, O.fixed_rate(rate, ztype, dims)))
    out.append(("rate8a", O.fixed_rate(8, ztype, dims, align=True)))
    for p in (5, 16, 33, 64):
        out.append(("prec%d" % p, O.fixed_precision(p)))
    for tol in (1e-1, 1e-4, 1e-9):
        out.append(("acc%g" % tol, O.fixed_accuracy(tol)))
    out.append(("rev", O.reversible()))
    out.append(("exp_small", O.config(1, 11, 64, -1074)))
    out.append(("exp_mixed", O.config(100, 600, 20, -30)))
    out.append(("exp_revmin", O.config(100, 16658, 64, -1075)))
    return out


def data_for(ztype, shape, kind, seed):
    dt = O.DTYPES[ztype]
    if kind == "smooth":
        return smooth(shape, dt, seed, scale=1e3 if np.dtype(dt).kind == "f" else 1.0)
    if kind == "walk":
        return random_walk(shape, dt, seed)
    return specials(shape, dt, seed)


def median_group_words(lens, start_bits, T):
    """The median count of stream words that workgroups of T blocks touch."""
    offs = start_bits + np.concatenate([[0], np.cumsum(lens.astype(np.int64))])
    b0 = np.arange(0, lens.size, T)
    b1 = np.minimum(b0 + T, lens.size)
    nw = (offs[b1] + 63) // 64 - offs[b0] // 64
    return max(int(np.median(nw)), 1)


def run_case(arr, ztype, cfg, start_bits):
    dims = arr.ndim
    n, err, ow = O.compress(arr, cfg, start_bits=start_bits)
    assert err == 0
    nwords = n // 8
    fixed = cfg.maxbits if is_fixed(cfg, ztype, dims) else 0
    rev = cfg.minexp < -1074
    ref, rn, rerr = O.decompress(ow, arr.shape, arr.dtype, cfg, start_bits=start_bits)
    assert rerr == 0
    isz = arr.dtype.itemsize
    u = np.uint32 if isz == 4 else np.uint64
    # the global-memory path, and the LDS path with workgroups of 3 blocks
    # (group boundaries inside the field) and of 64; variable-rate streams (whose
    # LDS path is the write pass; they decode from global memory) also with a
    # staging capacity of the median workgroup's words, so that dense
    # workgroups are written block by block beside staged ones
    if fixed:
        variants = [(0, 0), (3, 0), (64, 0)] if fixed <= 4096 else [(0, 0)]
    else:
        variants = [(0, 0), (3, 0), (64, 0), (3, None)]
    for lds_T, lds_cap in variants:
        if lds_cap is None:
            lds_cap = median_group_words(lens0, start_bits, lds_T)
        words = np.zeros(nwords + 4, dtype=np.uint64)
        words[: start_bits // 64] = PREFIX_WORD
        if start_bits % 64:
            words[start_bits // 64] = PREFIX_PART & ((1 << (start_bits % 64)) - 1)
        end, lens = H.compress(arr, ztype, cfg, words, start_bits, fixed, rev, lds_T, lds_cap)
        assert (end + 63) // 64 * 8 == n, f"stream length (lds_T={lds_T}, cap={lds_cap})"
        assert np.array_equal(words[:nwords], ow[:nwords]), f"stream words (lds_T={lds_T}, cap={lds_cap})"
        assert not words[nwords:].any()
        got, loaded = H.decompress(arr.shape, arr.dtype, ztype, cfg, words, start_bits, fixed,
                                   None if fixed else lens, rev, lds_T)
        assert np.array_equal(got.view(u), ref.view(u)), f"decoded values (lds_T={lds_T}, cap={lds_cap})"
        if lds_T == 0:
            lens0 = lens
        if lds_T and fixed:
            # a whole stream is read without loading past its end
            assert 0 < loaded <= nwords, f"loaded words (lds_T={lds_T}, cap={lds_cap})"


def test_lds_path_capacity_and_overread():
    """Fixed-rate LDS path against a stream cut short: writes past the
    capacity are dropped, reads past it return zero bits (bitstream/core.rs
    write and read at the end of the buffer)."""
    arr = smooth((8, 8, 12), np.float32, 5, scale=10.0)
    for rate, start in ((8, 0), (5, 0), (5, 13)):
        cfg = O.fixed_rate(rate, O.FLOAT, 3)
        n, err, ow = O.compress(arr, cfg, start_bits=start)
        cut = n // 8 - 3
        ref_words = ow[:cut].copy()
        on, oerr, ow2 = O.compress(arr, cfg, capacity=cut * 8, start_bits=start)
        words = np.zeros(cut, dtype=np.uint64)
        words[: start // 64] = PREFIX_WORD
        if start % 64:
            words[start // 64] = PREFIX_PART & ((1 << (start % 64)) - 1)
        H.compress(arr, O.FLOAT, cfg, words, start, cfg.maxbits, False, 4)
        assert np.array_equal(words, ow2[:cut])
        ref, _, _ = O.decompress(ref_words, arr.shape, arr.dtype, cfg, start_bits=start)
        got, loaded = H.decompress(arr.shape, arr.dtype, O.FLOAT, cfg, ref_words, start,
                                   cfg.maxbits, None, False, 4)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        assert loaded > cut


@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("ztype", TYPES)
def test_modes_match_oracle(dims, ztype):
    for shape in SHAPES[dims]:
        for kind in ("smooth", "specials") if ztype in (O.FLOAT, O.DOUBLE) else ("smooth", "walk", "specials"):
            arr = data_for(ztype, shape, kind, seed=hash((dims, ztype, shape, kind)) & 0xffff)
            for name, cfg in configs(ztype, dims):
                lossy_special = kind == "specials" and cfg.minexp >= -1074
                for start in (0, 13):
                    try:
                        run_case(arr, ztype, cfg, start)
                    except AssertionError as e:
                        raise AssertionError(f"{name} {kind} {shape} start={start}: {e}") from None


@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("ztype", TYPES)
def test_roundings_match_oracle(dims, ztype):
    shape = SHAPES[dims][0]
    arr = data_for(ztype, shape, "smooth", seed=7 * dims + ztype)
    for (rnd, tight), (name, base) in itertools.product(ROUNDINGS, configs(ztype, dims)):
        cfg = O.config(base.minbits, base.maxbits, base.maxprec, base.minexp, rnd, tight)
        try:
            run_case(arr, ztype, cfg, 0)
        except AssertionError as e:
            raise AssertionError(f"{name} rounding={rnd} tight={tight}: {e}") from None


def test_block_grid_division_is_exact():
    """The block-grid coordinates use a multiply-and-shift division by the
    grid's extents; it must equal integer division for every dividend."""
    rng = np.random.default_rng(21)
    edge = np.array([0, 1, 2, 3, 0x7fffffff, 0x80000000, 0x80000001, 0xfffffffe, 0xffffffff],
                    dtype=np.uint64)
    divisors = [1, 2, 3, 5, 7, 64, 127, 128, 129, 1000, 4097, 65535, 65536, 65537,
                (1 << 31) - 1, 1 << 31, (1 << 31) + 1, 0xfffffffe, 0xffffffff]
    divisors += [int(d) for d in rng.integers(1, 1 << 32, size=40, dtype=np.uint64)]
    for d in divisors:
        n = np.concatenate([edge, rng.integers(0, 1 << 32, size=4000, dtype=np.uint64),
                            np.array([k * d + r for k in (1, 2, 1000) for r in (0, d - 1)
                                      if k * d + r < (1 << 32)], dtype=np.uint64)])
        got = H.divide(n.astype(np.uint32), d)
        assert np.array_equal(got.astype(np.uint64), n // np.uint64(d)), d


def random_words(rng, n, density):
    """n words whose bits are one with probability `density`."""
    bits = rng.random((n, 64)) < density
    return np.packbits(bits, axis=1, bitorder="little").view(np.uint64).reshape(-1).copy()


@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("ztype", TYPES)
def test_random_streams_decode_as_the_oracle(dims, ztype):
    """Arbitrary bit streams decode as the oracle decodes them: budget cuts
    inside runs, runs that reach the last coefficient, tests longer than a
    window and every header, at sparse and dense bit densities."""
    rng = np.random.default_rng(dims * 10 + ztype)
    shape = {1: (37,), 2: (9, 13), 3: (5, 6, 7)}[dims]
    dt = O.DTYPES[ztype]
    u = np.uint32 if np.dtype(dt).itemsize == 4 else np.uint64
    cfgs = [O.fixed_rate(r, ztype, dims) for r in (1, 2, 5, 8, 13, 16, 24, 32)]
    cfgs += [O.fixed_precision(p) for p in (3, 12, 33)]
    cfgs += [O.fixed_accuracy(1e-3), O.reversible(), O.config(20, 300, 40, -60)]
    for cfg in cfgs:
        for rnd in (O.ROUND_NEVER, O.ROUND_LAST):
            c = O.config(cfg.minbits, cfg.maxbits, cfg.maxprec, cfg.minexp, rnd, False)
            for density in (0.5, 0.1, 0.02):
                nblocks = int(np.prod([(v + 3) // 4 for v in shape]))
                nwords = (nblocks * max(c.maxbits, 64) + 63) // 64 + 8
                nwords = min(nwords, nblocks * 300 + 64)
                words = random_words(rng, nwords, density)
                fixed = c.maxbits if is_fixed(c, ztype, dims) else 0
                rev = c.minexp < -1074
                ref, rn, rerr = O.decompress(words, shape, dt, c)
                lens = None
                if not fixed:
                    offs = O.block_offsets(words, shape, dt, c)
                    lens = np.diff(offs.astype(np.int64)).astype(np.uint32)
                got, _ = H.decompress(shape, dt, ztype, c, words, 0, fixed, lens, rev, 0)
                assert np.array_equal(got.view(u), ref.view(u)), (c.minbits, c.maxbits, c.maxprec,
                                                                  c.minexp, rnd, density)
                if fixed and fixed <= 4096:
                    got, _ = H.decompress(shape, dt, ztype, c, words, 0, fixed, None, rev, 3)
                    assert np.array_equal(got.view(u), ref.view(u))


@pytest.mark.parametrize("ztype", TYPES)
def test_many_3d_blocks_near_the_last_coefficient(ztype):
    """3-D blocks whose planes are parsed with few coefficients left (the
    decoder's 32-bit window, taken when every block of a wave has at most 16
    coefficients left, beside the 64-bit one) decode as the oracle decodes
    them: smooth fields at fixed rates where most blocks end with all 64
    coefficients significant, and arbitrary streams (budget cuts inside runs,
    runs reaching the last coefficient) at several bit densities."""
    rng = np.random.default_rng(300 + ztype)
    shape = (16, 20, 24)
    dt = O.DTYPES[ztype]
    u = np.uint32 if np.dtype(dt).itemsize == 4 else np.uint64
    nblocks = 4 * 5 * 6
    for rate in (4, 8, 12, 16, 24, 31):
        for rnd in (O.ROUND_NEVER, O.ROUND_LAST):
            base = O.fixed_rate(rate, ztype, 3)
            c = O.config(base.minbits, base.maxbits, base.maxprec, base.minexp, rnd, False)
            arr = data_for(ztype, shape, "smooth", rate)
            n, err, ow = O.compress(arr, c)
            assert err == 0
            ref, _, rerr = O.decompress(ow, shape, dt, c)
            assert rerr == 0
            for lds_T in (0, 64):
                got, _ = H.decompress(shape, dt, ztype, c, ow, 0, c.maxbits, None, False, lds_T)
                assert np.array_equal(got.view(u), ref.view(u)), (rate, rnd, lds_T)
            for density in (0.6, 0.5, 0.3, 0.1):
                words = random_words(rng, (nblocks * c.maxbits + 63) // 64 + 8, density)
                ref, _, rerr = O.decompress(words, shape, dt, c)
                got, _ = H.decompress(shape, dt, ztype, c, words, 0, c.maxbits, None, False, 64)
                assert np.array_equal(got.view(u), ref.view(u)), (rate, rnd, density)


@pytest.mark.parametrize("dims", [1, 2, 3, 4])
@pytest.mark.parametrize("ztype", TYPES)
def test_length_reader_matches_the_oracle(dims, ztype):
    """The length reader of index-less decodes (walk_offsets in zfpx_field.h,
    which the library runs on the host) finds the block offsets of the
    oracle's serial decode: on compressed fields of every variable-rate mode,
    min_bits padding included, from a start bit inside a word, and (1-3-D) on
    arbitrary bit streams, where budgets cut planes short and runs of zeros
    reach the last coefficient."""
    rng = np.random.default_rng(100 + 10 * dims + ztype)
    shape = {1: (37,), 2: (9, 13), 3: (5, 6, 7), 4: (3, 5, 4, 6)}[dims]
    dt = O.DTYPES[ztype]
    nblocks = int(np.prod([(v + 3) // 4 for v in shape]))
    cfgs = [O.fixed_precision(p) for p in (5, 16, 33, 64)]
    cfgs += [O.fixed_accuracy(t) for t in (1e-1, 1e-4, 1e-9)]
    cfgs += [O.reversible(), O.config(1, 11, 64, -1074), O.config(100, 600, 20, -30),
             O.config(100, 16658, 64, -1075), O.fixed_rate(8, ztype, dims)]
    kinds = ("smooth", "specials") if np.dtype(dt).kind == "f" else ("smooth", "walk")
    for cfg in cfgs:
        rev = cfg.minexp < -1074
        for kind in kinds:
            arr = data_for(ztype, shape, kind, seed=dims + 7 * ztype)
            if dims == 4:
                words, ref = O.compress_blocks(arr, cfg, 4 * 4 * 16658 // 64 + 64)
                got = H.walk(ztype, dims, cfg, words, 0, nblocks, rev)
                assert np.array_equal(got, ref), (cfg.minbits, cfg.maxbits, cfg.maxprec, kind)
                continue
            for sb in (0, 13):
                n, err, words = O.compress(arr, cfg, start_bits=sb)
                assert err == 0
                ref = O.block_offsets(words, shape, dt, cfg, start_bits=sb)
                got = H.walk(ztype, dims, cfg, words, sb, nblocks, rev)
                assert np.array_equal(got, ref), (cfg.minbits, cfg.maxbits, cfg.maxprec, kind, sb)
        if dims == 4:
            continue
        for density in (0.5, 0.1, 0.02):
            words = random_words(rng, nblocks * 80 + 64, density)
            ref = O.block_offsets(words, shape, dt, cfg)
            got = H.walk(ztype, dims, cfg, words, 0, nblocks, rev)
            assert np.array_equal(got, ref), (cfg.minbits, cfg.maxbits, cfg.maxprec, density)


@pytest.mark.parametrize("dims", [1, 2, 3])
@pytest.mark.parametrize("ztype", TYPES)
def test_one_pass_compress_matches_oracle(dims, ztype):
    """The one-pass variable-rate compress (prepare and count each block,
    scan, look back, code from the same registers, merge the words groups
    share) gives the oracle's stream and block lengths for every
    variable-rate mode and rounding, from a start bit inside a word with the
    caller's bits below it, into words it does not clear first: with the
    plan's workgroups, with groups of 64 and 3 blocks, and with an LDS room of
    3 words, which sends most groups straight to global memory."""
    rng = np.random.default_rng(7 * dims + ztype)
    cfgs = [c for name, c in configs(ztype, dims) if not is_fixed(c, ztype, dims)]
    for rnd, tight in ROUNDINGS[1:]:
        b = O.fixed_precision(20)
        cfgs.append(O.config(b.minbits, b.maxbits, b.maxprec, b.minexp, rnd, tight))
    kinds = ("smooth", "specials") if ztype in (O.FLOAT, O.DOUBLE) else ("smooth", "walk")
    for shape in SHAPES[dims]:
        for kind in kinds:
            arr = data_for(ztype, shape, kind, seed=int(rng.integers(1 << 30)))
            for cfg in cfgs:
                rev = cfg.minexp < -1074
                for sb in (0, 77):
                    n, err, ow = O.compress(arr, cfg, start_bits=sb)
                    assert err == 0
                    nwords = n // 8
                    offs = O.block_offsets(ow, shape, arr.dtype, cfg, start_bits=sb)
                    for T, cap in ((0, 0), (64, 0), (3, 0), (3, 3), (64, 3)):
                        words = np.full(nwords + 4, 0xa5a5a5a5a5a5a5a5, dtype=np.uint64)
                        words[: sb // 64] = PREFIX_WORD
                        if sb % 64:
                            words[sb // 64] = PREFIX_PART & ((1 << (sb % 64)) - 1) | (0xf0f0 << (sb % 64))
                        end, lens = H.compress_1p(arr, ztype, cfg, words, sb, rev, T, cap)
                        ctx = (shape, kind, cfg.minbits, cfg.maxbits, cfg.maxprec, cfg.minexp,
                               cfg.rounding, sb, T, cap)
                        assert end == offs[-1], ctx
                        assert np.array_equal(lens, np.diff(offs.astype(np.int64)).astype(np.uint32)), ctx
                        assert np.array_equal(words[:nwords], ow[:nwords]), ctx
                        assert (words[nwords:] == 0xa5a5a5a5a5a5a5a5).all(), ctx
