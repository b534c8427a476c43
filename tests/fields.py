"""Seeded synthetic fields for tests (numpy arrays, C order: x is the last axis)."""
import numpy as np


def smooth(shape, dtype, seed, scale=1.0):
    """A smooth random field: sum of a few random low-frequency waves plus noise."""
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.linspace(0, 1, n) for n in shape], indexing="ij")
    f = np.zeros(shape)
    for _ in range(4):
        phase = rng.uniform(0, 2 * np.pi)
        k = rng.uniform(0.5, 3.0, size=len(shape))
        arg = sum(kk * g for kk, g in zip(k, grids)) * 2 * np.pi + phase
        f += rng.normal() * np.sin(arg)
    f += 1e-3 * rng.normal(size=shape)
    f *= scale
    if np.dtype(dtype).kind == "f":
        return f.astype(dtype)
    return np.round(f * 1e6).astype(dtype)


def random_walk(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    steps = rng.integers(-3, 4, size=int(np.prod(shape)))
    return np.cumsum(steps).reshape(shape).astype(dtype)


def specials(shape, dtype, seed):
    """Values with NaN, infinities, signed zeros, subnormals and tiny blocks mixed in."""
    rng = np.random.default_rng(seed)
    a = smooth(shape, dtype, seed).ravel().copy()
    if np.dtype(dtype).kind == "f":
        fi = np.finfo(dtype)
        picks = [np.nan, np.inf, -np.inf, 0.0, -0.0, fi.tiny / 8, -fi.tiny / 3, fi.max, fi.tiny]
        # signaling NaNs (quiet bit clear), both signs
        u = np.uint32 if np.dtype(dtype).itemsize == 4 else np.uint64
        snan = [0x7fa00000, 0xff800001] if u is np.uint32 else [0x7ff4000000000000, 0xfff0000000000001]
        pool = np.concatenate([np.array(picks, dtype=dtype), np.array(snan, dtype=u).view(dtype)])
        idx = rng.choice(a.size, size=max(1, a.size // 10), replace=False)
        a[idx] = pool[rng.integers(0, pool.size, size=idx.size)]
        # a run of tiny values so whole blocks are tiny
        n = min(a.size, 64)
        a[:n] = (rng.normal(size=n) * float(fi.tiny) * 1e3).astype(dtype)
    else:
        ii = np.iinfo(dtype)
        idx = rng.choice(a.size, size=max(1, a.size // 10), replace=False)
        a[idx] = rng.choice(np.array([ii.min, ii.max, 0, -1, 1], dtype=dtype), size=idx.size)
    return a.reshape(shape)
