"""Seeded synthetic fields for benchmarks and tests.

grf(shape, seed): a Gaussian random field with power spectrum ~ k^-3
(spectral synthesis: white noise filtered by k^-3/2 in Fourier space), scaled
to unit variance, plus 1e-3 white noise.  Optionally 5 % of the 4^d blocks set
to zero and 1 % of values replaced by subnormals, as the large fixed-rate
workload specifies.  Generated in float64 (torch on the GPU when one is given,
numpy otherwise) and cast once to the requested dtype.
"""
import numpy as np


def _grf_numpy(shape, seed):
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    spec = np.fft.rfftn(noise)
    k2 = np.zeros(spec.shape)
    for ax, n in enumerate(shape):
        if ax == len(shape) - 1:
            k = np.fft.rfftfreq(n) * n
        else:
            k = np.fft.fftfreq(n) * n
        sh = [1] * len(shape)
        sh[ax] = k.size
        k2 = k2 + (k.reshape(sh)) ** 2
    k2[(0,) * len(shape)] = 1.0
    spec *= k2 ** (-3.0 / 4.0)  # amplitude ~ k^-3/2 -> power ~ k^-3
    spec[(0,) * len(shape)] = 0.0
    f = np.fft.irfftn(spec, s=shape)
    f /= f.std()
    f += 1e-3 * rng.standard_normal(shape)
    return f


def _grf_torch(shape, seed, device):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    noise = torch.randn(shape, generator=g, device=device, dtype=torch.float64)
    spec = torch.fft.rfftn(noise)
    k2 = torch.zeros(spec.shape, device=device, dtype=torch.float64)
    for ax, n in enumerate(shape):
        if ax == len(shape) - 1:
            k = torch.fft.rfftfreq(n, device=device, dtype=torch.float64) * n
        else:
            k = torch.fft.fftfreq(n, device=device, dtype=torch.float64) * n
        sh = [1] * len(shape)
        sh[ax] = k.numel()
        k2 = k2 + k.reshape(sh) ** 2
    del noise
    k2[(0,) * len(shape)] = 1.0
    spec *= k2.pow(-0.75)
    del k2
    spec[(0,) * len(shape)] = 0.0
    f = torch.fft.irfftn(spec, s=shape)
    del spec
    f /= f.std()
    f += 1e-3 * torch.randn(shape, generator=g, device=device, dtype=torch.float64)
    return f


def grf(shape, seed, dtype=np.float32, zero_blocks=0.0, subnormals=0.0, scale=1.0, device=None):
    """Return a numpy array (device=None) or a torch tensor on `device`."""
    shape = tuple(int(n) for n in shape)
    if device is None:
        f = _grf_numpy(shape, seed) * scale
        rng = np.random.default_rng(seed + 1)
        out = f.astype(dtype)
        if zero_blocks > 0:
            nb = [(n + 3) // 4 for n in shape]
            total = int(np.prod(nb))
            pick = rng.choice(total, size=int(total * zero_blocks), replace=False)
            coords = np.unravel_index(pick, nb)
            for c in zip(*coords):
                out[tuple(slice(4 * i, 4 * i + 4) for i in c)] = 0
        if subnormals > 0 and np.dtype(dtype).kind == "f":
            tiny = np.finfo(dtype).tiny
            idx = rng.choice(out.size, size=int(out.size * subnormals), replace=False)
            flat = out.reshape(-1)
            flat[idx] = (rng.uniform(-1, 1, size=idx.size) * tiny).astype(dtype)
        return out
    import torch
    f = _grf_torch(shape, seed, device)
    if scale != 1.0:
        f *= scale
    tdt = {np.float32: torch.float32, np.float64: torch.float64, np.int32: torch.int32,
           np.int64: torch.int64}[np.dtype(dtype).type]
    out = f.to(tdt)
    del f
    g = torch.Generator(device=device)
    g.manual_seed(seed + 1)
    if zero_blocks > 0:
        nb = [(n + 3) // 4 for n in shape]
        mask = torch.rand(nb, generator=g, device=device) < zero_blocks
        for ax in range(len(shape)):
            mask = mask.repeat_interleave(4, dim=ax)
        mask = mask[tuple(slice(0, n) for n in shape)]
        out[mask] = 0
        del mask
    if subnormals > 0 and out.is_floating_point():
        tiny = torch.finfo(out.dtype).tiny
        sel = torch.rand(shape, generator=g, device=device) < subnormals
        vals = (torch.rand(shape, generator=g, device=device, dtype=torch.float64) * 2 - 1) * tiny
        out[sel] = vals[sel].to(out.dtype)
        del sel, vals
    return out


def grf_tiles(seed, device, sizes=(1 << 20, 2048, 512)):
    """One periodic k^-3 tile per dimensionality (1-D, 2-D, 3-D), float64 on
    `device`: three FFTs in all, so that many arrays of arbitrary shape can be
    cut from them without a transform (and a runtime FFT kernel build) per
    shape."""
    return {d: _grf_torch((sizes[d - 1],) * d, seed + d, device) for d in (1, 2, 3)}


def grf_window(tiles, shape, seed, dtype=np.float32, scale=1.0):
    """An array of `shape` cut from the tile of its dimensionality at an
    origin drawn from `seed` (the tile is periodic, so windows wrap), scaled
    and cast to `dtype`.  Its spectrum is the tile's, ~k^-3."""
    import torch
    shape = tuple(int(n) for n in shape)
    t = tiles[len(shape)]
    rng = np.random.default_rng(seed)
    origin = [int(rng.integers(0, n)) for n in t.shape]
    idx = [torch.arange(o, o + n, device=t.device) % m for o, n, m in zip(origin, shape, t.shape)]
    w = t
    for ax, ix in enumerate(idx):
        w = w.index_select(ax, ix)
    sign = -1.0 if rng.integers(2) else 1.0
    tdt = {np.float32: torch.float32, np.float64: torch.float64, np.int32: torch.int32,
           np.int64: torch.int64}[np.dtype(dtype).type]
    return (w * (sign * scale)).to(tdt).contiguous()


# The mixed-array workloads of the batched and served forms: dimensionality
# uniform over 1-3, scalar type uniform over i32/i64/f32/f64, mode uniform over
# fixed rate 4-24, fixed precision 8-48 and fixed accuracy 1e-9..1e-2
# (log-uniform; integer arrays take full precision under fixed accuracy), each
# array cut from the k^-3 tiles with its own seed.
_MIX_TYPES = (("int32", np.int32, 2.0 ** 16), ("int64", np.int64, 2.0 ** 32),
              ("float", np.float32, 1.0), ("double", np.float64, 1.0))


def _mix_config(rng, ztype, dims):
    import zfp_hip as Z
    mode = int(rng.integers(3))
    if mode == 0:
        return Z.ZfpConfig.fixed_rate(float(rng.integers(4, 25)), ztype, dims), mode
    if mode == 1:
        return Z.ZfpConfig.fixed_precision(int(rng.integers(8, 49))), mode
    return Z.ZfpConfig.fixed_accuracy(10.0 ** rng.uniform(-9, -2)), mode


def batch_workload(n_arrays, seed=4096, device="cuda", note=None):
    """The batched workload: `n_arrays` arrays with each axis log-uniform in
    [3, 300] (many not multiples of 4).  Returns [(array, config, mode)] with
    mode 0 fixed rate, 1 fixed precision, 2 fixed accuracy."""
    import math

    import zfp_hip as Z
    ztypes = {"int32": Z.TYPE_INT32, "int64": Z.TYPE_INT64, "float": Z.TYPE_FLOAT,
              "double": Z.TYPE_DOUBLE}
    rng = np.random.default_rng(seed)
    tiles = grf_tiles(seed, device)
    out = []
    for i in range(n_arrays):
        dims = int(rng.integers(1, 4))
        shape = tuple(int(round(math.exp(rng.uniform(math.log(3), math.log(300))))) for _ in range(dims))
        name, npdt, scale = _MIX_TYPES[int(rng.integers(4))]
        cfg, mode = _mix_config(rng, ztypes[name], dims)
        out.append((grf_window(tiles, shape, seed + i, dtype=npdt, scale=scale), cfg, mode))
        if note and (i + 1) % 1024 == 0:
            note(f"{i + 1} arrays generated")
    return out


def serving_workload(n_requests, rate, seed=2000, device="cuda"):
    """The served workload: Poisson arrivals at `rate` requests/s, compress and
    decompress alternating (50/50), array sizes log-uniform from 16 KiB to
    256 MiB.  Returns [(kind, array, config, arrival_s)]."""
    import math

    import zfp_hip as Z
    ztypes = {"int32": Z.TYPE_INT32, "int64": Z.TYPE_INT64, "float": Z.TYPE_FLOAT,
              "double": Z.TYPE_DOUBLE}
    rng = np.random.default_rng(seed)
    tiles = grf_tiles(seed, device)
    out, t = [], 0.0
    for i in range(n_requests):
        dims = int(rng.integers(1, 4))
        name, npdt, scale = _MIX_TYPES[int(rng.integers(4))]
        esz = np.dtype(npdt).itemsize
        nbytes = math.exp(rng.uniform(math.log(16 << 10), math.log(256 << 20)))
        side = max(3, int(round((nbytes / esz) ** (1.0 / dims))))
        shape = tuple(max(3, int(side * rng.uniform(0.8, 1.25))) for _ in range(dims))
        cfg, _ = _mix_config(rng, ztypes[name], dims)
        x = grf_window(tiles, shape, seed + i, dtype=npdt, scale=scale)
        t += rng.exponential(1.0 / rate)
        out.append(("decompress" if i % 2 else "compress", x, cfg, t))
    return out
