"""Seeded synthetic inputs for the KODAMA.matrix workloads (row-major float32).

Every draw comes from a counter-based SplitMix64 stream turned into uniforms
(53-bit) and normals (Box-Muller in double, then cast to float32), so a dataset
is identical on any host and any numpy version.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _splitmix(seed: int, stream: int, count: int) -> np.ndarray:
    """Outputs count..: z_i = mix(seed ^ stream-key + (i + 1) * golden)."""
    base = np.uint64((seed * 0x100000001B3 + stream * 0xC2B2AE3D27D4EB4F) & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        z = base + (np.arange(1, count + 1, dtype=np.uint64) * _GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, stream: int, count: int) -> np.ndarray:
    return (_splitmix(seed, stream, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, count: int) -> np.ndarray:
    u = uniform(seed, stream, 2 * count)
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def randint(seed: int, stream: int, count: int, lo: int, hi: int) -> np.ndarray:
    """Integers uniform in [lo, hi] inclusive."""
    return lo + np.minimum((uniform(seed, stream, count) * (hi - lo + 1)).astype(np.int64), hi - lo)


def gaussian_mixture(n: int, p: int, centers: int, seed: int, spread: float = 3.0,
                     zipf_s: float | None = None, sigma: float = 1.0):
    """Isotropic mixture: means U[-spread, spread]^p, sd sigma; equal weights or
    Zipf(s) cluster sizes. Returns (x float32 [n, p], component int32 [n], 1-based)."""
    means = (uniform(seed, 1, centers * p) * 2.0 - 1.0).reshape(centers, p) * spread
    if zipf_s is None:
        w = np.full(centers, 1.0 / centers)
    else:
        w = 1.0 / np.arange(1, centers + 1, dtype=np.float64) ** zipf_s
        w /= w.sum()
    cum = np.cumsum(w)
    comp = np.searchsorted(cum, uniform(seed, 2, n) * cum[-1], side="right")
    comp = np.minimum(comp, centers - 1)
    x = means[comp] + sigma * normal(seed, 3, n * p).reshape(n, p)
    return x.astype(np.float32), (comp + 1).astype(np.int32)


def config1(n: int = 2000, p: int = 50, seed: int = 1):
    """5-component mixture, means U[-3,3]^50, unit variance, equal weights; initial
    labels uniform in {1..5} from seed 2."""
    x, comp = gaussian_mixture(n, p, 5, seed, spread=3.0)
    labels = randint(2, 1, n, 1, 5).astype(np.int32)
    return x, labels, comp


def config2(n: int = 10000, p: int = 50, seed: int = 3):
    """Latent Gaussian with per-dimension variance 1/i plus 8 clusters shifted by
    2 sigma along 3 random directions."""
    sd = 1.0 / np.sqrt(np.arange(1, p + 1, dtype=np.float64))
    base = normal(seed, 1, n * p).reshape(n, p) * sd
    comp = np.minimum((uniform(seed, 2, n) * 8).astype(np.int64), 7)
    dirs = normal(seed, 3, 3 * p).reshape(3, p)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    signs = (uniform(seed, 4, 8 * 3).reshape(8, 3) * 2.0 - 1.0)
    shift = 2.0 * (signs @ dirs) * sd
    x = base + shift[comp]
    return x.astype(np.float32), (comp + 1).astype(np.int32)


def config3(n: int = 100000, p: int = 50, seed: int = 4):
    """20-cluster Zipf(1.1) mixture, sigma 1, means U[-4,4]^50."""
    return gaussian_mixture(n, p, 20, seed, spread=4.0, zipf_s=1.1)


def config4(n: int = 1000000, p: int = 50, seed: int = 5):
    return gaussian_mixture(n, p, 20, seed, spread=4.0, zipf_s=1.1)


def config5(n: int = 200000, p: int = 50, seed: int = 8):
    """30-cluster Zipf mixture with 5% label noise on the initial labels."""
    x, comp = gaussian_mixture(n, p, 30, seed, spread=4.0, zipf_s=1.1)
    noisy = uniform(seed, 9, n) < 0.05
    labels = comp.copy()
    labels[noisy] = randint(seed, 10, int(noisy.sum()), 1, 30)
    return x, labels, comp
