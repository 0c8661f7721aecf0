"""Range-noise oracle (test infrastructure): NumPy restatement of the noise models of
src/rmagine_core/src/noise/GaussianNoise.cpp:18-31, RelGaussianNoise.cpp:20-34,
UniformDustNoise.cpp:19-50 (and their CUDA kernels, src/rmagine_cuda/src/noise/*.cu), with
the library's counter-based random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
"Parallel random numbers: as easy as 1, 2, 3", SC 2011) of counter (element index lo, hi,
epoch, model kind) under key (seed, 0x5EED).

Pinned by: the Philox known-answer vectors published with Random123 (tests/test_noise.py),
closed-form moments of the noise distributions, and the dust probabilities.
Only tests import this module.
"""
from __future__ import annotations

import numpy as np

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32_10(c, k0, k1):
    """c: (..., 4) uint32 counters; k0, k1: scalars or arrays.  Returns (..., 4) uint32."""
    c = np.asarray(c, dtype=np.uint64)
    x, y, z, w = (c[..., i] for i in range(4))
    k0 = np.asarray(k0, dtype=np.uint64) & MASK
    k1 = np.asarray(k1, dtype=np.uint64) & MASK
    for _ in range(10):
        p0 = np.uint64(M0) * x
        p1 = np.uint64(M1) * z
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(MASK)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(MASK)
        x, y, z, w = (hi1 ^ y ^ k0) & np.uint64(MASK), lo1, (hi0 ^ w ^ k1) & np.uint64(MASK), lo0
        k0 = (k0 + np.uint64(W0)) & np.uint64(MASK)
        k1 = (k1 + np.uint64(W1)) & np.uint64(MASK)
    return np.stack([x, y, z, w], axis=-1).astype(np.uint32)


def _words(n, kind, seed, epoch):
    i = np.arange(n, dtype=np.uint64)
    c = np.stack([i & np.uint64(MASK), i >> np.uint64(32), np.full(n, epoch, np.uint64),
                  np.full(n, kind, np.uint64)], axis=-1)
    return philox4x32_10(c, seed, 0x5EED).astype(np.uint64)


def _open0(x):  # (0, 1]
    return ((x >> np.uint64(8)).astype(np.float64) + 1.0) * 2.0 ** -24


def _open1(x):  # [0, 1)
    return (x >> np.uint64(8)).astype(np.float64) * 2.0 ** -24


def apply(ranges, kind, p0, p1, p2=0.0, max_range=10000.0, seed=42, epoch=0):
    """Noisy copy of `ranges` (float32), as rmh_noise_apply computes it."""
    r = np.asarray(ranges, dtype=np.float32).copy()
    w = _words(len(r), kind, seed, epoch)
    f32 = lambda v: np.float64(np.float32(v))  # parameters cross the ABI as float
    if kind == 2:  # uniform dust
        d = np.minimum(r, np.float32(max_range)).astype(np.float64)
        p_hit = 1.0 - np.power(1.0 - f32(p0), d)
        hit = _open1(w[:, 0]) < p_hit
        nr = (_open1(w[:, 1]) * d).astype(np.float32)
        p_ret = np.power(f32(p1), nr.astype(np.float64))
        ret = hit & (_open1(w[:, 2]) < p_ret)
        r[ret] = nr[ret]
        return r
    ok = r <= np.float32(max_range)
    n = np.sqrt(-2.0 * np.log(_open0(w[:, 0]))) * np.cos(6.283185307179586 * _open1(w[:, 1]))
    sd = np.full(len(r), f32(p1))
    if kind == 1:
        sd = sd * np.power(r.astype(np.float64), f32(p2))
    out = (r.astype(np.float64) + n * sd + f32(p0)).astype(np.float32)
    r[ok] = out[ok]
    return r
