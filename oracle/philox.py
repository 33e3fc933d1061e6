"""Philox4x32-10 counter-based Gaussian stream (test infrastructure only).

The reference draws its sketches from ``cp.random.RandomState(seed)``
(trg/ATRG_3d_new.py:90-91, utility/randomized_svd.py:14,19), i.e. cuRAND,
which cannot be reproduced without CUDA.  The device library instead
generates sketches with Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
"Parallel random numbers: as easy as 1, 2, 3", SC'11) and a Box-Muller
transform; this module is the same generator in NumPy so that the oracle and
the GPU draw the identical sketch for the same (seed, stream).

Element e of a sketch uses counter (e_lo, e_hi, stream, 0) and key
(seed_lo, seed_hi); the first two output words give
u_i = (x_i + 0.5) * 2^-32 and z = sqrt(-2 ln u0) * cos(2 pi u1).
Pinned by the Random123 known-answer vectors (tests/test_oracle_units.py).
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32 with 10 rounds; all args uint32 arrays/scalars."""
    c0 = np.asarray(c0, dtype=np.uint32).copy()
    c1 = np.asarray(c1, dtype=np.uint32).copy()
    c2 = np.asarray(c2, dtype=np.uint32).copy()
    c3 = np.asarray(c3, dtype=np.uint32).copy()
    k0 = np.asarray(k0, dtype=np.uint32).copy()
    k1 = np.asarray(k1, dtype=np.uint32).copy()
    with np.errstate(over="ignore"):
        for _ in range(10):
            p0 = M0 * c0.astype(np.uint64)
            p1 = M1 * c2.astype(np.uint64)
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & _MASK).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & _MASK).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            k0 = (k0 + W0).astype(np.uint32)
            k1 = (k1 + W1).astype(np.uint32)
    return c0, c1, c2, c3


def gaussian(n, seed, stream=0, offset=0):
    """n standard normals (float64) of stream ``stream`` under ``seed``."""
    e = np.arange(offset, offset + n, dtype=np.uint64)
    lo = (e & _MASK).astype(np.uint32)
    hi = (e >> np.uint64(32)).astype(np.uint32)
    seed = int(seed)
    x0, x1, _, _ = philox4x32_10(lo, hi, np.uint32(stream & 0xFFFFFFFF), np.uint32(0),
                                 np.uint32(seed & 0xFFFFFFFF), np.uint32((seed >> 32) & 0xFFFFFFFF))
    u0 = (x0.astype(np.float64) + 0.5) * (1.0 / 4294967296.0)
    u1 = (x1.astype(np.float64) + 0.5) * (1.0 / 4294967296.0)
    return np.sqrt(-2.0 * np.log(u0)) * np.cos(2.0 * np.pi * u1)


def standard_normal(shape, seed, stream=0):
    n = int(np.prod(shape)) if len(shape) else 1
    return gaussian(n, seed, stream).reshape(shape)
