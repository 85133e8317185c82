"""Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
"Parallel random numbers: as easy as 1, 2, 3", SC'11), vectorised in numpy.

Used by the synthetic system generators so that every box is a pure function
of (seed, stream, index).  The same round function is implemented in the HIP
kernels for on-device move proposals.
"""
from __future__ import annotations

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr: np.ndarray, key: np.ndarray) -> np.ndarray:
    """ctr: (n, 4) uint32, key: (n, 2) or (2,) uint32 -> (n, 4) uint32."""
    c = np.atleast_2d(np.asarray(ctr, dtype=np.uint32))
    k = np.broadcast_to(np.asarray(key, dtype=np.uint32), (c.shape[0], 2))
    x0, x1, x2, x3 = (c[:, i].astype(np.uint64) for i in range(4))
    k0 = k[:, 0].astype(np.uint64)
    k1 = k[:, 1].astype(np.uint64)
    for r in range(10):
        p0 = M0 * x0
        p1 = M1 * x2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        x0, x1, x2, x3 = hi1 ^ x1 ^ k0, lo1, hi0 ^ x3 ^ k1, lo0
        k0 = (k0 + np.uint64(W0)) & _MASK
        k1 = (k1 + np.uint64(W1)) & _MASK
    return np.stack([x0, x1, x2, x3], axis=-1).astype(np.uint32)


def stream(seed: int, tag: int, n: int) -> np.ndarray:
    """n*4 uint32 words of stream `tag` under `seed` (counter = block index)."""
    nb = (n + 3) // 4
    ctr = np.zeros((nb, 4), dtype=np.uint32)
    ctr[:, 0] = np.arange(nb, dtype=np.uint64).astype(np.uint32)
    ctr[:, 1] = np.uint32(tag & 0xFFFFFFFF)
    key = np.array([seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF], dtype=np.uint32)
    return philox4x32_10(ctr, key).reshape(-1)


def uniform(seed: int, tag: int, n: int) -> np.ndarray:
    """n doubles in [0,1) with 53 random bits each (two words per value)."""
    w = stream(seed, tag, 2 * n)[: 2 * n].astype(np.uint64).reshape(n, 2)
    a = w[:, 0] >> np.uint64(5)
    b = w[:, 1] >> np.uint64(6)
    return (a.astype(np.float64) * 67108864.0 + b.astype(np.float64)) / 9007199254740992.0
