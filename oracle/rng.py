"""Philox4x32-10 counter-based generator (oracle copy; test infrastructure only).

Published algorithm: J. Salmon, M. Moraes, R. Dror, D. Shaw, "Parallel random
numbers: as easy as 1, 2, 3", SC'11 -- the Philox4x32 bijection with 10 rounds,
multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments 0x9E3779B9 /
0xBB67AE85.

The reference draws its Metropolis uniforms and proposals with JAX threefry
(src/vmc/peps/standard/kernels.py:171-190). JAX is not available, so sample
sequences cannot be matched to the reference bit for bit; instead the GPU
sampler and this oracle share Philox with the same counter layout, so their
sample sequences can be compared draw for draw.

Counter layout of a sampler draw: ``(site, sweep, chain, 0)`` under key
``(seed_lo, seed_hi)``; ``u = ((x0 << 21) | (x1 >> 11)) * 2**-53`` is the
accept uniform, ``x2`` the proposal word (``delta = 1 + x2 % (d - 1)``).

Pinned by the published known-answer vectors of the Random123 distribution
(tests/test_oracle_rng.py).
"""
from __future__ import annotations

import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint64(0x9E3779B9)
_W1 = np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32(ctr, key, rounds: int = 10):
    """Philox4x32 on arrays: ``ctr`` (..., 4) uint32, ``key`` (..., 2) uint32."""
    c = [np.asarray(ctr[..., i], dtype=np.uint64) for i in range(4)]
    k0 = np.asarray(key[..., 0], dtype=np.uint64)
    k1 = np.asarray(key[..., 1], dtype=np.uint64)
    k0, k1 = np.broadcast_arrays(k0, k1)
    k0 = k0.copy()
    k1 = k1.copy()
    for r in range(rounds):
        if r > 0:
            k0 = (k0 + _W0) & _MASK
            k1 = (k1 + _W1) & _MASK
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
    return np.stack([x.astype(np.uint32) for x in c], axis=-1)


def seed_key(seed: int) -> np.ndarray:
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return np.array([seed & 0xFFFFFFFF, seed >> 32], dtype=np.uint32)


def sweep_draws(seed: int, sweep: int, site: int, chains: np.ndarray):
    """Accept uniforms (float64) and proposal words (uint32) for ``chains``."""
    chains = np.asarray(chains, dtype=np.uint32)
    ctr = np.zeros(chains.shape + (4,), dtype=np.uint32)
    ctr[..., 0] = np.uint32(site)
    ctr[..., 1] = np.uint32(sweep)
    ctr[..., 2] = chains
    x = philox4x32(ctr, seed_key(seed)[None, :])
    u53 = (x[..., 0].astype(np.uint64) << np.uint64(21)) | (
        x[..., 1].astype(np.uint64) >> np.uint64(11)
    )
    u = u53.astype(np.float64) * (2.0 ** -53)
    return u, x[..., 2]
