"""Synthetic PEPS inputs for benchmarks and examples.

Site tensors of shape (d, up, down, left, right) with i.i.d. N(0,1)/sqrt(D^4 d)
entries from PCG64 ``numpy.random.default_rng(seed)``, drawn site by site in
row-major order (a complex tensor draws its real part, then its imaginary
part); configurations are uniform occupancies from ``default_rng(seed + 1)``.
Lower precisions are cast from the FP64 draw so every precision sees the same
values.
"""
from __future__ import annotations

import numpy as np


def site_dims(r, c, R, C, D):
    """Bond dims (up, down, left, right) of an OBC site (src/vmc/peps/standard/model.py:27-35)."""
    return (1 if r == 0 else D, 1 if r == R - 1 else D, 1 if c == 0 else D, 1 if c == C - 1 else D)


def random_site_tensors(shape, bond_dim, phys_dim=2, seed=0, complex_=False):
    rng = np.random.default_rng(seed)
    R, C = shape
    scale = 1.0 / np.sqrt(float(bond_dim) ** 4 * phys_dim)
    out = []
    for r in range(R):
        row = []
        for c in range(C):
            dims = (phys_dim, *site_dims(r, c, R, C, bond_dim))
            t = rng.standard_normal(dims) * scale
            if complex_:
                t = t + 1j * (rng.standard_normal(dims) * scale)
            row.append(t)
        out.append(row)
    return out


def flatten(tensors):
    """Parameter vector in the reference's ravel_pytree order (row-major sites)."""
    return np.concatenate([np.ravel(t) for row in tensors for t in row])


def random_configs(n_sites, n_samples, seed=0, phys_dim=2):
    rng = np.random.default_rng(seed + 1)
    return rng.integers(0, phys_dim, size=(n_samples, n_sites)).astype(np.int32)
