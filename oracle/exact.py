"""Independent brute-force checks that pin the oracle (test infrastructure only).

* ``full_network_amplitude``: contracts the whole PEPS network for one
  configuration with a single ``np.einsum`` over all bonds (no boundary MPS,
  no row order) -- the value ``_apply_mpo_exact`` (src/vmc/peps/common/strategy.py:118-125)
  must reproduce.
* ``dense_hamiltonian``: builds H from bucketed terms by Kronecker products, with
  ``H[s, p] = <s|O|p>`` embedded at the term's sites (the NetKet dense-operator
  recipe of tests/test_local_estimate_exact.py:24-79).
* ``all_configs``: the 2^N (or d^N) basis in the reference's row-major order.
"""
from __future__ import annotations

import itertools
import string

import numpy as np


def full_network_amplitude(tensors, config, shape):
    """psi(config) by one einsum over the full network (tiny lattices only)."""
    n_rows, n_cols = shape
    letters = iter(string.ascii_letters)
    h = {}  # horizontal bond (r, c)-(r, c+1)
    v = {}  # vertical bond (r, c)-(r+1, c)
    for r in range(n_rows):
        for c in range(n_cols):
            if c + 1 < n_cols:
                h[(r, c)] = next(letters)
            if r + 1 < n_rows:
                v[(r, c)] = next(letters)
    operands, subs = [], []
    spins = np.asarray(config).reshape(shape)
    for r in range(n_rows):
        for c in range(n_cols):
            t = tensors[r][c][spins[r, c]]  # (up, down, left, right)
            idx = ""
            for (dim, lab) in (
                (0, v.get((r - 1, c))),
                (1, v.get((r, c))),
                (2, h.get((r, c - 1))),
                (3, h.get((r, c))),
            ):
                if lab is None:
                    t = t.take(0, axis=len(idx))
                else:
                    idx += lab
            operands.append(t)
            subs.append(idx)
    expr = ",".join(subs) + "->"
    return np.einsum(expr, *operands, optimize="greedy")


def all_configs(n_sites, phys_dim=2):
    """All configurations, site 0 most significant (row-major lattice order)."""
    return np.array(list(itertools.product(range(phys_dim), repeat=n_sites)), dtype=np.int32)


def _embed(op, sites_flat, n_sites, phys_dim):
    """Dense operator acting with ``op`` (<out|op|in>) on the given flat sites."""
    dim = phys_dim ** n_sites
    basis = all_configs(n_sites, phys_dim)
    H = np.zeros((dim, dim), dtype=np.result_type(op.dtype, np.float64))
    k = len(sites_flat)
    weights = phys_dim ** np.arange(n_sites - 1, -1, -1)
    for col, cfg in enumerate(basis):
        local_in = 0
        for s in sites_flat:
            local_in = local_in * phys_dim + cfg[s]
        for local_out in range(phys_dim ** k):
            amp = op[local_out, local_in]
            if amp == 0:
                continue
            new = cfg.copy()
            rem = local_out
            for s in reversed(sites_flat):
                new[s] = rem % phys_dim
                rem //= phys_dim
            H[int(new @ weights), col] += amp
    return H


def dense_hamiltonian(terms, shape, phys_dim=2):
    """Dense H (d^N x d^N) from an ``oracle.peps.BucketedTerms``."""
    n_rows, n_cols = shape
    n = n_rows * n_cols
    dim = phys_dim ** n
    H = np.zeros((dim, dim), dtype=np.complex128)
    flat = lambda r, c: r * n_cols + c
    for sites, diag in terms.diagonal:
        op = np.diag(np.asarray(diag, dtype=np.complex128))
        H += _embed(op, [flat(r, c) for r, c in sites], n, phys_dim)
    for r in range(n_rows):
        for c in range(n_cols):
            for op in terms.one_site[r][c]:
                H += _embed(np.asarray(op), [flat(r, c)], n, phys_dim)
            if c + 1 < n_cols:
                for op in terms.horizontal[r][c]:
                    H += _embed(np.asarray(op), [flat(r, c), flat(r, c + 1)], n, phys_dim)
            if r + 1 < n_rows:
                for op in terms.vertical[r][c]:
                    H += _embed(np.asarray(op), [flat(r, c), flat(r + 1, c)], n, phys_dim)
    return H
