"""Hamiltonian term tables used by the workloads (test infrastructure only).

Conventions follow the reference:
* Heisenberg bond operator ``sz_sz - exchange`` with exchange 2 on the flip
  block (tests/test_local_estimate_exact.py:32-50, NetKet sign-rule form);
* transverse-field Ising: ``diag_zz = -J [1, -1, -1, 1]`` on nearest-neighbour
  pairs and ``op_x = -h sigma_x`` on every site
  (examples/ground_states/ising.py:30-47).
"""
from __future__ import annotations

import numpy as np

from oracle.peps import BucketedTerms

HEISENBERG_BOND = np.array(
    [[1, 0, 0, 0], [0, -1, -2, 0], [0, -2, -1, 0], [0, 0, 0, 1]], dtype=np.float64
)


def heisenberg(shape, J=1.0):
    n_rows, n_cols = shape
    t = BucketedTerms(shape)
    op = J * HEISENBERG_BOND
    for r in range(n_rows):
        for c in range(n_cols):
            if c + 1 < n_cols:
                t.horizontal[r][c].append(op)
            if r + 1 < n_rows:
                t.vertical[r][c].append(op)
    return t


def tfim(shape, J=1.0, h=3.04):
    n_rows, n_cols = shape
    t = BucketedTerms(shape)
    diag_zz = -J * np.array([1.0, -1.0, -1.0, 1.0])
    op_x = -h * np.array([[0.0, 1.0], [1.0, 0.0]])
    for r in range(n_rows):
        for c in range(n_cols):
            t.one_site[r][c].append(op_x)
            if c + 1 < n_cols:
                t.diagonal.append((((r, c), (r, c + 1)), diag_zz))
            if r + 1 < n_rows:
                t.diagonal.append((((r, c), (r + 1, c)), diag_zz))
    return t
