"""Rydberg Hamiltonian for the blockade PEPS (mirror of src/vmc/peps/blockade/hamiltonian.py:22-79)."""
from __future__ import annotations

import numpy as np

from vmc_amd.operators import DiagonalTerm, LocalHamiltonian, OneSiteTerm

__all__ = ["rydberg_hamiltonian"]


def rydberg_hamiltonian(shape, Omega: float, Delta: float, V_nnn: float = 0.0) -> LocalHamiltonian:
    """H = (Omega/2) sum_i sigma^x_i - Delta sum_i n_i + V_nnn sum_<<ij>> n_i n_j.
    Nearest-neighbour n_i n_j terms vanish under the blockade and are omitted;
    the next-nearest neighbours are the diagonal (r+1, c+1) and anti-diagonal
    (r+1, c-1) pairs."""
    n_rows, n_cols = shape
    sigma_x = np.array([[0.0, 1.0], [1.0, 0.0]])
    n_diag = np.array([0.0, -Delta])
    nn_diag = np.array([0.0, 0.0, 0.0, V_nnn])
    terms = []
    for r in range(n_rows):
        for c in range(n_cols):
            terms.append(OneSiteTerm(row=r, col=c, op=0.5 * Omega * sigma_x))
            terms.append(DiagonalTerm(sites=((r, c),), diag=n_diag))
    if V_nnn != 0.0:
        for r in range(n_rows):
            for c in range(n_cols):
                if r + 1 < n_rows and c + 1 < n_cols:
                    terms.append(DiagonalTerm(sites=((r, c), (r + 1, c + 1)), diag=nn_diag))
                if r + 1 < n_rows and c - 1 >= 0:
                    terms.append(DiagonalTerm(sites=((r, c), (r + 1, c - 1)), diag=nn_diag))
    return LocalHamiltonian(shape=tuple(shape), terms=tuple(terms))
