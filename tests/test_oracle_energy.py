"""Pins the oracle's local energies (src/vmc/peps/common/energy.py:123-224).

E_loc(s) == (H psi)(s) / psi(s) with H built densely by Kronecker embedding
(recipe of tests/test_local_estimate_exact.py:24-79, tol 1e-9). The reference
applies off-diagonal operators as op[:, s] (energy.py:88, 179, 200), i.e. the
transpose of <s|O|p>; a non-symmetric operator pins that convention.
"""
import numpy as np
import pytest

from oracle import exact as X
from oracle import hamiltonians as HH
from oracle import peps as P


def _check(terms, shape, D, seed, strat=None, dense_terms=None):
    T = P.random_peps(shape, D, seed=seed)
    cfgs = X.all_configs(shape[0] * shape[1])
    strat = strat or P.NoTruncation()
    amp = P.value(T, cfgs, shape, strat)
    e = P.local_estimate(T, cfgs, shape, strat, terms, amp)
    H = X.dense_hamiltonian(dense_terms or terms, shape)
    ref = (H @ amp) / amp
    assert np.max(np.abs(e - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("shape", [(2, 3), (3, 2), (2, 2)])
def test_heisenberg_matches_dense(shape):
    _check(HH.heisenberg(shape), shape, 3, 1)


@pytest.mark.parametrize("shape", [(2, 3), (3, 3)])
def test_tfim_matches_dense(shape):
    _check(HH.tfim(shape, J=1.0, h=3.04), shape, 2, 2)


def test_horizontal_only_and_vertical_only():
    shape = (2, 3)
    h = HH.heisenberg(shape)
    hor = P.BucketedTerms(shape, horizontal=h.horizontal)
    ver = P.BucketedTerms(shape, vertical=h.vertical)
    _check(hor, shape, 3, 3)
    _check(ver, shape, 3, 4)


def test_transposed_operator_convention():
    shape = (2, 2)
    rng = np.random.default_rng(9)
    op1 = rng.standard_normal((2, 2))
    op2 = rng.standard_normal((4, 4))
    terms = P.BucketedTerms(shape)
    terms.one_site[0][1].append(op1)
    terms.horizontal[1][0].append(op2)
    terms.vertical[0][0].append(op2)
    dense = P.BucketedTerms(shape)
    dense.one_site[0][1].append(op1.T)
    dense.horizontal[1][0].append(op2.T)
    dense.vertical[0][0].append(op2.T)
    _check(terms, shape, 2, 5, dense_terms=dense)


def test_zipup_energy_with_exact_chi():
    shape = (3, 3)
    _check(HH.heisenberg(shape), shape, 2, 6, strat=P.ZipUp(8))


def test_complex_operators_match_dense():
    """Complex, non-Hermitian one-site / two-site / diagonal terms on a complex
    PEPS (local_terms.py:22-200 accept any dtype): E_loc == (H^T-convention psi)/psi."""
    shape, D = (2, 3), 2
    T = P.random_peps(shape, D, seed=8, dtype=np.complex128)
    cfgs = X.all_configs(6)
    rng = np.random.default_rng(10)
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    op1, op2, op3, dg = c(2, 2), c(4, 4), c(4, 4), c(4)
    terms, dense = P.BucketedTerms(shape), P.BucketedTerms(shape)
    terms.one_site[1][2].append(op1)
    terms.horizontal[0][1].append(op2)
    terms.vertical[0][0].append(op3)
    terms.diagonal.append((((1, 0), (1, 1)), dg))
    dense.one_site[1][2].append(op1.T)
    dense.horizontal[0][1].append(op2.T)
    dense.vertical[0][0].append(op3.T)
    dense.diagonal.append((((1, 0), (1, 1)), dg))
    amp = P.value(T, cfgs, shape, P.NoTruncation())
    e = P.local_estimate(T, cfgs, shape, P.NoTruncation(), terms, amp)
    H = X.dense_hamiltonian(dense, shape)
    ref = (H @ amp) / amp
    assert np.iscomplexobj(e) and np.max(np.abs(ref.imag)) > 1e-3
    assert np.max(np.abs(e - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))


def test_complex_tables_layout():
    """to_tables keeps complex values (complex128 arrays, term order) for vmc_plan_set_terms_c."""
    from vmc_amd.operators import HorizontalTwoSiteTerm, LocalHamiltonian, OneSiteTerm
    from vmc_amd.operators.local_terms import to_tables

    op1 = np.array([[0, 1j], [-1j, 0]])
    op2 = np.eye(4) * (1 + 2j)
    t = to_tables(LocalHamiltonian(shape=(2, 2), terms=(OneSiteTerm(0, 1, op1), HorizontalTwoSiteTerm(1, 0, op2))), 2)
    assert t["one_ops"].dtype == np.complex128 and t["h_ops"].dtype == np.complex128
    assert np.array_equal(t["one_sites"], [1]) and np.array_equal(t["h_sites"], [2])
    assert np.array_equal(t["one_ops"], op1.ravel()) and np.array_equal(t["h_ops"], op2.ravel())
    real = to_tables(LocalHamiltonian(shape=(2, 2), terms=(OneSiteTerm(0, 1, op1.real),)), 2)
    assert real["one_ops"].dtype == np.float64
