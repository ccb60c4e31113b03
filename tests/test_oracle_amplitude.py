"""Pins the oracle's amplitudes (oracle.peps) with checks that do not reuse its code.

* NoTruncation (src/vmc/peps/common/strategy.py:118-125) == one-einsum contraction
  of the whole network (oracle.exact.full_network_amplitude);
* ZipUp / DensityMatrix / Variational with chi >= the exact boundary rank ==
  NoTruncation (recipe of tests/test_peps_strategy_equivalence.py:16-41);
* psi is multilinear in the site tensors and invariant under bond gauges
  A1 X, X^-1 A2 (Wu 2025, eqs. 7-8).
"""
import numpy as np
import pytest

from oracle import exact as X
from oracle import peps as P


def _amp(T, cfgs, shape, strat):
    return P.value(T, cfgs, shape, strat)


@pytest.mark.parametrize("shape,D,seed", [((3, 3), 3, 1), ((2, 3), 3, 2), ((2, 4), 2, 7)])
def test_no_truncation_matches_full_network(shape, D, seed):
    T = P.random_peps(shape, D, seed=seed)
    cfgs = X.all_configs(shape[0] * shape[1])
    amp = _amp(T, cfgs, shape, P.NoTruncation())
    ref = np.array([X.full_network_amplitude(T, c, shape) for c in cfgs])
    assert np.max(np.abs(amp - ref)) / np.max(np.abs(ref)) < 1e-12


def test_no_truncation_complex_matches_full_network():
    shape = (2, 3)
    T = P.random_peps(shape, 2, seed=4, dtype=np.complex128)
    cfgs = X.all_configs(6)
    amp = _amp(T, cfgs, shape, P.NoTruncation())
    ref = np.array([X.full_network_amplitude(T, c, shape) for c in cfgs])
    assert np.max(np.abs(amp - ref)) / np.max(np.abs(ref)) < 1e-12


def test_strategies_match_no_truncation():
    shape, D = (3, 3), 4
    T = P.random_peps(shape, D, seed=0)
    cfgs = P.random_configs(9, 64, seed=0)
    ref = _amp(T, cfgs, shape, P.NoTruncation())
    for strat in (P.ZipUp(16), P.DensityMatrix(16), P.Variational(16, n_sweeps=2)):
        amp = _amp(T, cfgs, shape, strat)
        assert np.max(np.abs(amp - ref)) / np.max(np.abs(ref)) < 1e-7, strat


def test_zipup_truncation_converges_with_chi():
    shape, D = (4, 4), 3
    T = P.random_peps(shape, D, seed=3)
    cfgs = P.random_configs(16, 16, seed=3)
    ref = _amp(T, cfgs, shape, P.NoTruncation())
    errs = [np.max(np.abs(_amp(T, cfgs, shape, P.ZipUp(chi)) - ref)) / np.max(np.abs(ref))
            for chi in (2, 4, 27)]
    assert errs[0] > 1e-6            # chi = 2 truncates
    assert errs[2] < 1e-11           # chi = D^3 keeps every singular value
    assert errs[1] < errs[0]


def test_multilinear_in_site_tensors():
    shape = (3, 3)
    T = P.random_peps(shape, 3, seed=5)
    cfgs = P.random_configs(9, 16, seed=5)
    base = _amp(T, cfgs, shape, P.ZipUp(9))
    T2 = [[t.copy() for t in row] for row in T]
    T2[1][2] *= 3.0
    assert np.allclose(_amp(T2, cfgs, shape, P.ZipUp(9)), 3.0 * base, rtol=1e-10, atol=0)
    T3 = [[0.5 * t for t in row] for row in T]
    assert np.allclose(_amp(T3, cfgs, shape, P.ZipUp(9)), 0.5 ** 9 * base, rtol=1e-10, atol=0)


def test_bond_gauge_invariance():
    shape, D = (3, 3), 3
    T = P.random_peps(shape, D, seed=6)
    cfgs = P.random_configs(9, 16, seed=6)
    base = _amp(T, cfgs, shape, P.NoTruncation())
    rng = np.random.default_rng(0)
    Xm = rng.standard_normal((D, D)) + 3 * np.eye(D)
    Xi = np.linalg.inv(Xm)
    T2 = [[t.copy() for t in row] for row in T]
    # horizontal bond (1,0)-(1,1): right index of A[1][0], left index of A[1][1]
    T2[1][0] = np.einsum("pudlr,rs->pudls", T2[1][0], Xm)
    T2[1][1] = np.einsum("sr,pudrx->pudsx", Xi, T2[1][1])
    # vertical bond (0,2)-(1,2): down index of A[0][2], up index of A[1][2]
    T2[0][2] = np.einsum("pudlr,ds->pusl r".replace(" ", ""), T2[0][2], Xm)
    T2[1][2] = np.einsum("su,pudlr->psdlr", Xi, T2[1][2])
    amp = _amp(T2, cfgs, shape, P.NoTruncation())
    assert np.max(np.abs(amp - base)) / np.max(np.abs(base)) < 1e-11


def test_contract_bottom_is_matrix_product():
    rng = np.random.default_rng(1)
    mats = [rng.standard_normal((2, 1, 1, 3)), rng.standard_normal((2, 3, 1, 4)), rng.standard_normal((2, 4, 1, 1))]
    ref = np.array([(mats[0][z, :, 0, :] @ mats[1][z, :, 0, :] @ mats[2][z, :, 0, :])[0, 0] for z in range(2)])
    assert np.allclose(P.contract_bottom(tuple(mats)), ref, rtol=1e-14)


def _mps_mpo_pair(seed, Dl=1, P_=3, Dmid=6, Dr=1, W=3, Q=3, complex_=False):
    """Two-site boundary (1, P, Dmid)(Dmid, P, 1) and row MPO (1, W, P, Q)(W, 1, P, Q), batch 1."""
    rng = np.random.default_rng(seed)

    def g(*s):
        x = rng.standard_normal(s)
        return x + 1j * rng.standard_normal(s) if complex_ else x

    mps = (g(1, Dl, P_, Dmid), g(1, Dmid, P_, Dr))
    mpo = (g(1, Dl, W, P_, Q), g(1, W, Dr, P_, Q))
    return mps, mpo


def _dense_state(tensors):
    """Contract an MPS (z, l, p, r) chain into the (left-most l) x ... vector of batch 0."""
    v = tensors[0][0]
    for t in tensors[1:]:
        v = np.tensordot(v, t[0], axes=([-1], [0]))
    return v.reshape(-1)


@pytest.mark.parametrize("complex_", [False, True])
def test_variational_two_sites_converges_to_best_rank_k(complex_):
    """One-site ALS on two sites alternates projections onto the row and column
    spaces, so many sweeps converge to the Eckart-Young optimum: the rank-k
    truncated SVD of the target MPO . MPS written as a Q x Q matrix."""
    Q = 6
    mps, mpo = _mps_mpo_pair(11, Q=Q, complex_=complex_)
    target = _dense_state(P.apply_mpo_exact(mps, mpo)).reshape(Q, Q)
    k = 2
    U, S, Vh = np.linalg.svd(target)
    assert S[k] > 1e-3 * S[0]  # the truncation discards weight
    best = (U[:, :k] * S[:k]) @ Vh[:k]
    res = P.apply_mpo_variational(mps, mpo, k, n_sweeps=300)
    got = _dense_state(res).reshape(Q, Q)
    assert np.linalg.norm(got - best) / np.linalg.norm(best) < 1e-9
    # two sweeps already approximate the target about as well as the optimum
    res2 = _dense_state(P.apply_mpo_variational(mps, mpo, k, n_sweeps=2)).reshape(Q, Q)
    assert np.linalg.norm(res2 - target) < 1.5 * np.linalg.norm(best - target)


def test_variational_independent_of_qr_phase_convention(monkeypatch):
    """Any QR differs from LAPACK's by a diagonal phase gauge D (Q D, D^H R);
    the compressed state, hence the amplitude, must not depend on it."""
    shape, D = (4, 4), 3
    T = P.random_peps(shape, D, seed=8, dtype=np.complex128)
    cfgs = P.random_configs(16, 8, seed=8)
    base = _amp(T, cfgs, shape, P.Variational(4, n_sweeps=2))
    exact = _amp(T, cfgs, shape, P.NoTruncation())
    assert np.max(np.abs(base - exact)) / np.max(np.abs(exact)) > 1e-6  # chi = 4 truncates
    orig = P.qr_reduced
    rng = np.random.default_rng(1)

    def phased(a):
        q, r = orig(a)
        ph = np.exp(1j * rng.uniform(0, 2 * np.pi, size=q.shape[:-2] + (q.shape[-1],)))
        return q * ph[..., None, :], np.conj(ph)[..., :, None] * r

    monkeypatch.setattr(P, "qr_reduced", phased)
    amp = _amp(T, cfgs, shape, P.Variational(4, n_sweeps=2))
    assert np.max(np.abs(amp - base)) / np.max(np.abs(base)) < 1e-10
