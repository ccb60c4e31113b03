"""Pins the oracle's truncating strategies where the exact answer is known.

* A designed 2 x 3 PEPS (oracle.designed) whose row-0 thetas are monomial, so
  each cut keeps the chi largest entries: ZipUp (strategy.py:128-150) and
  DensityMatrix (strategy.py:153-176) must reproduce the closed-form amplitudes
  of their own keep rules through two truncating cuts. Zip-up with the
  singular values placed on the new tensor instead of the carry fails.
* The truncated env-gradient (energy.py:259-301) obeys a row identity: psi is
  linear in each site tensor of row r given the fixed top and bottom
  boundaries, so sum(A[s] * dpsi/dA[s]) is the same number psi_r for every site
  of row r, and psi_{R-1} is the top-sweep amplitude.
* DensityMatrix keeps min(chi, n) eigenvectors; where rank(theta) < k < n the
  kept null-space vectors are arbitrary. Rotating the null-space basis leaves
  the amplitude unchanged when chi <= D (no such cut) and changes it otherwise,
  which is why device parity for DensityMatrix is asserted at chi <= D only.
"""
import numpy as np
import pytest

from oracle import designed as DS
from oracle import peps as P


def _designed_amps(T, strat):
    return P.value(T, DS.all_configs(), DS.SHAPE, strat)


@pytest.mark.parametrize("chi", [2, 3])
@pytest.mark.parametrize("seed", range(8))
def test_designed_zipup_and_density_match_closed_form(seed, chi):
    T, w = DS.monomial_peps(seed)
    cfgs = DS.all_configs()
    for strat, rule in ((P.ZipUp(chi), "zipup"), (P.DensityMatrix(chi), "density")):
        amp = _designed_amps(T, strat)
        ref = DS.amplitudes(w, chi, rule, cfgs)
        assert np.max(np.abs(amp - ref)) <= 1e-12 * np.max(np.abs(ref)), (rule, amp, ref)


def test_designed_rules_differ():
    """The two keep rules disagree on some probed configuration for most seeds,
    so the closed-form comparison above distinguishes them."""
    cfgs = DS.all_configs()
    differ = 0
    for seed in range(8):
        for chi in (2, 3):
            _, w = DS.monomial_peps(seed)
            a = DS.amplitudes(w, chi, "zipup", cfgs)
            b = DS.amplitudes(w, chi, "density", cfgs)
            differ += int(np.max(np.abs(a - b)) > 1e-3)
    assert differ >= 9


class _ZipUpWrongSide:
    """Zip-up with S on the new tensor (new = U S, carry = V^H): a planted mistake."""

    def __init__(self, chi):
        self.truncate_bond_dimension = chi

    def apply(self, mps, mpo):
        new, carry = [], None
        n = len(mps)
        for i, (m, w) in enumerate(zip(mps, mpo)):
            theta = P.contract_theta(m, w, carry)
            z, left, q, Dr, wr = theta.shape
            if i == n - 1:
                new.append(theta.reshape(z, left, q, Dr * wr))
                break
            U, S, Vh = np.linalg.svd(theta.reshape(z, left * q, Dr * wr), full_matrices=False)
            k = min(self.truncate_bond_dimension, S.shape[-1])
            new.append((U[:, :, :k] * S[:, None, :k]).reshape(z, left, q, k))
            carry = Vh[:, :k, :].reshape(z, k, Dr, wr)
        return tuple(new)


def test_wrong_singular_value_side_is_detected():
    cfgs = DS.all_configs()
    caught = 0
    for seed in range(8):
        for chi in (2, 3):
            T, w = DS.monomial_peps(seed)
            ref = DS.amplitudes(w, chi, "zipup", cfgs)
            amp = _designed_amps(T, _ZipUpWrongSide(chi))
            caught += int(np.max(np.abs(amp - ref)) > 1e-6 * np.max(np.abs(ref)))
    assert caught >= 9


def _row_euler_sums(T, cfgs, shape, strat, D):
    amp, grad, p = P.value_and_grad(T, cfgs, shape, strat, D)
    spins = cfgs.reshape(-1, *shape)
    sums = np.zeros((len(cfgs), shape[0], shape[1]), dtype=grad.dtype)
    off = 0
    for r in range(shape[0]):
        for c in range(shape[1]):
            a = T[r][c][spins[:, r, c]].reshape(len(cfgs), -1)
            g = grad[:, off:off + a.shape[1]]
            sums[:, r, c] = np.sum(a * g, axis=1)
            off += a.shape[1]
    return amp, sums


@pytest.mark.parametrize("strat", [P.ZipUp(3), P.ZipUp(5), P.DensityMatrix(3)])
def test_truncated_env_gradient_row_identity(strat):
    shape, D = (4, 4), 3
    T = P.random_peps(shape, D, seed=11)
    cfgs = P.random_configs(16, 12, seed=11)
    amp, sums = _row_euler_sums(T, cfgs, shape, strat, D)
    ref = P.value(T, cfgs, shape, P.NoTruncation())
    assert np.max(np.abs(amp - ref)) > 1e-6 * np.max(np.abs(ref))  # chi truncates
    for r in range(shape[0]):
        row = sums[:, r, :]
        assert np.max(np.abs(row - row[:, :1])) <= 1e-10 * np.max(np.abs(row))
    assert np.max(np.abs(sums[:, -1, 0] - amp)) <= 1e-10 * np.max(np.abs(amp))


def test_row_identity_detects_swapped_gradient_blocks():
    shape, D = (3, 3), 3
    T = P.random_peps(shape, D, seed=12)
    cfgs = P.random_configs(9, 8, seed=12)
    amp, grad, _ = P.value_and_grad(T, cfgs, shape, P.ZipUp(3), D)
    spins = cfgs.reshape(-1, *shape)
    # row 1 sites 0 and 2 have equal shapes (D, D, 1, D): swap their blocks
    sizes = [T[r][c][0].size for r in range(3) for c in range(3)]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    g = grad.copy()
    i0, i2 = 3, 5
    g[:, offs[i0]:offs[i0 + 1]], g[:, offs[i2]:offs[i2 + 1]] = grad[:, offs[i2]:offs[i2 + 1]], grad[:, offs[i0]:offs[i0 + 1]]
    s = []
    for i in (i0, i2):
        r, c = divmod(i, 3)
        a = T[r][c][spins[:, r, c]].reshape(len(cfgs), -1)
        s.append(np.sum(a * g[:, offs[i]:offs[i + 1]], axis=1))
    assert np.max(np.abs(s[0] - s[1])) > 1e-6 * np.max(np.abs(amp))


class _DensityRotatedNull:
    """DensityMatrix with a random orthogonal rotation of the theta^H theta null space."""

    def __init__(self, chi, seed):
        self.truncate_bond_dimension = chi
        self.rng = np.random.default_rng(seed)

    def apply(self, mps, mpo):
        new, carry = [], None
        n = len(mps)
        for i, (m, w) in enumerate(zip(mps, mpo)):
            theta = P.contract_theta(m, w, carry)
            z, left, q, Dr, wr = theta.shape
            theta = theta.reshape(z, left * q, Dr * wr)
            if i == n - 1:
                new.append(theta.reshape(z, left, q, Dr * wr))
                break
            rho = np.conj(np.swapaxes(theta, 1, 2)) @ theta
            evals, evecs = np.linalg.eigh(rho)
            k = min(self.truncate_bond_dimension, rho.shape[-1])
            vk = np.empty((z, rho.shape[-1], k), dtype=theta.dtype)
            for b in range(z):
                ev = evals[b]
                order = np.argsort(ev)[::-1]
                vec = evecs[b][:, order]
                null = ev[order] <= 1e-12 * max(ev.max(), 1e-300)
                if null.sum() > 1:
                    Qr, _ = np.linalg.qr(self.rng.standard_normal((null.sum(), null.sum())))
                    vec[:, null] = vec[:, null] @ Qr
                vk[b] = vec[:, :k]
            new.append((theta @ vk).reshape(z, left, q, k))
            carry = np.conj(np.swapaxes(vk, 1, 2)).reshape(z, k, Dr, wr)
        return tuple(new)


def test_density_matrix_null_space_ambiguity():
    shape, D = (4, 4), 3
    T = P.random_peps(shape, D, seed=13)
    cfgs = P.random_configs(16, 8, seed=13)
    # chi <= D: no cut keeps null-space vectors whose choice a later cut sees
    a = P.value(T, cfgs, shape, P.DensityMatrix(3))
    b = P.value(T, cfgs, shape, _DensityRotatedNull(3, seed=1))
    assert np.max(np.abs(a - b)) <= 1e-10 * np.max(np.abs(a))
    # chi = 4 > D: row 1 column 0 has rank 3 < k = 4 < n = 9, and later cuts truncate
    a = P.value(T, cfgs, shape, P.DensityMatrix(4))
    b = P.value(T, cfgs, shape, _DensityRotatedNull(4, seed=1))
    assert np.max(np.abs(a - b)) > 1e-4 * np.max(np.abs(a))
