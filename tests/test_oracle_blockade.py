"""Pins of the blockade-PEPS oracle (oracle/blockade.py), CPU only.

* Per-slice charge masks equal _assemble_site's per-sample masks on valid
  samples (D0 != D1, real and complex), and entries a mask removes do not
  change any amplitude.
* Invalid samples (two adjacent occupied sites) have amplitude and gradient 0.
* Gradients: central finite differences of the exact amplitude, scattered
  through the combined slice indices, equal the sliced env-gradients.
* Local energies: the Rydberg E_loc of the estimate equals <x|H|psi>/psi(x)
  with H built densely on the valid (independent-set) basis, for every basis
  state, on 2x3 and 3x3 lattices (two-row windows, the last row's one-row
  windows and the kL/kU rewiring of the flipped site's neighbours all enter).
* The sampler visits only valid states and its histogram matches |psi|^2.
"""
import itertools

import numpy as np
import pytest

from oracle import blockade as B
from oracle import peps as P


def _all_valid(shape):
    n = shape[0] * shape[1]
    occ = np.array(list(itertools.product((0, 1), repeat=n)), np.int64).reshape(-1, *shape)
    return occ[B.valid(occ)]


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_slice_masks_equal_per_sample_masks(dtype):
    cfg = B.BLConfig((3, 3), 3, 1)
    T = B.random_tensors(cfg, seed=4, dtype=dtype)
    occ = _all_valid((3, 3))
    strat = P.ZipUp(81)
    flat = B.value_and_grad(cfg, T, occ, strat)[0]
    lit = B.literal_amplitude(cfg, T, occ, strat)
    assert np.max(np.abs(flat - lit)) <= 1e-13 * np.max(np.abs(lit))
    # entries a mask removes (bond index >= D1 on the outgoing bonds of an
    # occupied site) never matter; unmasked ones do
    T2 = [[t.copy() for t in row] for row in T]
    for r in range(3):
        for c in range(3):
            t = T2[r][c]
            if t.shape[3] > 1:
                t[1, 0, :, 1:, :, :] += 5.0
            if t.shape[5] > 1:
                t[1, 0, :, :, :, 1:] += 3.0
    assert np.max(np.abs(B.literal_amplitude(cfg, T2, occ, strat) - lit)) <= 1e-13 * np.max(np.abs(lit))
    assert np.max(np.abs(B.value_and_grad(cfg, T2, occ, strat)[0] - lit)) <= 1e-13 * np.max(np.abs(lit))
    T3 = [[t.copy() for t in row] for row in T]
    T3[1][1][1, 0, 1, 0, 0, 0] += 1.0  # up bond index 1 < D0 of an occupied site: kept
    assert np.max(np.abs(B.literal_amplitude(cfg, T3, occ, strat) - lit)) > 1e-6 * np.max(np.abs(lit))
    mask = cfg.mask()
    assert mask is not None


def test_invalid_samples_have_zero_amplitude_and_gradient():
    cfg = B.BLConfig((2, 3), 2, 2)
    T = B.random_tensors(cfg, seed=1)
    occ = np.array([[[1, 1, 0], [0, 0, 0]], [[1, 0, 0], [1, 0, 0]], [[1, 0, 1], [0, 1, 0]]], np.int64)
    amp, grad, _ = B.value_and_grad(cfg, T, occ, P.NoTruncation())
    assert list(B.valid(occ)) == [False, False, True]
    assert amp[0] == 0 and amp[1] == 0 and amp[2] != 0
    assert not grad[:2].any()


def test_gradient_matches_finite_differences():
    cfg = B.BLConfig((2, 3), 2, 2)
    T = B.random_tensors(cfg, seed=2)
    occ = _all_valid((2, 3))[[1, 5, 9, 12]]
    strat = P.NoTruncation()
    amp, grad, p = B.value_and_grad(cfg, T, occ, strat)
    assert np.array_equal(p, B.slices(cfg, occ))  # the reference's index = the slice on valid samples
    eps = 1e-6
    off = 0
    for r in range(2):
        for c in range(3):
            t = T[r][c]
            n_per = int(np.prod(t.shape[2:]))
            for k in range(0, n_per, 3):
                for z in range(occ.shape[0]):
                    sl = p[z, r * 3 + c]
                    idx = np.unravel_index(sl * n_per + k, t.shape)
                    Tp = [[x.copy() for x in row] for row in T]
                    Tm = [[x.copy() for x in row] for row in T]
                    Tp[r][c][idx] += eps
                    Tm[r][c][idx] -= eps
                    fd = (B.value_and_grad(cfg, Tp, occ[z:z + 1], strat)[0][0]
                          - B.value_and_grad(cfg, Tm, occ[z:z + 1], strat)[0][0]) / (2 * eps)
                    assert abs(fd - grad[z, off + k]) <= 1e-6 * max(1.0, abs(fd))
            off += n_per


def _dense_h(shape, basis, omega, delta, vnnn):
    """H of hamiltonian.py on the independent-set basis (rows: |x'>, columns: |x>)."""
    n_rows, n_cols = shape
    index = {tuple(b.ravel()): i for i, b in enumerate(basis)}
    H = np.zeros((len(basis), len(basis)))
    for j, b in enumerate(basis):
        occ = b.ravel()
        e = -delta * occ.sum()
        for r in range(n_rows):
            for c in range(n_cols):
                for rr, cc in ((r + 1, c + 1), (r + 1, c - 1)):
                    if 0 <= rr < n_rows and 0 <= cc < n_cols:
                        e += vnnn * b[r, c] * b[rr, cc]
        H[j, j] = e
        for s in range(occ.size):
            f = occ.copy()
            f[s] = 1 - f[s]
            i = index.get(tuple(f))
            if i is not None:
                H[i, j] += 0.5 * omega
    return H


@pytest.mark.parametrize("shape,D0,D1", [((2, 3), 2, 2), ((3, 3), 2, 1)])
def test_rydberg_eloc_matches_dense_hamiltonian(shape, D0, D1):
    cfg = B.BLConfig(shape, D0, D1)
    T = B.random_tensors(cfg, seed=6)
    basis = _all_valid(shape)
    strat = P.NoTruncation()
    omega, delta, vnnn = 1.3, 0.7, 0.45
    amp, _, _ = B.value_and_grad(cfg, T, basis, strat)
    H = _dense_h(shape, basis, omega, delta, vnnn)
    ref = (H.T @ amp) / amp  # sum_x' H[x', x] psi(x') / psi(x): the op[:, s] convention
    terms = B.rydberg_terms(shape, omega, delta, vnnn)
    z = basis.shape[0]
    cache = B.init_cache(cfg, T, basis, strat)
    occ, amp_t, tops = B.transition(cfg, T, basis, cache, strat, seed=0, sweep=0, chains=np.arange(z))
    # estimate on the unchanged basis: reuse the top boundaries of a forward pass
    ft = B.flat_tensors(cfg, T)
    amp_f, tops = P.forward_with_cache(ft, B.slices(cfg, basis).reshape(z, *shape), strat)
    _, _, e, _ = B.estimate(cfg, T, basis, amp_f, tops, strat, terms)
    assert np.max(np.abs(e - ref)) <= 1e-9 * np.max(np.abs(ref))
    assert B.valid(occ).all()


def test_sampler_histogram_matches_amplitudes():
    shape = (2, 3)
    cfg = B.BLConfig(shape, 2, 2)
    T = B.random_tensors(cfg, seed=8)
    basis = _all_valid(shape)
    strat = P.NoTruncation()
    amp = B.value_and_grad(cfg, T, basis, strat)[0]
    prob = np.abs(amp) ** 2 / np.sum(np.abs(amp) ** 2)
    terms = B.rydberg_terms(shape, 1.0, 0.5)
    occ0 = B.random_independent_sets(shape, 2048, seed=3)
    _, _, hist = B.mc_sampler(cfg, T, occ0, strat, terms, seed=11, n_steps=12)
    samples = np.concatenate(hist["samples"][2:], axis=0).reshape(-1, *shape)
    assert B.valid(samples).all()
    keys = {tuple(b.ravel()): i for i, b in enumerate(basis)}
    counts = np.zeros(len(basis))
    for s in samples.reshape(samples.shape[0], -1):
        counts[keys[tuple(s)]] += 1
    assert np.max(np.abs(counts / counts.sum() - prob)) < 0.02
