"""Pins the oracle's sequential-sampling sweep (src/vmc/peps/standard/kernels.py:131-245).

* empirical histogram vs exact |psi|^2 (recipe of tests/test_sequential_sampling.py:126-150);
* sampler outputs: p == config repeated per site, o == env-grad/amp of the
  evaluator at the sampled configuration, E_loc == local_estimate at the
  sampled configuration (transition/estimate cache turnover is exact when
  nothing is truncated).
"""
import numpy as np

from oracle import exact as X
from oracle import hamiltonians as HH
from oracle import peps as P


def _sample(T, shape, strat, terms, n_chains, n_steps, seed):
    cfg0 = P.random_configs(shape[0] * shape[1], n_chains, seed=seed + 17)
    cache = P.init_cache(T, cfg0, shape, strat)
    _, _, hist = P.mc_sampler(T, cfg0, cache, shape, strat, terms, seed, n_steps)
    return hist


def test_histogram_matches_exact_probabilities():
    shape = (2, 3)
    T = P.random_peps(shape, 2, seed=42)
    basis = X.all_configs(6)
    probs = np.abs(P.value(T, basis, shape, P.NoTruncation())) ** 2
    probs /= probs.sum()
    hist = _sample(T, shape, P.NoTruncation(), P.BucketedTerms(shape), 64, 640, seed=42)
    samples = hist["samples"].reshape(-1, 6)
    idx = samples @ (2 ** np.arange(5, -1, -1))
    emp = np.bincount(idx, minlength=64) / len(idx)
    assert np.max(np.abs(emp - probs)) < 4e-3


def test_sampler_outputs_match_evaluator():
    shape = (2, 3)
    D = 2
    T = P.random_peps(shape, D, seed=3)
    terms = HH.heisenberg(shape)
    hist = _sample(T, shape, P.ZipUp(8), terms, 8, 3, seed=3)
    samples = hist["samples"].reshape(-1, 6)
    amp, grad, p = P.value_and_grad(T, samples, shape, P.ZipUp(8), D)
    assert np.allclose(hist["amp"].reshape(-1), amp, rtol=1e-12, atol=0)
    assert np.allclose(hist["o"].reshape(len(samples), -1), grad / amp[:, None], rtol=1e-9, atol=1e-12)
    assert np.array_equal(hist["p"].reshape(len(samples), -1), p)
    e = P.local_estimate(T, samples, shape, P.ZipUp(8), terms, amp)
    assert np.allclose(hist["eloc"].reshape(-1), e, rtol=1e-9, atol=1e-12)
