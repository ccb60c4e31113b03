"""Generic rollout builder (mirror of src/vmc/core/sampling.py:22-82)."""
from __future__ import annotations

import logging
import math

import torch

logger = logging.getLogger(__name__)


def make_mc_sampler(transition, estimate):
    """sampling.py:22-64: scan over sweeps of the batched transition/estimate pair."""

    def mc_sweep(tensors, config_state, key, cache):
        config_next, key_next, context = transition(tensors, config_state, key, cache)
        cache_next, local_estimates = estimate(tensors, config_next, context)
        return config_next, key_next, cache_next, local_estimates

    def mc_sampler(tensors, config_states, chain_keys, cache, *, n_steps: int):
        samples, estimates = [], []
        carry = (config_states, chain_keys, cache)
        for _ in range(n_steps):
            cfg, key, cache_c = carry
            cfg_next, key_next, cache_next, est = mc_sweep(tensors, cfg, key, cache_c)
            carry = (cfg_next, key_next, cache_next)
            samples.append(cfg_next)
            estimates.append(est)
        stacked = type(estimates[0])(*[
            (torch.stack(vals) if vals[0] is not None else None)
            for vals in zip(*estimates)
        ])
        return carry, (torch.stack(samples), stacked)

    return mc_sampler


def _sample_counts(n_samples: int, n_chains: int):
    """sampling.py:76-78."""
    chain_length = math.ceil(n_samples / n_chains)
    return n_samples, n_chains, chain_length, chain_length * n_chains


def _trim_samples(samples, total_samples: int, num_samples: int):
    """sampling.py:81-82: step-major flatten, keep the first num_samples."""
    return samples.reshape((total_samples,) + tuple(samples.shape[2:]))[:num_samples]
