"""MC kernel bundle (mirror of src/vmc/peps/standard/kernels.py:32-247).

``build_mc_kernels(model, operator)`` returns ``(init_cache, transition,
estimate)`` with the reference's roles. They are batch-level (the reference's
per-chain functions only ever run under ``jax.vmap``): each call runs one gfx950
launch sequence over all chains.

Keys: the reference threads a JAX key per chain; here a :class:`ChainKeys`
(seed, sweep) selects the Philox4x32-10 counter block of a sweep, and
``transition`` returns the key advanced by one sweep.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, NamedTuple

import torch

from vmc_amd.operators.local_terms import LocalHamiltonian, to_tables
from vmc_amd.peps.compat import _amp_from_log, plan_for, scatter_full


class Cache(NamedTuple):
    """Persistent cache across sweeps: bottom boundaries of every row (device buffer)."""

    bottom_envs: Any


class Context(NamedTuple):
    """Transient transition output consumed by estimate(): log psi (B, 2) of the
    truncated top sweep, and the top boundaries of every row, held in a device
    workspace that this Context owns (with the cache the estimate rewrites)."""

    amp: torch.Tensor
    top_envs: Any


@dataclass(frozen=True)
class _TopEnvs:
    workspace: Any  # device buffer holding the top boundaries written by transition
    cache: Any


class LocalEstimates(NamedTuple):
    local_log_derivatives: torch.Tensor
    local_estimate: torch.Tensor
    active_slice_indices: torch.Tensor | None
    amp: torch.Tensor | None = None


@dataclass(frozen=True)
class ChainKeys:
    seed: int
    sweep: int = 0

    def next(self) -> "ChainKeys":
        return ChainKeys(self.seed, self.sweep + 1)


def _flat(model, tensors):
    if tensors is None:
        return model.flat_params()
    if isinstance(tensors, torch.Tensor):
        return tensors.contiguous()
    return torch.cat([torch.as_tensor(t).reshape(-1) for row in tensors for t in row]).to(
        device=model.device, dtype=model.dtype).contiguous()


def build_mc_kernels(model, operator: LocalHamiltonian, *, full_gradient: bool = False):
    """kernels.py:91-247; GIPEPS and BlockadePEPS models dispatch to their own
    kernels (the plum registrations of gi/kernels.py:17-23 and
    blockade/kernels.py:17-23)."""
    from vmc_amd.peps.blockade.model import BlockadePEPS
    from vmc_amd.peps.gi.model import GIPEPS

    if isinstance(model, GIPEPS):
        from vmc_amd.peps.gi.kernels import build_mc_kernels as gi_build

        return gi_build(model, operator, full_gradient=full_gradient)
    if isinstance(model, BlockadePEPS):
        from vmc_amd.peps.blockade.kernels import build_mc_kernels as bl_build

        return bl_build(model, operator, full_gradient=full_gradient)
    # a plan of this bundle's own: its operator tables are not shared with any
    # other bundle or local_estimate call
    plan = plan_for(model, fresh=True, terms=to_tables(operator, model.phys_dim) if operator.terms else None)
    n_sites = model.shape[0] * model.shape[1]

    def init_cache(tensors, config_states) -> Cache:
        cfg = torch.as_tensor(config_states).reshape(-1, n_sites).to(device=model.device, dtype=torch.int32)
        cache = plan.new_cache(cfg.shape[0])
        plan.init_cache(_flat(model, tensors), cfg.contiguous(), cache)
        return Cache(bottom_envs=cache)

    def transition(tensors, config_states, key: ChainKeys, cache: Cache):
        cfg = torch.as_tensor(config_states).reshape(-1, n_sites).to(
            device=model.device, dtype=torch.int32).clone().contiguous()
        ws = plan.new_sweep_workspace(cfg.shape[0])
        log_amp = plan.transition(_flat(model, tensors), cfg, cache.bottom_envs, key.seed, key.sweep, ws=ws)
        return cfg, key.next(), Context(amp=log_amp, top_envs=_TopEnvs(ws, cache.bottom_envs))

    def estimate(tensors, config_next, context: Context):
        cache = context.top_envs.cache
        o, p, eloc = plan.estimate(_flat(model, tensors), config_next, cache, context.amp,
                                   energy=bool(operator.terms), ws=context.top_envs.workspace)
        if eloc is None:
            eloc = torch.zeros((config_next.shape[0],), dtype=plan.energy_dtype, device=model.device)
        if full_gradient:
            o = scatter_full(model, o, config_next)
            p = None
        # estimate() rewrites the cache buffer in place with the new bottom
        # boundaries (the reference returns a new Cache).
        return Cache(bottom_envs=cache), LocalEstimates(
            local_log_derivatives=o, local_estimate=eloc, active_slice_indices=p,
            amp=_amp_from_log(context.amp, model.dtype.is_complex))

    return init_cache, transition, estimate
