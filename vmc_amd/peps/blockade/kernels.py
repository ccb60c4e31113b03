"""Blockade-PEPS MC kernels (mirror of src/vmc/peps/blockade/kernels.py:17-117).

``transition`` runs the device sweep (row-pair flips, the last row, then
gradients, energies and the new bottom boundaries in the same launch
sequence) and ``estimate`` returns its results; the pair keeps the
reference's contract for ``make_mc_sampler``. Only the diagonal and one-site
terms of the operator enter the local energy, as in the reference's estimate
(model.py:313-330).
"""
from __future__ import annotations

import torch

from vmc_amd.operators.local_terms import LocalHamiltonian, to_tables
from vmc_amd.peps.blockade.compat import bl_plan_for
from vmc_amd.peps.compat import _amp_from_log
from vmc_amd.peps.kernels import Cache, ChainKeys, Context, LocalEstimates


def _flat(model, tensors):
    if tensors is None:
        return model.flat_params()
    if isinstance(tensors, torch.Tensor):
        return tensors.contiguous()
    return torch.cat([torch.as_tensor(t).reshape(-1) for row in tensors for t in row]).to(
        device=model.device, dtype=model.dtype).contiguous()


def build_mc_kernels(model, operator: LocalHamiltonian, *, full_gradient: bool = False):
    plan = bl_plan_for(model.shape, model.config, model.strategy, model.dtype, model.device, fresh=True)
    energy = bool(operator.terms)
    if energy:
        plan.set_terms(to_tables(operator, 2))
    n_sites = model.shape[0] * model.shape[1]

    def init_cache(tensors, config_states) -> Cache:
        s = torch.as_tensor(config_states).reshape(-1, n_sites).to(device=model.device, dtype=torch.int32)
        cache = plan.new_cache(s.shape[0])
        plan.init_cache(_flat(model, tensors), s.contiguous(), cache)
        return Cache(bottom_envs=cache)

    def transition(tensors, config_states, key: ChainKeys, cache: Cache):
        s = torch.as_tensor(config_states).reshape(-1, n_sites).to(
            device=model.device, dtype=torch.int32).clone().contiguous()
        la, o, p, e = plan.mc_sweep(_flat(model, tensors), s, cache.bottom_envs, key.seed, key.sweep, energy=energy)
        if e is None:
            e = torch.zeros((s.shape[0],), dtype=plan.energy_dtype, device=model.device)
        return s, key.next(), Context(amp=la, top_envs=(cache.bottom_envs, o, p, e))

    def estimate(tensors, config_next, context: Context):
        cache, o, p, e = context.top_envs
        if full_gradient:
            full = torch.zeros((o.shape[0], model.flat_params().numel()), dtype=o.dtype, device=o.device)
            off_full, off_sl = 0, 0
            for n, nsl in zip(model.params_per_site(), model.sliced_dims()):
                idx = p[:, off_sl].long()
                cols = off_full + idx[:, None] * n + torch.arange(n, device=o.device)[None, :]
                full.scatter_(1, cols, o[:, off_sl:off_sl + n])
                off_sl += n
                off_full += n * nsl
            o, p = full, None
        return Cache(bottom_envs=cache), LocalEstimates(
            local_log_derivatives=o, local_estimate=e, active_slice_indices=p,
            amp=_amp_from_log(context.amp, model.dtype.is_complex))

    return init_cache, transition, estimate
