"""gi_apply and the GI evaluator (mirror of src/vmc/peps/gi/compat.py:18-99)."""
from __future__ import annotations

import torch

from vmc_amd.engine import GIPlan
from vmc_amd.peps.compat import _amp_from_log
from vmc_amd.peps.strategy import device_strategy

_PLANS: dict = {}


def gi_plan_for(model, *, fresh=False) -> GIPlan:
    """Shared plan of the model's shapes; ``fresh=True`` (each build_mc_kernels
    bundle) returns a plan of the caller's own for its operator tables."""
    code = device_strategy(model.strategy)
    n_sweeps = getattr(model.strategy, "n_sweeps", 0)
    deg = tuple(int(d) for d in model.config.degeneracy_per_charge)
    key = (model.shape, model.N, model.phys_dim, model.dmax, model.Qx, model.charge_of_site, deg,
           model.strategy.truncate_bond_dimension, code, n_sweeps, model.dtype, str(model.device))
    plan = None if fresh else _PLANS.get(key)
    if plan is None:
        plan = GIPlan(model.shape, model.N, model.phys_dim, model.dmax, model.strategy.truncate_bond_dimension,
                      model.charge_of_site, Qx=model.Qx, strategy=code, dtype=model.dtype, device=model.device,
                      n_sweeps=n_sweeps, degeneracy_per_charge=deg)
        if not fresh:
            _PLANS[key] = plan
    return plan


def _flat(model, flat_params):
    return model.flat_params() if flat_params is None else flat_params.contiguous()


def gi_value_and_grad(model, samples, flat_params=None):
    """(amp, sliced dpsi/dA, combined slice indices p) for GI samples (n, n_sample).
    Samples violating Gauss law have amplitude 0 and zero gradient."""
    plan = gi_plan_for(model)
    s = torch.as_tensor(samples).reshape(-1, plan.n_sample)
    la, o, p = plan.amp_grad(_flat(model, flat_params), s, want_p=True)
    amp = _amp_from_log(la, model.dtype.is_complex)
    return amp, o.to(amp.dtype) * amp[:, None], p


class _GIApply(torch.autograd.Function):
    """Custom VJP of gi/compat.py:18-56 (holomorphic: conj(dpsi/dA) * g)."""

    @staticmethod
    def forward(ctx, flat_params, model, samples):
        plan = gi_plan_for(model)
        s = torch.as_tensor(samples).reshape(-1, plan.n_sample)
        la, o, p = plan.amp_grad(flat_params.detach().contiguous(), s, want_p=True)
        amp = _amp_from_log(la, flat_params.dtype.is_complex)
        ctx.model = model
        ctx.save_for_backward(o, amp, p)
        return amp

    @staticmethod
    def backward(ctx, g):
        o, amp, p = ctx.saved_tensors
        model = ctx.model
        dpsi = o.to(amp.dtype) * amp[:, None]
        if dpsi.is_complex():
            dpsi = dpsi.conj()
        sliced = dpsi * g[:, None]
        # scatter each site's slice gradient to its (sigma, cfg) slot of the full tensor
        full = torch.zeros(model.flat_params().numel(), dtype=o.dtype, device=o.device)
        off_full, off_sl = 0, 0
        for n, nsl in zip(model.params_per_site(), model.sliced_dims()):
            idx = p[:, off_sl].long()
            tgt = off_full + idx[:, None] * n + torch.arange(n, device=o.device)[None, :]
            full.index_add_(0, tgt.reshape(-1), sliced[:, off_sl:off_sl + n].reshape(-1).to(full.dtype))
            off_sl += n
            off_full += n * nsl
        return full, None, None


def gi_apply(model, flat_params, samples):
    """gi/compat.py:60-99: amplitudes of GI samples (0 where Gauss law fails),
    differentiable in flat_params."""
    return _GIApply.apply(_flat(model, flat_params), model, samples)
