"""blockade_apply and the blockade evaluator (mirror of src/vmc/peps/blockade/compat.py:17-62)."""
from __future__ import annotations

import torch

from vmc_amd.engine import BLPlan
from vmc_amd.peps.compat import _amp_from_log
from vmc_amd.peps.strategy import device_strategy

_PLANS: dict = {}


def bl_plan_for(shape, config, strategy, dtype, device, *, fresh=False) -> BLPlan:
    """Shared plan of these shapes; ``fresh=True`` (each build_mc_kernels bundle)
    returns a plan of the caller's own for its operator tables."""
    code = device_strategy(strategy)
    n_sweeps = getattr(strategy, "n_sweeps", 0)
    key = (tuple(shape), int(config.D0), int(config.D1), strategy.truncate_bond_dimension, code, n_sweeps, dtype,
           str(device))
    plan = None if fresh else _PLANS.get(key)
    if plan is None:
        plan = BLPlan(shape, config.D0, config.D1, strategy.truncate_bond_dimension, strategy=code, dtype=dtype,
                      device=device, n_sweeps=n_sweeps)
        if not fresh:
            _PLANS[key] = plan
    return plan


def _flat(tensors, dtype, device):
    if isinstance(tensors, torch.Tensor):
        return tensors.contiguous()
    return torch.cat([torch.as_tensor(t).reshape(-1) for row in tensors for t in row]).to(
        device=device, dtype=dtype).contiguous()


class _BlockadeApply(torch.autograd.Function):
    """blockade_apply with its holomorphic VJP: conj(dpsi/dA) * g on the active
    (n, cfg) slice of every site, zero for invalid samples."""

    @staticmethod
    def forward(ctx, flat_params, plan, samples, ppsite, sdims):
        la, o, p = plan.amp_grad(flat_params.detach().contiguous(), samples, want_p=True)
        amp = _amp_from_log(la, flat_params.dtype.is_complex)
        ctx.meta = (ppsite, sdims, flat_params.numel())
        ctx.save_for_backward(o, amp, p)
        return amp

    @staticmethod
    def backward(ctx, g):
        o, amp, p = ctx.saved_tensors
        ppsite, sdims, nparams = ctx.meta
        dpsi = torch.nan_to_num(o.to(amp.dtype) * amp[:, None])
        if dpsi.is_complex():
            dpsi = dpsi.conj()
        sliced = dpsi * g[:, None]
        full = torch.zeros(nparams, dtype=o.dtype, device=o.device)
        off_full, off_sl = 0, 0
        for n, nsl in zip(ppsite, sdims):
            idx = p[:, off_sl].long()
            tgt = off_full + idx[:, None] * n + torch.arange(n, device=o.device)[None, :]
            full.index_add_(0, tgt.reshape(-1), sliced[:, off_sl:off_sl + n].reshape(-1).to(full.dtype))
            off_sl += n
            off_full += n * nsl
        return full, None, None, None, None


def _site_tables(shape):
    n_rows, n_cols = shape
    ppsite, sdims = [], []
    for r in range(n_rows):
        for c in range(n_cols):
            ppsite.append(1)
            sdims.append(2 ** (int(r > 0) + int(c > 0)) * 2)
    return ppsite, sdims


def blockade_apply(tensors, sample, shape, config, strategy, *, model=None):
    """compat.py:37-62: amplitudes of occupation samples (0 where two adjacent
    sites are occupied), differentiable in the flat parameters. ``tensors`` is
    the flat parameter vector or the nested site tensors."""
    dtype = model.dtype if model is not None else config.dtype
    device = model.device if model is not None else torch.device("cuda")
    flat = _flat(tensors, dtype, device)
    plan = bl_plan_for(shape, config, strategy, dtype, device)
    single = torch.as_tensor(sample).ndim == 1
    s = torch.as_tensor(sample).reshape(-1, plan.n_sample)
    n_rows, n_cols = shape
    dm = max(config.D0, config.D1)
    ppsite = [(dm if r > 0 else 1) * (dm if r < n_rows - 1 else 1) * (dm if c > 0 else 1) * (dm if c < n_cols - 1 else 1)
              for r in range(n_rows) for c in range(n_cols)]
    _, sdims = _site_tables(shape)
    amp = _BlockadeApply.apply(flat, plan, s, ppsite, sdims)
    return amp[0] if single else amp


def blockade_value_and_grad(model, samples, flat_params=None):
    """(amp, sliced dpsi/dA, combined slice indices p) of occupation samples;
    invalid samples have amplitude 0 and zero gradient."""
    plan = bl_plan_for(model.shape, model.config, model.strategy, model.dtype, model.device)
    flat = model.flat_params() if flat_params is None else flat_params.contiguous()
    s = torch.as_tensor(samples).reshape(-1, plan.n_sample)
    la, o, p = plan.amp_grad(flat, s, want_p=True)
    amp = _amp_from_log(la, model.dtype.is_complex)
    return amp, o.to(amp.dtype) * amp[:, None], p
