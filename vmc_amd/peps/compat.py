"""Evaluator surface (mirror of src/vmc/peps/standard/compat.py:31-194).

Every function runs the gfx950 engine through the C ABI (vmc_amp_grad); there is
no CPU fallback. Amplitudes are produced as log psi on the device because at
10x10 |psi| ~ 1e-85; ``_value``/``_value_and_grad`` exponentiate in FP64 to
return the reference's raw amplitudes and dpsi/dA.
"""
from __future__ import annotations

import math

import torch

from vmc_amd.engine import AmpGradPlan
from vmc_amd.peps.strategy import device_strategy

_PLANS: dict = {}


def plan_for(model, dtype=None, *, fresh=False, terms=None) -> AmpGradPlan:
    """Device plan of the model's shapes and strategy.

    Plans without terms are shared per shape (they hold no caller state). A plan
    with operator terms is shared only by callers with the same tables (keyed by
    their bytes), and ``fresh=True`` returns a plan of the caller's own, as each
    ``build_mc_kernels`` bundle uses, so no call replaces another's Hamiltonian.
    """
    dtype = dtype or model.dtype
    code = device_strategy(model.strategy)
    n_sweeps = getattr(model.strategy, "n_sweeps", 0)
    key = (model.shape, model.bond_dim, model.phys_dim, model.strategy.truncate_bond_dimension, code, n_sweeps,
           dtype, str(model.device))

    def make():
        return AmpGradPlan(model.shape, model.bond_dim, model.strategy.truncate_bond_dimension,
                           phys_dim=model.phys_dim, strategy=code, dtype=dtype, device=model.device,
                           n_sweeps=n_sweeps)

    if fresh:
        plan = make()
        if terms is not None:
            plan.set_terms(terms)
        return plan
    if terms is not None:
        key = key + (_tables_digest(terms),)
    plan = _PLANS.get(key)
    if plan is None:
        plan = make()
        if terms is not None:
            plan.set_terms(terms)
        _PLANS[key] = plan
    return plan


def _tables_digest(tables):
    import hashlib

    import numpy as np

    h = hashlib.sha1()
    for k in sorted(tables):
        v = tables[k]
        h.update(k.encode())
        if v is not None:
            a = np.ascontiguousarray(np.asarray(v))
            h.update(str(a.dtype).encode())
            h.update(a.tobytes())
    return h.hexdigest()


def spin_to_occupancy(spins):
    """src/vmc/utils/utils.py:55-57 (idempotent on {0,1})."""
    return torch.div(torch.as_tensor(spins) + 1, 2, rounding_mode="floor").to(torch.int32)


def _configs(model, sample):
    s = torch.as_tensor(sample)
    if s.ndim == 1:
        s = s[None]
    s = s.reshape(s.shape[0], -1)
    return spin_to_occupancy(s).to(model.device)


def _log_complex(log_amp):
    return torch.complex(log_amp[:, 0], log_amp[:, 1])


def log_value(model, samples):
    """log psi (complex128) for a batch: what PEPS.__call__ returns (model.py:82-87)."""
    plan = plan_for(model)
    la, _, _ = plan.amp_grad(model.flat_params(), _configs(model, samples), grad=False)
    return _log_complex(la)


def _amp_from_log(la, complex_=False):
    mag = torch.exp(la[:, 0])
    if complex_:
        return torch.polar(mag, la[:, 1])
    return torch.where(la[:, 1] != 0, -mag, mag)


def _value(model, sample):
    """compat.py:67-78: raw amplitudes (FP64)."""
    single = torch.as_tensor(sample).ndim == 1
    plan = plan_for(model)
    la, _, _ = plan.amp_grad(model.flat_params(), _configs(model, sample), grad=False)
    amp = _amp_from_log(la, model.dtype.is_complex)
    return amp[0] if single else amp


def _value_and_grad(model, sample, *, full_gradient=False):
    """compat.py:94-133: (amp, dpsi/dA, p); dpsi/dA is not divided by amp."""
    single = torch.as_tensor(sample).ndim == 1
    cfg = _configs(model, sample)
    plan = plan_for(model)
    la, o, p = plan.amp_grad(model.flat_params(), cfg, want_p=True)
    amp = _amp_from_log(la, model.dtype.is_complex)
    grad = o.to(amp.dtype) * amp[:, None]
    if full_gradient:
        grad = scatter_full(model, grad, cfg)
        p = None
    if single:
        return amp[0], grad[0], (None if p is None else p[0])
    return amp, grad, p


def _grad(model, sample, *, full_gradient=False):
    """compat.py:81-91."""
    _, g, p = _value_and_grad(model, sample, full_gradient=full_gradient)
    return g, p


def scatter_full(model, sliced, cfg):
    """Scatter sliced per-site blocks into the (d, ...) full-parameter layout."""
    B = sliced.shape[0]
    d = model.phys_dim
    parts = []
    off = 0
    for site, n in enumerate(model.params_per_site()):
        full = torch.zeros((B, d, n), dtype=sliced.dtype, device=sliced.device)
        full[torch.arange(B, device=sliced.device), cfg[:, site].long()] = sliced[:, off:off + n]
        parts.append(full.reshape(B, -1))
        off += n
    return torch.cat(parts, dim=1)


def local_estimate(model, samples, operator, amps=None):
    """compat.py:136-194: local energies of given samples (device top sweep +
    bottom-up energy pass; the denominator is the device's own top-sweep psi)."""
    from vmc_amd.operators.local_terms import to_tables

    plan = plan_for(model, terms=to_tables(operator, model.phys_dim))
    _, eloc = plan.local_estimate(model.flat_params(), _configs(model, samples))
    return eloc


class _PepsApply(torch.autograd.Function):
    """Custom VJP of peps_apply (compat.py:31-64): grad[r][c][s_rc] = g * env_grad."""

    @staticmethod
    def forward(ctx, flat_params, configs, model):
        plan = plan_for(model, flat_params.dtype)
        la, o, _ = plan.amp_grad(flat_params.detach().contiguous(), configs)
        amp = _amp_from_log(la, flat_params.dtype.is_complex)
        ctx.save_for_backward(o, amp, configs)
        ctx.model = model
        return amp

    @staticmethod
    def backward(ctx, g):
        o, amp, configs = ctx.saved_tensors
        # holomorphic psi: torch's backward takes conj(dpsi/dA) * g
        dpsi = o.to(amp.dtype) * amp[:, None]
        if dpsi.is_complex():
            dpsi = dpsi.conj()
        sliced = dpsi * g[:, None]
        full = scatter_full(ctx.model, sliced, configs).sum(dim=0)
        return full.to(o.dtype), None, None


def peps_apply(model, flat_params, samples):
    """Amplitudes differentiable w.r.t. the flat parameter vector (torch.autograd)."""
    return _PepsApply.apply(flat_params, _configs(model, samples), model)
