"""Sample and Jacobian helpers (mirror of src/vmc/utils/vmc_utils.py:34-99).

``build_dense_jac`` is the dense centred Jacobian of log psi that the
NetKet-facing DenseSR consumes. For a PEPS or GI-PEPS it is the device's
full-layout log-derivative dpsi/dA / psi (the adjoint kernels, the same numbers
jax.jacrev of log psi gives, holomorphic); any other torch-differentiable
apply function goes through torch.func.jacrev.
"""
from __future__ import annotations

import torch

__all__ = ["flatten_samples", "batched_eval", "build_dense_jac", "ravel_params"]


def flatten_samples(samples):
    """vmc_utils.py:34-37: flatten every leading axis, keep the site axis."""
    s = torch.as_tensor(samples)
    return s.reshape(-1, s.shape[-1])


def batched_eval(eval_fn, samples, *, batch_size: int):
    """vmc_utils.py:40-59: eval_fn over fixed-size chunks, outputs concatenated
    and trimmed to the number of samples."""
    s = torch.as_tensor(samples)
    n = s.shape[0]
    outs = [eval_fn(s[i:i + batch_size]) for i in range(0, n, batch_size)]
    return torch.cat(outs, dim=0)[:n]


def ravel_params(params):
    """Flatten a parameter pytree (tensor, nested list/tuple, or dict in sorted
    key order like jax's ravel_pytree) into one vector; returns (flat, unravel)."""
    leaves, rebuild = _flatten(params)
    shapes = [t.shape for t in leaves]
    sizes = [t.numel() for t in leaves]
    flat = torch.cat([t.reshape(-1) for t in leaves]) if leaves else torch.zeros(0)

    def unravel(v):
        parts, off = [], 0
        for shape, n in zip(shapes, sizes):
            parts.append(v[off:off + n].reshape(shape))
            off += n
        return rebuild(parts)

    return flat, unravel


def _flatten(tree):
    if isinstance(tree, torch.Tensor):
        return [tree], lambda parts: parts[0]
    if isinstance(tree, dict):
        keys = sorted(tree)
        subs = [_flatten(tree[k]) for k in keys]
        counts = [len(l) for l, _ in subs]

        def rebuild(parts):
            out, off = {}, 0
            for k, (_, rb), c in zip(keys, subs, counts):
                out[k] = rb(parts[off:off + c])
                off += c
            return out

        return [x for l, _ in subs for x in l], rebuild
    if isinstance(tree, (list, tuple)):
        subs = [_flatten(t) for t in tree]
        counts = [len(l) for l, _ in subs]
        kind = type(tree)

        def rebuild(parts):
            out, off = [], 0
            for (_, rb), c in zip(subs, counts):
                out.append(rb(parts[off:off + c]))
                off += c
            return kind(out) if kind is tuple else out

        return [x for l, _ in subs for x in l], rebuild
    t = torch.as_tensor(tree)
    return [t], lambda parts: parts[0]


def _peps_of(apply_fun):
    from vmc_amd.peps.gi.model import GIPEPS
    from vmc_amd.peps.model import PEPS

    obj = getattr(apply_fun, "__self__", apply_fun)
    return obj if isinstance(obj, (PEPS, GIPEPS)) else None


def build_dense_jac(apply_fun, params, model_state, samples, *, holomorphic: bool = True):
    """vmc_utils.py:73-99: dense centred Jacobian d log psi / d params, one row per
    sample, columns in the parameter pytree's flat order; not rescaled by
    1 / sqrt(n_samples)."""
    samples = flatten_samples(samples)
    model = _peps_of(apply_fun)
    if model is not None:
        flat, _ = ravel_params(params) if params is not None else (model.flat_params(), None)
        flat = flat.to(device=model.device, dtype=model.dtype).contiguous()
        jac = _peps_log_jacobian(model, flat, samples)
    else:
        flat, unravel = ravel_params(params)

        def logpsi(v, x):
            return apply_fun({"params": unravel(v), **(model_state or {})}, x[None])[0]

        jac = torch.func.vmap(torch.func.jacrev(logpsi), in_dims=(None, 0))(flat, samples)
        if jac.is_complex() and not holomorphic:
            raise NotImplementedError("non-holomorphic Jacobians of complex outputs")
    return jac - jac.mean(dim=0, keepdim=True)


def _peps_log_jacobian(model, flat, samples):
    """Full-layout d log psi / dA of a PEPS or GI-PEPS from the device adjoint."""
    from vmc_amd.peps.gi.model import GIPEPS

    if isinstance(model, GIPEPS):
        from vmc_amd.peps.gi.compat import gi_plan_for

        plan = gi_plan_for(model)
        s = samples.to(device=model.device, dtype=torch.int32).reshape(-1, plan.n_sample)
        _, o, p = plan.amp_grad(flat, s, want_p=True)
        n = o.shape[0]
        full = torch.zeros((n, flat.numel()), dtype=o.dtype, device=o.device)
        off_full, off_sl = 0, 0
        for npar, nsl in zip(model.params_per_site(), model.sliced_dims()):
            idx = p[:, off_sl].long()
            cols = off_full + idx[:, None] * npar + torch.arange(npar, device=o.device)[None, :]
            full.scatter_(1, cols, o[:, off_sl:off_sl + npar])
            off_sl += npar
            off_full += npar * nsl
        return full
    from vmc_amd.peps.compat import _configs, plan_for, scatter_full

    cfg = _configs(model, samples)
    _, o, _ = plan_for(model).amp_grad(flat, cfg)
    return scatter_full(model, o, cfg)
