"""Sliced Jacobian and quantum geometric tensor (mirror of src/vmc/qgt)."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from vmc_amd._lib import LIB, check

__all__ = ["SliceOrdering", "SiteOrdering", "Jacobian", "SlicedJacobian", "QGT", "ParameterSpace",
           "SampleSpace", "jacobian_mean"]


@dataclass(frozen=True)
class SliceOrdering:
    """jacobian.py:21-33 (k outer). The device builds site-major columns; the
    result equals SliceOrdering followed by _reorder_updates."""


@dataclass(frozen=True)
class SiteOrdering:
    """jacobian.py:36-46 (sites outer)."""

    params_per_site: tuple


@dataclass(frozen=True)
class ParameterSpace:
    """O^H O formulation (qgt.py:162-166)."""


@dataclass(frozen=True)
class SampleSpace:
    """O O^H formulation (qgt.py:169-173)."""


@dataclass
class Jacobian:
    O: torch.Tensor


@dataclass
class SlicedJacobian:
    """jacobian.py:56-93: o (Ns, sum pps), p (Ns, sum pps) int8, sliced dims per site."""

    o: torch.Tensor
    p: torch.Tensor
    sliced_dims: tuple
    ordering: object = SliceOrdering()
    params_per_site: tuple = ()

    @classmethod
    def from_samples(cls, model, samples, ordering=SliceOrdering()):
        from vmc_amd.peps.compat import plan_for, _configs

        if hasattr(model, "sliced_dims"):  # GIPEPS: samples carry links
            from vmc_amd.peps.gi.compat import gi_plan_for

            plan = gi_plan_for(model)
            _, o, p = plan.amp_grad(model.flat_params(), torch.as_tensor(samples).reshape(-1, plan.n_sample),
                                    want_p=True)
            return cls(o, p, tuple(model.sliced_dims()), ordering, tuple(model.params_per_site()))
        plan = plan_for(model)
        _, o, p = plan.amp_grad(model.flat_params(), _configs(model, samples), want_p=True)
        return cls(o, p, (model.phys_dim,) * (model.shape[0] * model.shape[1]), ordering,
                   tuple(model.params_per_site()))


def jacobian_mean(jac):
    """jacobian.py:96-140 (site-major column order)."""
    if isinstance(jac, Jacobian):
        return jac.O.mean(dim=0)
    d = max(jac.sliced_dims)
    parts = []
    off = 0
    for n in jac.params_per_site:
        o = jac.o[:, off:off + n].double()
        p = jac.p[:, off:off + n]
        for k in range(d):
            parts.append(torch.where(p == k, o, torch.zeros_like(o)).mean(dim=0))
        off += n
    return torch.cat(parts)


_O_DTYPES = {torch.float32: 0, torch.float64: 1, torch.complex64: 2, torch.complex128: 3}


def slices_arg(jac):
    """(max slice count, int32 per-site counts or None when uniform)."""
    sd = tuple(int(x) for x in jac.sliced_dims)
    if len(set(sd)) <= 1:
        return (sd[0] if sd else 1), None
    return max(sd), torch.tensor(sd, dtype=torch.int32)


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def sr_call_args(jac):
    """(o, p, params_per_site, vmc_dtype of o) for the SR entry points."""
    pps = torch.tensor(list(jac.params_per_site), dtype=torch.int32)
    o = jac.o.contiguous()
    p = jac.p.contiguous()
    if o.dtype not in _O_DTYPES:
        raise TypeError(f"log-derivatives of dtype {o.dtype} are not supported")
    return o, p, pps, _O_DTYPES[o.dtype]


@dataclass
class QGT:
    """qgt.py:176-189: lazy QGT; to_dense runs on the device (vmc_qgt_dense)."""

    jac: SlicedJacobian
    space: object = ParameterSpace()

    def to_dense(self):
        jac = self.jac
        if isinstance(jac, Jacobian):
            # a full Jacobian is one slice over all columns (p = 0)
            O = jac.O.contiguous()
            jac = SlicedJacobian(O, torch.zeros(O.shape, dtype=torch.int8, device=O.device), (1,),
                                 params_per_site=(O.shape[1],))
        o, p, pps, odt = sr_call_args(jac)
        ns, npl = o.shape
        d, dsite = slices_arg(jac)
        space = 0 if isinstance(self.space, ParameterSpace) else 1
        npx = d * npl if dsite is None else int((dsite.long() * pps.long()).sum())
        n = npx if space == 0 else ns
        cplx = o.is_complex()
        out = torch.empty((n, n), dtype=torch.complex128 if cplx else torch.float64, device=o.device)
        nbytes = int(LIB.vmc_sr_workspace_size(ns, npl, d, space, int(cplx)))
        ws = torch.empty(nbytes, dtype=torch.uint8, device=o.device)
        stream = torch.cuda.current_stream(o.device).cuda_stream
        check(LIB.vmc_qgt_dense(ns, npl, d, ctypes.c_void_p(pps.data_ptr()), len(pps), _vp(dsite),
                                ctypes.c_void_p(o.data_ptr()), odt, ctypes.c_void_p(p.data_ptr()), space,
                                ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
                                ctypes.c_size_t(nbytes), ctypes.c_void_p(stream)))
        return out

    def __matmul__(self, v):
        dense = self.to_dense()
        return dense @ v.to(dense.dtype)
