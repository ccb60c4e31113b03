"""SR preconditioner (mirror of src/vmc/preconditioners/preconditioners.py:41-330)."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from vmc_amd import _lib
from vmc_amd._lib import LIB, check
from vmc_amd.qgt import ParameterSpace, SampleSpace, SliceOrdering, SlicedJacobian, _vp, slices_arg, sr_call_args

__all__ = ["DirectSolve", "QRSolve", "SRPreconditioner", "solve_cg", "solve_cholesky", "solve_svd"]


def solve_cholesky(mat, rhs):
    """solvers.py:24-27. Runs on the device inside SRPreconditioner.apply
    (vmc_sr_solve_ex, VMC_SOLVE_CHOLESKY); pass it as DirectSolve(solver=...)."""
    raise NotImplementedError("device solvers run inside SRPreconditioner.apply (DirectSolve(solver=...))")


def solve_cg(mat, rhs, *, maxiter: int = 1000, tol: float = 1e-5):
    """solvers.py:16-21: conjugate gradients from 0 until |r| <= tol |b| or maxiter
    steps, on the device (VMC_SOLVE_CG); functools.partial keywords are honoured."""
    raise NotImplementedError("device solvers run inside SRPreconditioner.apply (DirectSolve(solver=...))")


def solve_svd(mat, rhs, *, rcond: float = 1e-12):
    """solvers.py:30-37: pseudo-inverse solve with cutoff rcond * max(S), on the
    device (VMC_SOLVE_SVD: the Cholesky solve when diag_shift certifies that no
    singular value falls under the cutoff, otherwise an error)."""
    raise NotImplementedError("device solvers run inside SRPreconditioner.apply (DirectSolve(solver=...))")


def _solver_args(solver):
    """(vmc_solver code, tol or rcond, maxiter) of a DirectSolve solver."""
    import functools
    import inspect

    kw = {}
    fn = solver
    if isinstance(solver, functools.partial):
        fn, kw = solver.func, dict(solver.keywords or {})
    if fn is solve_cholesky or fn == "cholesky":
        return _lib.VMC_SOLVE_CHOLESKY, 0.0, 0
    defaults = {k: v.default for k, v in inspect.signature(fn).parameters.items()
                if v.default is not inspect.Parameter.empty} if callable(fn) else {}
    if fn is solve_cg:
        a = {**defaults, **kw}
        return _lib.VMC_SOLVE_CG, float(a["tol"]), int(a["maxiter"])
    if fn is solve_svd:
        a = {**defaults, **kw}
        return _lib.VMC_SOLVE_SVD, float(a["rcond"]), 0
    raise NotImplementedError(f"{solver!r} has no device counterpart: use solve_cholesky, solve_cg or solve_svd")


@dataclass(frozen=True)
class DirectSolve:
    """preconditioners.py:41-45: dense solve of the shifted S or T on the device
    with solve_cholesky (default), solve_cg or solve_svd."""

    solver: object = solve_cholesky


@dataclass(frozen=True)
class QRSolve:
    """preconditioners.py:48-53: pivoted-QR (rank-revealing) solve on the device,
    ParameterSpace and full Jacobians only, diag_shift unused (as the reference)."""

    rcond: float | None = None
    min_norm: bool = True


def _params_complex(model, params):
    if params is not None and hasattr(params, "is_complex"):
        return bool(params.is_complex())
    dtype = getattr(model, "dtype", torch.float64)
    return bool(dtype.is_complex)


class SRPreconditioner:
    def __init__(self, space=ParameterSpace(), strategy=DirectSolve(), diag_shift: float = 1e-2,
                 gauge_config=None, ordering=SliceOrdering()):
        self.space = space
        self.strategy = strategy
        self.diag_shift = float(diag_shift)
        self.gauge_config = gauge_config
        self.ordering = ordering
        self.uses_local_energies = True
        self._metrics = {}

    @property
    def residual_error(self):
        return self._metrics.get("residual_error")

    def apply(self, model, params, samples, o, p, local_energies, *, grad_factor=1.0):
        """preconditioners.py:268-330 -> (site-major update vector, metrics)."""
        if self.gauge_config is not None:
            # preconditioners.py:288-292: the reference calls the projection, which raises
            from vmc_amd.gauge import compute_gauge_projection

            compute_gauge_projection(self.gauge_config, model, params, return_info=True)
        if isinstance(self.strategy, QRSolve):
            return self._apply_qr(model, params, o, p, local_energies, grad_factor)
        if p is None:
            # full Jacobian (Jacobian(o), preconditioners.py:296-297): one slice over
            # all parameters, so S = O_c^H O_c / Ns over the full columns
            o = o.contiguous()
            odt = {torch.float32: 0, torch.float64: 1, torch.complex64: 2, torch.complex128: 3}[o.dtype]
            ns, npl = o.shape
            p = torch.zeros((ns, npl), dtype=torch.int8, device=o.device)
            pps = torch.tensor([npl], dtype=torch.int32)
            d, dsite = 1, None
            npx = npl
        else:
            sd = (tuple(model.sliced_dims()) if hasattr(model, "sliced_dims")
                  else (model.phys_dim,) * (model.shape[0] * model.shape[1]))
            jac = SlicedJacobian(o, p, sd, self.ordering, tuple(model.params_per_site()))
            o, p, pps, odt = sr_call_args(jac)
            ns, npl = o.shape
            d, dsite = slices_arg(jac)
            npx = d * npl if dsite is None else int((dsite.long() * pps.long()).sum())
        solver, tol, maxiter = _solver_args(self.strategy.solver)
        space = 0 if isinstance(self.space, ParameterSpace) else 1
        gf = complex(grad_factor)
        e = local_energies.reshape(-1).to(device=o.device)
        e_cplx = e.is_complex()
        e = e.to(torch.complex128 if e_cplx else torch.float64).contiguous()
        cplx = o.is_complex() or e_cplx or gf.imag != 0.0
        # preconditioners.py:321-329 + tree_cast: real parameters take a real update
        real_params = not _params_complex(model, params)
        x_dtype = torch.float64 if real_params or not cplx else torch.complex128
        x = torch.empty(npx, dtype=x_dtype, device=o.device)
        metrics = torch.zeros(2, dtype=torch.float64, device=o.device)
        nbytes = int(LIB.vmc_sr_workspace_size(ns, npl, d, space, int(cplx)))
        ws = torch.empty(nbytes, dtype=torch.uint8, device=o.device)
        stream = torch.cuda.current_stream(o.device).cuda_stream
        check(LIB.vmc_sr_solve_ex(ns, npl, d, ctypes.c_void_p(pps.data_ptr()), len(pps), _vp(dsite),
                                  ctypes.c_void_p(o.data_ptr()), odt, ctypes.c_void_p(p.data_ptr()),
                                  ctypes.c_void_p(e.data_ptr()), int(e_cplx), space, ctypes.c_double(self.diag_shift),
                                  ctypes.c_double(gf.real), ctypes.c_double(gf.imag), int(real_params), solver,
                                  ctypes.c_double(tol), int(maxiter), ctypes.c_void_p(x.data_ptr()),
                                  ctypes.c_void_p(metrics.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
                                  ctypes.c_size_t(nbytes), ctypes.c_void_p(stream)))
        m = metrics.cpu()
        info = int(m[1].item())
        self._metrics = {"residual_error": float(m[0].item())}
        if solver == _lib.VMC_SOLVE_CG:
            self._metrics["cg_steps"] = info
        elif info != 0:
            raise RuntimeError(f"Cholesky failed at pivot {info}: matrix not positive definite")
        return x, self._metrics

    def _apply_qr(self, model, params, o, p, local_energies, grad_factor):
        """preconditioners.py:166-236 (QRSolve dispatches)."""
        if not isinstance(self.space, ParameterSpace):
            raise NotImplementedError("QRSolve not supported for SampleSpace")
        if p is not None:
            raise NotImplementedError("QRSolve not supported for SlicedJacobian")
        o = o.contiguous()
        ns, n = o.shape
        odt = {torch.float32: 0, torch.float64: 1, torch.complex64: 2, torch.complex128: 3}[o.dtype]
        gf = complex(grad_factor)
        e = local_energies.reshape(-1).to(device=o.device)
        e_cplx = e.is_complex()
        e = e.to(torch.complex128 if e_cplx else torch.float64).contiguous()
        cplx = o.is_complex() or e_cplx or gf.imag != 0.0
        real_params = not _params_complex(model, params)
        x = torch.empty(n, dtype=torch.float64 if real_params or not cplx else torch.complex128, device=o.device)
        metrics = torch.zeros(2, dtype=torch.float64, device=o.device)
        nbytes = int(LIB.vmc_sr_qr_workspace_size(ns, n, int(cplx)))
        ws = torch.empty(nbytes, dtype=torch.uint8, device=o.device)
        stream = torch.cuda.current_stream(o.device).cuda_stream
        rcond = -1.0 if self.strategy.rcond is None else float(self.strategy.rcond)
        check(LIB.vmc_sr_qr_solve(ns, n, ctypes.c_void_p(o.data_ptr()), odt, ctypes.c_void_p(e.data_ptr()),
                                  int(e_cplx), ctypes.c_double(gf.real), ctypes.c_double(gf.imag),
                                  ctypes.c_double(rcond), int(bool(self.strategy.min_norm)), int(real_params),
                                  ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(metrics.data_ptr()),
                                  ctypes.c_void_p(ws.data_ptr()), ctypes.c_size_t(nbytes), ctypes.c_void_p(stream)))
        m = metrics.cpu()
        self._metrics = {"residual_error": float(m[0].item()), "rank": int(m[1].item())}
        return x, self._metrics
