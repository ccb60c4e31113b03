"""NetKet-facing QGT adapters (mirror of src/vmc/qgt/netket_compat.py:21-78).

NetKet is not a dependency here: QGTOperator keeps the LinearOperator surface
the reference subclasses (``@``, ``solve`` / ``_solve``, ``to_dense``) and
DenseSR keeps the AbstractLinearPreconditioner surface (``lhs_constructor``,
``__call__(vstate, gradient)``) over any variational-state object that exposes
``samples``, ``parameters``, ``model_state`` and ``_apply_fun`` like
nk.vqs.MCState. The QGT is the device one (vmc_qgt_dense for the dense S).
"""
from __future__ import annotations

from typing import Any, Callable

import torch

from vmc_amd.qgt import QGT, Jacobian, ParameterSpace
from vmc_amd.utils.vmc_utils import build_dense_jac, flatten_samples, ravel_params

__all__ = ["QGTOperator", "DenseSR", "cg"]


def cg(A, b, x0=None, *, tol: float = 1e-5, atol: float = 0.0, maxiter: int | None = None):
    """Conjugate gradients with the stopping rule of jax.scipy.sparse.linalg.cg
    (|r|^2 <= max(tol^2 |b|^2, atol^2), at most maxiter = 10 n steps) over
    A @ v; returns (x, None) like the reference's solver."""
    n = b.numel()
    maxiter = 10 * n if maxiter is None else int(maxiter)
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    r = b - (A @ x) if x0 is not None else b.clone()
    p = r.clone()
    rs = torch.vdot(r, r).real
    thresh = max(tol * tol * float(torch.vdot(b, b).real), atol * atol)
    k = 0
    while float(rs) > thresh and k < maxiter:
        ap = A @ p
        alpha = rs / torch.vdot(p, ap).real
        x = x + alpha * p
        r = r - alpha * ap
        rs_new = torch.vdot(r, r).real
        p = r + (rs_new / rs) * p
        rs = rs_new
        k += 1
    return x, None


class QGTOperator:
    """netket_compat.py:21-46: the QGT (ParameterSpace) plus diag_shift as a linear
    operator over parameter pytrees."""

    def __init__(self, _qgt: QGT | None = None, diag_shift: float = 0.0, _params_structure: Any = None):
        self._qgt = _qgt
        self.diag_shift = float(diag_shift)
        self._params_structure = _params_structure
        self._dense = None

    def _S(self):
        if self._dense is None:
            self._dense = self._qgt.to_dense()
        return self._dense

    def __matmul__(self, v):
        flat, unravel = ravel_params(v)
        S = self._S()
        flat = flat.to(device=S.device, dtype=S.dtype)
        return unravel(S @ flat + self.diag_shift * flat)

    def _solve(self, solve_fun: Callable, y, *, x0=None):
        y_flat, unravel = ravel_params(y)
        S = self._S()
        y_flat = y_flat.to(device=S.device, dtype=S.dtype)
        if x0 is not None:
            x0 = ravel_params(x0)[0].to(device=S.device, dtype=S.dtype)
        out, info = solve_fun(self, y_flat, x0=x0)
        return unravel(out), info

    def solve(self, solve_fun: Callable, y, *, x0=None):
        return self._solve(solve_fun, y, x0=x0)

    def to_dense(self):
        S = self._S()
        return S + self.diag_shift * torch.eye(S.shape[0], dtype=S.dtype, device=S.device)


class DenseSR:
    """netket_compat.py:49-78: SR preconditioner whose left-hand side is the
    device QGT of the dense centred Jacobian of the variational state."""

    def __init__(self, diag_shift: float = 1e-2, solver: Callable = cg, holomorphic: bool = True):
        self.diag_shift = float(diag_shift)
        self.solver = solver
        self.holomorphic = bool(holomorphic)
        self._lhs = None
        self.x0 = None
        self.info = None

    def lhs_constructor(self, vstate, step=None) -> QGTOperator:
        samples = flatten_samples(vstate.samples)
        O = build_dense_jac(vstate._apply_fun, vstate.parameters, vstate.model_state, samples,
                            holomorphic=self.holomorphic)
        qgt = QGT(Jacobian(O), space=ParameterSpace())
        return QGTOperator(_qgt=qgt, diag_shift=self.diag_shift, _params_structure=vstate.parameters)

    def __call__(self, vstate, gradient, step=None):
        """AbstractLinearPreconditioner.__call__: solve (S + diag_shift) x = gradient."""
        self._lhs = self.lhs_constructor(vstate, step)
        self.x0, self.info = self._lhs.solve(self.solver, gradient)
        return self.x0
