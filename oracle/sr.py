"""FP64 restatement of the sliced Jacobian, QGT, SR solves and integrators.

Test infrastructure only. Follows
* src/vmc/qgt/jacobian.py (SliceOrdering, SiteOrdering, jacobian_mean),
* src/vmc/qgt/qgt.py (sliced blocks, matvecs, to_dense in parameter/sample space),
* src/vmc/qgt/solvers.py (cholesky / svd / cg),
* src/vmc/preconditioners/preconditioners.py (DirectSolve, QRSolve, reorder,
  SRPreconditioner.apply),
* src/vmc/drivers/integrators.py (Euler, RK4).

Pinning (tests/test_oracle_sr.py): sliced blocks == explicit full Jacobian
built from the sampled slices; S and T Hermitian PSD; sample/parameter-space
solution duality O^H (O O^H + e)^-1 == (O^H O + e)^-1 O^H; solvers vs
numpy.linalg.solve; RK4 vs expm and Euler first order. QRSolve's rank-revealing
path: parity unpinned beyond the full-rank case (no reference test exists).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.linalg as sla


@dataclass(frozen=True)
class SliceOrdering:
    pass


@dataclass(frozen=True)
class SiteOrdering:
    params_per_site: tuple


def sliced_blocks(o, p, sliced_dims, ordering):
    """qgt.py:62-87 (_iter_sliced_blocks): list of (masked block, width)."""
    if isinstance(ordering, SliceOrdering):
        return [(np.where(p == k, o, 0), o.shape[1]) for k in range(max(sliced_dims))]
    out = []
    i = 0
    for site, n in enumerate(ordering.params_per_site):
        for k in range(sliced_dims[site]):
            out.append((np.where(p[:, i:i + n] == k, o[:, i:i + n], 0), n))
        i += n
    return out


def dense_blocks(o, p, sliced_dims, ordering):
    """qgt.py:106-109: concatenated masked blocks (the expanded Jacobian)."""
    return np.concatenate([b for b, _ in sliced_blocks(o, p, sliced_dims, ordering)], axis=1)


def jacobian_mean(O=None, o=None, p=None, sliced_dims=None, ordering=None):
    """jacobian.py:96-140."""
    if O is not None:
        return O.mean(axis=0)
    return np.concatenate([b.mean(axis=0) for b, _ in sliced_blocks(o, p, sliced_dims, ordering)])


def qgt_dense_parameter(O):
    """qgt.py:173-180 / 197-205: S = O^H O / Ns - conj(mean) mean^T."""
    ns = O.shape[0]
    mean = O.mean(axis=0)
    return (O.conj().T @ O) / ns - np.conj(mean)[:, None] * mean[None, :]


def qgt_dense_sample(O):
    """qgt.py:182-195 / 206-223: centred T = O O^H / Ns with the reference's centring terms."""
    ns = O.shape[0]
    mean = O.mean(axis=0)
    G = (O @ O.conj().T) / ns
    ones = np.ones(ns)
    w = O @ np.conj(mean)
    G = G - (w[:, None] * ones[None, :]) / ns
    G = G - (ones[:, None] * np.conj(w)[None, :]) / ns
    G = G + (np.vdot(mean, mean) / ns) * np.ones((ns, ns))
    return G


def qgt_matvec_parameter(O, v):
    """qgt.py:116-122: S v."""
    ns = O.shape[0]
    mean = O.mean(axis=0)
    return (O.conj().T @ (O @ v)) / ns - np.conj(mean) * np.dot(mean, v)


def qgt_matvec_sample(O, v):
    """qgt.py:125-135: T v."""
    ns = O.shape[0]
    mean = O.mean(axis=0)
    r = (O @ (O.conj().T @ v)) / ns
    w = O @ np.conj(mean)
    vs = np.sum(v)
    r = r - w * (vs / ns)
    r = r - np.ones(ns) * (np.vdot(w, v) / ns)
    r = r + np.ones(ns) * (np.vdot(mean, mean) * vs / ns)
    return r


def adjoint_matvec(O, v):
    """preconditioners.py:100-114: O_c^H v with O_c = O - mean."""
    return O.conj().T @ v - np.conj(O.mean(axis=0)) * np.sum(v)


def solve_cholesky(mat, rhs):
    """solvers.py:25-28."""
    return sla.cho_solve(sla.cho_factor(mat), rhs)


def solve_svd(mat, rhs, rcond=1e-12):
    """solvers.py:31-37."""
    U, S, Vh = np.linalg.svd(mat, full_matrices=False)
    cutoff = rcond * np.max(S)
    S_inv = np.where(S > cutoff, 1.0 / np.where(S > cutoff, S, 1.0), 0.0)
    return Vh.conj().T @ (S_inv * (U.conj().T @ rhs))


def solve_cg(mat, rhs, maxiter=1000, tol=1e-5):
    """solvers.py:16-22: conjugate gradients on a Hermitian PD matrix."""
    x = np.zeros_like(rhs)
    r = rhs - mat @ x
    pdir = r.copy()
    rs = np.vdot(r, r).real
    bnorm = np.linalg.norm(rhs)
    for _ in range(maxiter):
        if np.sqrt(rs) <= tol * bnorm:
            break
        Ap = mat @ pdir
        alpha = rs / np.vdot(pdir, Ap).real
        x = x + alpha * pdir
        r = r - alpha * Ap
        rs_new = np.vdot(r, r).real
        pdir = r + (rs_new / rs) * pdir
        rs = rs_new
    return x


def direct_solve_parameter(O, dv, diag_shift, solver=solve_cholesky):
    """preconditioners.py:117-133."""
    rhs = adjoint_matvec(O, dv)
    S = qgt_dense_parameter(O)
    x = solver(S + diag_shift * np.eye(S.shape[0]), rhs)
    resid = np.linalg.norm(S @ x - rhs) ** 2 / np.linalg.norm(rhs) ** 2
    return x, {"residual_error": resid}


def direct_solve_sample(O, dv, diag_shift, solver=solve_cholesky):
    """preconditioners.py:136-152."""
    T = qgt_dense_sample(O)
    y = solver(T + diag_shift * np.eye(T.shape[0]), dv)
    x = adjoint_matvec(O, y)
    resid = np.linalg.norm(T @ y - dv) ** 2 / np.linalg.norm(dv) ** 2
    return x, {"residual_error": resid}


def qr_solve(O, dv, rcond=None, min_norm=True):
    """preconditioners.py:166-214: pivoted-QR min-norm solve (full Jacobian only)."""
    Oc = O - O.mean(axis=0)[None, :]
    q, r, piv = sla.qr(Oc, mode="economic", pivoting=True)
    y = q.conj().T @ dv
    rank = int(np.linalg.matrix_rank(r) if rcond is None else np.linalg.matrix_rank(r, rtol=rcond))
    n = r.shape[1]
    n2 = n - rank
    x = np.zeros(n, dtype=np.result_type(O.dtype, dv.dtype))
    if rank > 0:
        r11, r12, y1 = r[:rank, :rank], r[:rank, rank:], y[:rank]
        if min_norm and n2 > 0:
            a = sla.solve_triangular(r11, y1, lower=False)
            B = sla.solve_triangular(r11, r12, lower=False)
            lhs = B.conj().T @ B + np.eye(n2)
            x2 = sla.solve(lhs, B.conj().T @ a, assume_a="pos")
            x1 = a - B @ x2
        else:
            x1 = sla.solve_triangular(r11, y1, lower=False)
            x2 = np.zeros(n2, dtype=x.dtype)
        x[piv] = np.concatenate([x1, x2])
    rhs = Oc.conj().T @ dv
    resid = Oc.conj().T @ (Oc @ x) - rhs
    return x, {"residual_error": np.linalg.norm(resid) ** 2 / np.linalg.norm(rhs) ** 2, "rank": rank}


def reorder_slice_to_site(x, pps, sliced_dims):
    """preconditioners.py:61-86: k-major -> site-major permutation."""
    total = sum(pps)
    perm = []
    off = 0
    for site, n in enumerate(pps):
        for k in range(sliced_dims[site]):
            base = k * total + off
            perm.extend(range(base, base + n))
        off += n
    return x[np.asarray(perm)]


def sr_apply(o, p, local_energies, pps, sliced_dims, *, space="parameter", ordering=None,
             diag_shift=1e-2, grad_factor=1.0, real_params=True, solver=solve_cholesky):
    """preconditioners.py:268-330 (DirectSolve, no gauge projection).

    Returns the site-major update vector and the solver metrics.
    """
    ns = o.shape[0]
    ordering = ordering or SliceOrdering()
    dv = (local_energies.reshape(-1) - np.mean(local_energies)) / ns
    dv = grad_factor * dv
    if p is None:
        O = o
    else:
        O = dense_blocks(o, p, sliced_dims, ordering)
    if space == "parameter":
        x, metrics = direct_solve_parameter(O, dv, diag_shift, solver)
    else:
        x, metrics = direct_solve_sample(O, dv, diag_shift, solver)
    if p is not None and isinstance(ordering, SliceOrdering):
        x = reorder_slice_to_site(x, pps, sliced_dims)
    if real_params and np.isrealobj(grad_factor):
        x = 2.0 * x.real
    return x, metrics


def euler_step(f, y, t, dt):
    """integrators.py:45-63."""
    return y + dt * f(y, t), t + dt


def rk4_step(f, y, t, dt):
    """integrators.py:66-109."""
    k1 = f(y, t)
    k2 = f(y + 0.5 * dt * k1, t + 0.5 * dt)
    k3 = f(y + 0.5 * dt * k2, t + 0.5 * dt)
    k4 = f(y + dt * k3, t + dt)
    return y + dt * (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0, t + dt
