"""FP64 NumPy restatement of the reference PEPS hot path (test infrastructure only).

Batched over samples: every boundary/environment tensor carries a leading batch
axis ``z``. Pairwise einsum orders are the reference's ``optimize=[...]`` lists
(NumPy's ``['einsum_path', ...]`` uses the same operand-list convention), and
truncations use LAPACK SVD / eigh / QR exactly as the reference does.

Layouts (reference conventions):
* site tensor ``A[r][c]``: ``(d, up, down, left, right)`` (src/vmc/peps/standard/model.py:27-35)
* row-MPO tensor ``w``: ``(left, right, up, down)`` (src/vmc/peps/common/contraction.py:22-27)
* boundary-MPS tensor ``m``: ``(left, phys, right)`` (src/vmc/peps/common/strategy.py:273-275)
* configs: occupancy 0/1, row-major (src/vmc/utils/utils.py:55-57)

Pinning (tests/test_oracle_*.py):
* amplitudes (NoTruncation, ZipUp, DensityMatrix, Variational): brute-force
  full-network einsum (oracle.exact), strategy equivalence at chi >= exact rank
  (tol 1e-7 recipe of tests/test_peps_strategy_equivalence.py), multilinearity
  and bond-gauge invariance;
* gradients: central finite differences (recipe of
  tests/test_gradients_finite_diff.py), per-site Euler identity
  sum(A[s] * env_grad) == amp;
* local energies: dense-Hamiltonian brute force (recipe of
  tests/test_local_estimate_exact.py, tol 1e-9);
* sampler: exact |psi|^2 histogram (recipe of tests/test_sequential_sampling.py).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import NamedTuple

import numpy as np

from oracle.rng import sweep_draws


def _path(*pairs):
    return ["einsum_path", *pairs]


# --------------------------------------------------------------------------- #
# Model construction
# --------------------------------------------------------------------------- #


def site_dims(row, col, n_rows, n_cols, bond_dim):
    """src/vmc/peps/standard/model.py:27-35."""
    up = 1 if row == 0 else bond_dim
    down = 1 if row == n_rows - 1 else bond_dim
    left = 1 if col == 0 else bond_dim
    right = 1 if col == n_cols - 1 else bond_dim
    return up, down, left, right


def params_per_site(shape, bond_dim):
    """src/vmc/utils/smallo.py:23-35."""
    n_rows, n_cols = shape
    out = []
    for r in range(n_rows):
        for c in range(n_cols):
            u, d, l, rr = site_dims(r, c, n_rows, n_cols, bond_dim)
            out.append(u * d * l * rr)
    return out


def random_peps(shape, bond_dim, phys_dim=2, seed=0, dtype=np.float64):
    """Synthetic site tensors: i.i.d. N(0,1)/sqrt(D^4 d), PCG64 ``default_rng(seed)``.

    Sites are drawn in row-major order; a complex tensor draws its real part and
    then its imaginary part with the same scale. Float32/complex64 callers cast
    from this FP64 draw so both precisions see identical values.
    """
    rng = np.random.default_rng(seed)
    n_rows, n_cols = shape
    scale = 1.0 / np.sqrt(float(bond_dim) ** 4 * phys_dim)
    cplx = np.issubdtype(np.dtype(dtype), np.complexfloating)
    tensors = []
    for r in range(n_rows):
        row = []
        for c in range(n_cols):
            dims = (phys_dim, *site_dims(r, c, n_rows, n_cols, bond_dim))
            t = rng.standard_normal(dims) * scale
            if cplx:
                t = t + 1j * (rng.standard_normal(dims) * scale)
            row.append(t.astype(np.complex128 if cplx else np.float64))
        tensors.append(row)
    return tensors


def random_configs(n_sites, n_samples, seed=0, phys_dim=2):
    """Uniform configs from ``default_rng(seed + 1)``, int32 (n_samples, n_sites)."""
    rng = np.random.default_rng(seed + 1)
    return rng.integers(0, phys_dim, size=(n_samples, n_sites)).astype(np.int32)


def spin_to_occupancy(spins):
    """src/vmc/utils/utils.py:55-57 (idempotent on {0,1})."""
    return ((np.asarray(spins) + 1) // 2).astype(np.int32)


# --------------------------------------------------------------------------- #
# Boundary-MPS strategies (src/vmc/peps/common/strategy.py)
# --------------------------------------------------------------------------- #


def contract_theta(m, w, carry):
    """strategy.py:95-115. m (z,Dl,p,Dr), w (z,wl,wr,p,q), carry (z,k,Dl,wl)|None."""
    if carry is not None:
        theta = np.einsum(
            "zkdl,zdpr,zlwpq->zkqrw", carry, m, w, optimize=_path((0, 1), (0, 1))
        )
        return theta
    theta = np.einsum("zdpr,zlwpq->zdlqrw", m, w)
    z, Dl, wl, q, Dr, wr = theta.shape
    return theta.reshape(z, Dl * wl, q, Dr, wr)


def apply_mpo_exact(mps, mpo):
    """strategy.py:118-125."""
    out = []
    for m, w in zip(mps, mpo):
        c = np.einsum("zapb,zlrpq->zalqbr", m, w)
        z, a, l, q, b, r = c.shape
        out.append(c.reshape(z, a * l, q, b * r))
    return tuple(out)


def apply_mpo_zip_up(mps, mpo, chi):
    """strategy.py:128-150: on-the-fly SVD truncation, keep min(chi, rank bound)."""
    new = []
    carry = None
    n = len(mps)
    for i, (m, w) in enumerate(zip(mps, mpo)):
        theta = contract_theta(m, w, carry)
        z, left, q, Dr, wr = theta.shape
        if i == n - 1:
            new.append(theta.reshape(z, left, q, Dr * wr))
            break
        mat = theta.reshape(z, left * q, Dr * wr)
        U, S, Vh = np.linalg.svd(mat, full_matrices=False)
        k = min(chi, S.shape[-1])
        new.append(U[:, :, :k].reshape(z, left, q, k))
        carry = (S[:, :k, None].astype(Vh.dtype) * Vh[:, :k, :]).reshape(z, k, Dr, wr)
    return tuple(new)


def apply_mpo_density_matrix(mps, mpo, chi):
    """strategy.py:153-176: top-k eigenvectors of theta^H theta."""
    new = []
    carry = None
    n = len(mps)
    for i, (m, w) in enumerate(zip(mps, mpo)):
        theta = contract_theta(m, w, carry)
        z, left, q, Dr, wr = theta.shape
        theta = theta.reshape(z, left * q, Dr * wr)
        if i == n - 1:
            new.append(theta.reshape(z, left, q, Dr * wr))
            break
        rho = np.conj(np.swapaxes(theta, 1, 2)) @ theta
        evals, evecs = np.linalg.eigh(rho)
        k = min(chi, rho.shape[-1])
        order = np.argsort(evals, axis=-1)[:, ::-1][:, :k]
        vecs_k = np.take_along_axis(evecs, order[:, None, :], axis=2)
        new.append((theta @ vecs_k).reshape(z, left, q, k))
        carry = np.conj(np.swapaxes(vecs_k, 1, 2)).reshape(z, k, Dr, wr)
    return tuple(new)


def qr_reduced(a):
    """src/vmc/utils/factorizations.py:9-16: Householder (geqrf) reduced QR."""
    q, r = np.linalg.qr(a, mode="reduced")
    return q, r


def _init_compressed_mps(mps, mpo, Dc):
    """strategy.py:222-254."""
    result = []
    carry = None
    n = len(mps)
    for i, (m, w) in enumerate(zip(mps, mpo)):
        theta = contract_theta(m, w, carry)
        z, left, q, Dr, wr = theta.shape
        if i == n - 1:
            result.append(theta.reshape(z, left, q, Dr * wr))
            break
        mat = theta.reshape(z, left * q, Dr * wr)
        Q, R = qr_reduced(mat)
        k = min(Dc, Q.shape[-1])
        result.append(Q[:, :, :k].reshape(z, left, q, k))
        carry = R[:, :k, :].reshape(z, k, Dr, wr)
    return result


def _build_right_envs(mps, mpo, result, dtype):
    """strategy.py:257-289."""
    n = len(mps)
    z = mps[0].shape[0]
    envs = [None] * n
    R = np.ones((z, 1, 1, 1), dtype=dtype)
    envs[n - 1] = R
    for i in range(n - 2, -1, -1):
        R = np.einsum(
            "zaqb,zrwb,zlwpq,zdpr->zdla",
            np.conj(result[i + 1]), R, mpo[i + 1], mps[i + 1],
            optimize=_path((1, 0), (2, 0), (1, 0)),
        )
        envs[i] = R
    return envs


def _variational_sweep_lr(mps, mpo, result, right_envs, dtype):
    """strategy.py:292-356."""
    n = len(mps)
    z = mps[0].shape[0]
    new_result = list(result)
    left_envs = [None] * n
    L = np.ones((z, 1, 1, 1), dtype=dtype)
    for i in range(n):
        left_envs[i] = L
        optimal = np.einsum(
            "zdpr,zdla,zlwpq,zrwb->zaqb", mps[i], L, mpo[i], right_envs[i],
            optimize=_path((1, 0), (2, 0), (1, 0)),
        )
        if i == n - 1:
            new_result[i] = optimal
            break
        _, left, q, right = optimal.shape
        A, C = qr_reduced(optimal.reshape(z, left * q, right))
        new_tensor = A.reshape(z, left, q, A.shape[-1])
        new_result[i] = new_tensor
        new_result[i + 1] = np.einsum("zab,zbpr->zapr", C, new_result[i + 1])
        L = np.einsum(
            "zdpr,zdla,zlwpq,zaqb->zrwb", mps[i], L, mpo[i], np.conj(new_tensor),
            optimize=_path((1, 0), (2, 0), (1, 0)),
        )
    return new_result, left_envs


def _variational_sweep_rl(mps, mpo, result, left_envs, dtype):
    """strategy.py:359-427."""
    n = len(mps)
    z = mps[0].shape[0]
    new_result = list(result)
    right_envs = [None] * n
    R = np.ones((z, 1, 1, 1), dtype=dtype)
    for i in range(n - 1, -1, -1):
        right_envs[i] = R
        optimal = np.einsum(
            "zdpr,zdla,zlwpq,zrwb->zaqb", mps[i], left_envs[i], mpo[i], R,
            optimize=_path((1, 0), (2, 0), (1, 0)),
        )
        if i == 0:
            new_result[i] = optimal
            break
        _, left, q, right = optimal.shape
        mat = optimal.reshape(z, left, q * right)
        Q_t, R_t = qr_reduced(np.swapaxes(mat, 1, 2))
        B = np.swapaxes(Q_t, 1, 2)
        C = np.swapaxes(R_t, 1, 2)
        new_tensor = B.reshape(z, B.shape[1], q, right)
        new_result[i] = new_tensor
        new_result[i - 1] = np.einsum("zlpr,zrb->zlpb", new_result[i - 1], C)
        R = np.einsum(
            "zaqb,zrwb,zlwpq,zdpr->zdla", np.conj(new_tensor), R, mpo[i], mps[i],
            optimize=_path((1, 0), (2, 0), (1, 0)),
        )
    return new_result, right_envs


def apply_mpo_variational(mps, mpo, chi, n_sweeps=2):
    """strategy.py:179-219: QR-initialised one-site ALS compression."""
    dtype = mps[0].dtype
    result = _init_compressed_mps(mps, mpo, chi)
    right_envs = _build_right_envs(mps, mpo, result, dtype)
    for _ in range(n_sweeps):
        result, left_envs = _variational_sweep_lr(mps, mpo, result, right_envs, dtype)
        result, right_envs = _variational_sweep_rl(mps, mpo, result, left_envs, dtype)
    return tuple(result)


class NoTruncation:
    """strategy.py:44-51."""

    truncate_bond_dimension = 1

    def apply(self, mps, mpo):
        return apply_mpo_exact(mps, mpo)


@dataclass
class ZipUp:
    """strategy.py:54-61."""

    truncate_bond_dimension: int

    def apply(self, mps, mpo):
        return apply_mpo_zip_up(mps, mpo, self.truncate_bond_dimension)


@dataclass
class DensityMatrix:
    """strategy.py:64-71."""

    truncate_bond_dimension: int

    def apply(self, mps, mpo):
        return apply_mpo_density_matrix(mps, mpo, self.truncate_bond_dimension)


@dataclass
class Variational:
    """strategy.py:74-92."""

    truncate_bond_dimension: int
    n_sweeps: int = 2

    def apply(self, mps, mpo):
        return apply_mpo_variational(mps, mpo, self.truncate_bond_dimension, self.n_sweeps)


# --------------------------------------------------------------------------- #
# Contraction primitives (src/vmc/peps/common/contraction.py)
# --------------------------------------------------------------------------- #


def build_row_mpo(tensors, spins_row, row):
    """contraction.py:22-27: per-sample physical slice -> (z, l, r, u, d)."""
    out = []
    for col, t in enumerate(tensors[row]):
        sl = t[spins_row[:, col]]  # (z, u, d, l, r)
        out.append(np.transpose(sl, (0, 3, 4, 1, 2)))
    return tuple(out)


def ones_boundary(z, n_cols, dtype):
    return tuple(np.ones((z, 1, 1, 1), dtype=dtype) for _ in range(n_cols))


def contract_bottom(mps):
    """contraction.py:30-35."""
    z = mps[0].shape[0]
    state = np.ones((z, 1), dtype=mps[0].dtype)
    for site in mps:
        state = np.einsum("za,zab->zb", state, site[:, :, 0, :])
    return state[:, 0]


def forward_with_cache(tensors, spins, strategy):
    """contraction.py:38-56. spins (z, n_rows, n_cols)."""
    z, n_rows, n_cols = spins.shape
    dtype = tensors[0][0].dtype
    top_envs = [None] * n_rows
    boundary = ones_boundary(z, n_cols, dtype)
    for row in range(n_rows):
        top_envs[row] = boundary
        mpo = build_row_mpo(tensors, spins[:, row], row)
        boundary = strategy.apply(boundary, mpo)
    return contract_bottom(boundary), top_envs


def apply_mpo_from_below(bottom_mps, mpo, strategy):
    """contraction.py:59-65."""
    return strategy.apply(bottom_mps, tuple(np.transpose(w, (0, 1, 2, 4, 3)) for w in mpo))


def compute_right_envs(top_env, mpo_row, bottom_env, dtype):
    """contraction.py:68-86."""
    n_cols = len(mpo_row)
    z = mpo_row[0].shape[0]
    envs = [None] * n_cols
    envs[n_cols - 1] = np.ones((z, 1, 1, 1), dtype=dtype)
    for c in range(n_cols - 2, -1, -1):
        envs[c] = np.einsum(
            "zaub,zcduv,zevf,zbdf->zace",
            top_env[c + 1], mpo_row[c + 1], bottom_env[c + 1], envs[c + 1],
            optimize=_path((0, 3), (0, 2), (0, 1)),
        )
    return envs


def metropolis_ratio(prob_cur, prob_flip):
    """contraction.py:88-94."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(
            prob_cur > 0.0,
            prob_flip / np.where(prob_cur > 0.0, prob_cur, 1.0),
            np.where(prob_flip > 0.0, np.inf, 0.0),
        )


def left_env_step(left_env, top, mpo, bot):
    """Inline left-environment update, energy.py:114-119 / 201-205, kernels.py:197-204."""
    return np.einsum(
        "zace,zaub,zcduv,zevf->zbdf", left_env, top, mpo, bot,
        optimize=_path((0, 1), (0, 2), (0, 1)),
    )


# --------------------------------------------------------------------------- #
# Gradients and energies (src/vmc/peps/common/energy.py)
# --------------------------------------------------------------------------- #


def compute_single_gradient(left_env, right_env, top, bot):
    """energy.py:259-273 -> (z, up, down, left, right)."""
    g = np.einsum(
        "zace,zaub,zevf,zbdf->zcuvd", left_env, top, bot, right_env,
        optimize=_path((0, 1), (0, 1), (0, 1)),
    )
    return np.transpose(g, (0, 2, 3, 1, 4))


def compute_all_row_gradients(top_mps, bottom_mps, mpo):
    """energy.py:98-120."""
    n_cols = len(mpo)
    z = mpo[0].shape[0]
    dtype = mpo[0].dtype
    right_envs = compute_right_envs(top_mps, mpo, bottom_mps, dtype)
    grads = []
    left_env = np.ones((z, 1, 1, 1), dtype=dtype)
    for c in range(n_cols):
        grads.append(compute_single_gradient(left_env, right_envs[c], top_mps[c], bottom_mps[c]))
        left_env = left_env_step(left_env, top_mps[c], mpo[c], bottom_mps[c])
    return grads


def compute_all_gradients(tensors, spins, strategy, top_envs):
    """energy.py:276-301 (without the bottom-env cache option)."""
    z, n_rows, n_cols = spins.shape
    dtype = tensors[0][0].dtype
    grads = [None] * n_rows
    bottom = ones_boundary(z, n_cols, dtype)
    for row in range(n_rows - 1, -1, -1):
        mpo = build_row_mpo(tensors, spins[:, row], row)
        grads[row] = compute_all_row_gradients(top_envs[row], bottom, mpo)
        if row > 0:
            bottom = apply_mpo_from_below(bottom, mpo, strategy)
    return grads


def compute_right_envs_2row(top_env, mpo0, mpo1, bottom_env, dtype):
    """energy.py:24-44."""
    n_cols = len(mpo0)
    z = mpo0[0].shape[0]
    envs = [None] * n_cols
    envs[n_cols - 1] = np.ones((z, 1, 1, 1, 1), dtype=dtype)
    for c in range(n_cols - 2, -1, -1):
        envs[c] = np.einsum(
            "zaub,zlruv,zxyvw,zewf,zbryf->zalxe",
            top_env[c + 1], mpo0[c + 1], mpo1[c + 1], bottom_env[c + 1], envs[c + 1],
            optimize=_path((0, 4), (0, 3), (0, 2), (0, 1)),
        )
    return envs


def compute_row_pair_vertical_energy(
    top_mps, bottom_mps, mpo0, mpo1, tensors_row0, tensors_row1,
    spins_row0, spins_row1, terms_row, amp, phys_dim,
):
    """energy.py:47-95. ``terms_row[c]`` is a list of (d^2, d^2) operators."""
    z = amp.shape[0]
    out_dtype = np.result_type(amp.dtype, mpo0[0].dtype)
    if not any(terms_row):
        return np.zeros((z,), dtype=out_dtype)
    n_cols = len(mpo0)
    dtype = mpo0[0].dtype
    right_envs = compute_right_envs_2row(top_mps, mpo0, mpo1, bottom_mps, dtype)
    left_env = np.ones((z, 1, 1, 1, 1), dtype=dtype)
    energy = np.zeros((z,), dtype=out_dtype)
    for c in range(n_cols):
        if terms_row[c]:
            amps_edge = np.einsum(
                "zalmg,zaub,puvlr,qvwmn,zgwf,zbrnf->zpq",
                left_env, top_mps[c], tensors_row0[c], tensors_row1[c],
                bottom_mps[c], right_envs[c],
                optimize=_path((0, 1), (2, 3), (0, 2), (1, 2), (0, 1)),
            )
            col_idx = spins_row0[:, c] * phys_dim + spins_row1[:, c]
            amps_flat = amps_edge.reshape(z, -1)
            for op in terms_row[c]:
                energy = energy + np.einsum("zp,zp->z", op[:, col_idx].T, amps_flat) / amp
        left_env = np.einsum(
            "zalxe,zaub,zlruv,zxyvw,zewf->zbryf",
            left_env, top_mps[c], mpo0[c], mpo1[c], bottom_mps[c],
            optimize=_path((0, 1), (0, 3), (0, 2), (0, 1)),
        )
    return energy


@dataclass
class BucketedTerms:
    """Operator terms grouped by lattice location (src/vmc/operators/local_terms.py:165-200).

    diagonal: list of (sites tuple, diag vector); one_site[r][c], horizontal[r][c],
    vertical[r][c]: lists of dense operators.
    """

    shape: tuple
    diagonal: list = field(default_factory=list)
    one_site: list = None
    horizontal: list = None
    vertical: list = None

    def __post_init__(self):
        n_rows, n_cols = self.shape
        if self.one_site is None:
            self.one_site = [[[] for _ in range(n_cols)] for _ in range(n_rows)]
        if self.horizontal is None:
            self.horizontal = [[[] for _ in range(max(n_cols - 1, 0))] for _ in range(n_rows)]
        if self.vertical is None:
            self.vertical = [[[] for _ in range(n_cols)] for _ in range(max(n_rows - 1, 0))]


def compute_all_env_grads_and_energy(
    tensors, spins, amp, strategy, top_envs, terms: BucketedTerms, collect_grads=True
):
    """energy.py:123-224."""
    z, n_rows, n_cols = spins.shape
    dtype = tensors[0][0].dtype
    phys_dim = tensors[0][0].shape[0]
    env_grads = [[None] * n_cols for _ in range(n_rows)] if collect_grads else []
    bottom_cache = [None] * n_rows
    out_dtype = np.result_type(amp.dtype, dtype)
    energy = np.zeros((z,), dtype=out_dtype)

    for sites, diag in terms.diagonal:
        idx = np.zeros((z,), dtype=np.int64)
        for (r, c) in sites:
            idx = idx * phys_dim + spins[:, r, c]
        energy = energy + np.asarray(diag)[idx]

    bottom = ones_boundary(z, n_cols, dtype)
    next_row_mpo = None
    for row in range(n_rows - 1, -1, -1):
        bottom_cache[row] = bottom
        top = top_envs[row]
        mpo = build_row_mpo(tensors, spins[:, row], row)
        right_envs = compute_right_envs(top, mpo, bottom, dtype)
        left_env = np.ones((z, 1, 1, 1), dtype=dtype)
        for c in range(n_cols):
            site_terms = terms.one_site[row][c]
            if collect_grads or site_terms:
                env_grad = compute_single_gradient(left_env, right_envs[c], top[c], bottom[c])
                if collect_grads:
                    env_grads[row][c] = env_grad
                if site_terms:
                    amps_site = np.einsum("pudlr,zudlr->zp", tensors[row][c], env_grad)
                    s = spins[:, row, c]
                    for op in site_terms:
                        energy = energy + np.einsum("zp,zp->z", op[:, s].T, amps_site) / amp
            if c < n_cols - 1 and terms.horizontal[row][c]:
                amps_edge = np.einsum(
                    "zace,zaub,zedf,pudcr,qvwrx,zbvg,zfwi,zgxi->zpq",
                    left_env, top[c], bottom[c], tensors[row][c], tensors[row][c + 1],
                    top[c + 1], bottom[c + 1], right_envs[c + 1],
                    optimize=_path((0, 1), (1, 6), (0, 5), (1, 3), (1, 2), (1, 2), (0, 1)),
                )
                col_idx = spins[:, row, c] * phys_dim + spins[:, row, c + 1]
                amps_flat = amps_edge.reshape(z, -1)
                for op in terms.horizontal[row][c]:
                    energy = energy + np.einsum("zp,zp->z", op[:, col_idx].T, amps_flat) / amp
            left_env = left_env_step(left_env, top[c], mpo[c], bottom[c])
        if row < n_rows - 1:
            energy = energy + compute_row_pair_vertical_energy(
                top, bottom_cache[row + 1], mpo, next_row_mpo,
                tensors[row], tensors[row + 1], spins[:, row], spins[:, row + 1],
                terms.vertical[row], amp, phys_dim,
            )
        bottom = apply_mpo_from_below(bottom, mpo, strategy)
        next_row_mpo = mpo
    return env_grads, energy, bottom_cache


# --------------------------------------------------------------------------- #
# Evaluator (src/vmc/peps/standard/compat.py)
# --------------------------------------------------------------------------- #


def value(tensors, samples, shape, strategy):
    """compat.py:67-78 (``_value``): amplitudes of (z, n_sites) samples."""
    spins = spin_to_occupancy(samples).reshape(-1, *shape)
    amp, _ = forward_with_cache(tensors, spins, strategy)
    return amp


def value_and_grad(tensors, samples, shape, strategy, bond_dim, full_gradient=False):
    """compat.py:94-133: (amp, dpsi/dA, p) -- grads NOT divided by amp."""
    spins = spin_to_occupancy(samples).reshape(-1, *shape)
    z = spins.shape[0]
    n_rows, n_cols = shape
    amp, top_envs = forward_with_cache(tensors, spins, strategy)
    env_grads = compute_all_gradients(tensors, spins, strategy, top_envs)
    grad_parts, p_parts = [], []
    for r in range(n_rows):
        for c in range(n_cols):
            g = env_grads[r][c].reshape(z, -1)
            if full_gradient:
                d = tensors[r][c].shape[0]
                full = np.zeros((z, d, g.shape[1]), dtype=g.dtype)
                full[np.arange(z), spins[:, r, c]] = g
                grad_parts.append(full.reshape(z, -1))
            else:
                grad_parts.append(g)
                p_parts.append(np.repeat(spins[:, r, c][:, None], g.shape[1], axis=1).astype(np.int8))
    grad = np.concatenate(grad_parts, axis=1)
    if full_gradient:
        return amp, grad, None
    return amp, grad, np.concatenate(p_parts, axis=1)


def local_estimate(tensors, samples, shape, strategy, terms: BucketedTerms, amps):
    """compat.py:136-194."""
    spins = spin_to_occupancy(samples).reshape(-1, *shape)
    _, top_envs = forward_with_cache(tensors, spins, strategy)
    _, energy, _ = compute_all_env_grads_and_energy(
        tensors, spins, np.asarray(amps), strategy, top_envs, terms, collect_grads=False
    )
    return energy


# --------------------------------------------------------------------------- #
# MC kernels (src/vmc/peps/standard/kernels.py) and sampler (src/vmc/core/sampling.py)
# --------------------------------------------------------------------------- #


class Cache(NamedTuple):
    bottom_envs: tuple


class Context(NamedTuple):
    amp: np.ndarray
    top_envs: tuple


class LocalEstimates(NamedTuple):
    local_log_derivatives: np.ndarray
    local_estimate: np.ndarray
    active_slice_indices: np.ndarray | None
    amp: np.ndarray | None = None


def init_cache(tensors, config_states, shape, strategy):
    """kernels.py:115-129: bottom boundaries of each chain's config."""
    spins = np.asarray(config_states).reshape(-1, *shape)
    z, n_rows, n_cols = spins.shape
    dtype = tensors[0][0].dtype
    envs = [None] * n_rows
    env = ones_boundary(z, n_cols, dtype)
    for row in range(n_rows - 1, -1, -1):
        envs[row] = env
        mpo = build_row_mpo(tensors, spins[:, row], row)
        env = apply_mpo_from_below(env, mpo, strategy)
    return Cache(bottom_envs=tuple(envs))


def transition(tensors, config_states, cache, shape, strategy, seed, sweep, chain_ids):
    """kernels.py:131-209: one sequential top->bottom Metropolis sweep.

    Draws come from Philox (oracle.rng) instead of JAX threefry: draw for site
    ``i`` of sweep ``sweep`` of chain ``b`` uses counter ``(i, sweep, b, 0)``.
    """
    spins = np.array(config_states).reshape(-1, *shape).copy()
    z, n_rows, n_cols = spins.shape
    dtype = tensors[0][0].dtype
    top_env = ones_boundary(z, n_cols, dtype)
    top_cache = [None] * n_rows
    for row in range(n_rows):
        top_cache[row] = top_env
        bottom_env = cache.bottom_envs[row]
        mpo_row = build_row_mpo(tensors, spins[:, row], row)
        right_envs = compute_right_envs(top_env, mpo_row, bottom_env, dtype)
        left_env = np.ones((z, 1, 1, 1), dtype=dtype)
        amp_cur = np.einsum(
            "zace,zaub,zcduv,zevf,zbdf->z",
            left_env, top_env[0], mpo_row[0], bottom_env[0], right_envs[0],
            optimize=_path((0, 1), (1, 2), (1, 2), (0, 1)),
        )
        updated = []
        for col in range(n_cols):
            t = tensors[row][col]
            d = t.shape[0]
            cur = spins[:, row, col]
            u, word = sweep_draws(seed, sweep, row * n_cols + col, chain_ids)
            if d == 1:
                flip = cur
            elif d == 2:
                flip = 1 - cur
            else:
                delta = 1 + (word % np.uint32(d - 1)).astype(np.int32)
                flip = (cur + delta) % d
            mpo_cur = mpo_row[col]
            mpo_flip = np.transpose(t[flip], (0, 3, 4, 1, 2))
            amp_flip = np.einsum(
                "zace,zaub,zcduv,zevf,zbdf->z",
                left_env, top_env[col], mpo_flip, bottom_env[col], right_envs[col],
                optimize=_path((0, 1), (1, 2), (1, 2), (0, 1)),
            )
            ratio = metropolis_ratio(np.abs(amp_cur) ** 2, np.abs(amp_flip) ** 2)
            accept = u < np.minimum(1.0, ratio)
            spins[:, row, col] = np.where(accept, flip, cur)
            amp_cur = np.where(accept, amp_flip, amp_cur)
            mpo_sel = np.where(accept[:, None, None, None, None], mpo_flip, mpo_cur)
            updated.append(mpo_sel)
            left_env = left_env_step(left_env, top_env[col], mpo_sel, bottom_env[col])
        top_env = strategy.apply(top_env, tuple(updated))
    amp = contract_bottom(top_env)
    return spins.reshape(z, -1), Context(amp=amp, top_envs=tuple(top_cache))


def assemble_log_derivatives(tensors, shape, env_grads, config_state, amp, full_gradient):
    """kernels.py:54-88."""
    n_rows, n_cols = shape
    z = amp.shape[0]
    spins = np.asarray(config_state).reshape(z, n_rows, n_cols)
    parts, p_parts = [], []
    for r in range(n_rows):
        for c in range(n_cols):
            g = env_grads[r][c].reshape(z, -1)
            if full_gradient:
                d = tensors[r][c].shape[0]
                full = np.zeros((z, d, g.shape[1]), dtype=g.dtype)
                full[np.arange(z), spins[:, r, c]] = g
                parts.append(full.reshape(z, -1))
            else:
                parts.append(g)
                p_parts.append(np.repeat(spins[:, r, c][:, None], g.shape[1], axis=1).astype(np.int8))
    o = np.concatenate(parts, axis=1) / amp[:, None]
    return o, (None if full_gradient else np.concatenate(p_parts, axis=1))


def estimate(tensors, config_next, context, shape, strategy, terms, full_gradient=False):
    """kernels.py:211-245."""
    spins = np.asarray(config_next).reshape(-1, *shape)
    env_grads, eloc, bottom_next = compute_all_env_grads_and_energy(
        tensors, spins, context.amp, strategy, context.top_envs, terms, collect_grads=True
    )
    o, p = assemble_log_derivatives(tensors, shape, env_grads, config_next, context.amp, full_gradient)
    return Cache(bottom_envs=tuple(bottom_next)), LocalEstimates(o, eloc, p, context.amp)


def mc_sampler(tensors, config_states, cache, shape, strategy, terms, seed, n_steps,
               full_gradient=False, sweep0=0):
    """sampling.py:22-64: scan over sweeps of the batched (vmapped) mc_sweep.

    Returns final (configs, cache) and step-major histories
    samples (n_steps, chains, n_sites), o, eloc, p, amp.
    """
    z = np.asarray(config_states).shape[0]
    chain_ids = np.arange(z, dtype=np.uint32)
    configs = np.asarray(config_states)
    hist = {"samples": [], "o": [], "eloc": [], "p": [], "amp": []}
    for t in range(n_steps):
        configs, ctx = transition(tensors, configs, cache, shape, strategy, seed, sweep0 + t, chain_ids)
        cache, est = estimate(tensors, configs, ctx, shape, strategy, terms, full_gradient)
        hist["samples"].append(configs.copy())
        hist["o"].append(est.local_log_derivatives)
        hist["eloc"].append(est.local_estimate)
        hist["p"].append(est.active_slice_indices)
        hist["amp"].append(est.amp)
    out = {k: (np.stack(v) if v[0] is not None else None) for k, v in hist.items()}
    return configs, cache, out


def sample_counts(n_samples, n_chains):
    """sampling.py:76-78."""
    chain_length = -(-n_samples // n_chains)
    return n_samples, n_chains, chain_length, chain_length * n_chains


def trim_samples(samples, total_samples, num_samples):
    """sampling.py:81-82."""
    return samples.reshape((total_samples,) + samples.shape[2:])[:num_samples]
