"""FP64 restatement of the blockade PEPS path (src/vmc/peps/blockade/*).

Test infrastructure only: imported by tests/, never by the product path.

Follows
* src/vmc/peps/blockade/model.py: BlockadePEPSConfig and its charge masks
  (:43-66), tensor shapes (:100-121), site assembly (:157-212), the blockade
  check (:220-239), row MPOs (:267-292), estimate (:295-617), transition
  (:620-665), the single-row sweep (:668-817), the row-pair sweep (:820-994)
  and the two-row windows (:997-1033);
* src/vmc/peps/blockade/compat.py:17-62 (blockade_apply: zero amplitude for a
  sample with two adjacent occupied sites);
* src/vmc/peps/blockade/kernels.py:50-117 (sliced log-derivatives, combined
  slice index n * nc + kL * stride + kU);
* src/vmc/peps/blockade/hamiltonian.py:22-79 (rydberg_hamiltonian).

A site's effective tensor is the slice (n, cfg) of the stored
(2, nc, mu_u, mu_d, mu_l, mu_r) tensor, cfg = kL * stride + kU for n = 0 and
0 for n = 1, times the charge masks of its bonds (kU, n, kL, n). For every
configuration the amplitude, sampler and estimate ever evaluate (valid ones:
n = 1 implies kL = kU = 0) the masks are fixed by the slice, so the slices can
be masked once (`flat_tensors`); `literal_amplitude` applies the masks per
sample exactly as _assemble_site does, and the tests check the two agree.

Draws (the reference splits JAX threefry keys) come from oracle.rng Philox:
the flip of site (r, c) in sweep t for chain b uses counter
(BL_BASE + r * n_cols + c, t, b, 0) and its uniform is the accept draw, in the
row-pair sweeps and in the last row's single-row sweep alike.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from oracle import gi as GI
from oracle import peps as P
from oracle import rng as R

BL_BASE = 4 << 20

_ROW2_PATH = P._path((0, 1), (0, 4), (0, 3), (0, 2), (2, 3), (0, 2), (0, 1))  # model.py:713-717
_ROW1_PATH = P._path((0, 1), (1, 2), (1, 2), (0, 1))  # model.py:757-760


@dataclass(frozen=True)
class BLConfig:
    """model.py:43-66."""

    shape: tuple
    D0: int
    D1: int

    @property
    def dmax(self):
        return max(self.D0, self.D1)

    def mask(self):
        """mask_per_charge: (2, Dmax) bool, None when both sectors are full."""
        dm = self.dmax
        if self.D0 == dm and self.D1 == dm:
            return None
        return np.arange(dm)[None, :] < np.array([self.D0, self.D1])[:, None]


def nc_of(shape, r, c):
    """2 ** (incoming bonds): model.py:104-107."""
    return 2 ** (int(r > 0) + int(c > 0))


def site_shape(config, r, c):
    n_rows, n_cols = config.shape
    dm = config.dmax
    return (2, nc_of(config.shape, r, c), dm if r > 0 else 1, dm if r < n_rows - 1 else 1,
            dm if c > 0 else 1, dm if c < n_cols - 1 else 1)


def random_tensors(config, seed=0, dtype=np.float64):
    """N(0, 1) / sqrt(Dmax^4 * 2) entries, sites in row-major order."""
    rng = np.random.default_rng(seed)
    scale = 1.0 / np.sqrt(config.dmax ** 4 * 2.0)
    out = []
    for r in range(config.shape[0]):
        row = []
        for c in range(config.shape[1]):
            t = rng.standard_normal(site_shape(config, r, c)) * scale
            if np.iscomplexobj(np.zeros(1, dtype)):
                t = t + 1j * rng.standard_normal(site_shape(config, r, c)) * scale
            row.append(t.astype(dtype))
        out.append(row)
    return out


def valid(occ):
    """compat.py:47-50: no two adjacent occupied sites. occ (z, n_rows, n_cols)."""
    bad_h = np.any(occ[:, :, 1:] * occ[:, :, :-1], axis=(1, 2))
    bad_v = np.any(occ[:, 1:, :] * occ[:, :-1, :], axis=(1, 2))
    return ~(bad_h | bad_v)


def random_independent_sets(shape, n, seed=0):
    """model.py:242-264 with a numpy generator: greedy row-major occupation of
    free sites, each with probability 1/2."""
    rng = np.random.default_rng(seed)
    n_rows, n_cols = shape
    occ = np.zeros((n, n_rows, n_cols), np.int64)
    for r in range(n_rows):
        for c in range(n_cols):
            free = np.ones(n, bool)
            if c > 0:
                free &= occ[:, r, c - 1] == 0
            if r > 0:
                free &= occ[:, r - 1, c] == 0
            occ[:, r, c] = (rng.random(n) < 0.5) & free
    return occ


def _cfg(n, kl, ku, r):
    """model.py:188-191."""
    return np.where(n == 0, kl * (2 if r > 0 else 1) + ku, 0)


def _neighbours(occ, r, c):
    z = occ.shape[0]
    kl = occ[:, r, c - 1] if c > 0 else np.zeros(z, np.int64)
    ku = occ[:, r - 1, c] if r > 0 else np.zeros(z, np.int64)
    return kl, ku


def slices(config, occ):
    """Combined slice n * nc + cfg of every site, (z, n_sites)."""
    n_rows, n_cols = config.shape
    out = np.zeros((occ.shape[0], n_rows * n_cols), np.int64)
    for r in range(n_rows):
        for c in range(n_cols):
            kl, ku = _neighbours(occ, r, c)
            n = occ[:, r, c]
            out[:, r * n_cols + c] = n * nc_of(config.shape, r, c) + _cfg(n, kl, ku, r)
    return out


def active_slice_indices(config, occ):
    """kernels.py:93-103: n * nc + kL * stride + kU (equal to the slice for valid samples)."""
    n_rows, n_cols = config.shape
    out = np.zeros((occ.shape[0], n_rows * n_cols), np.int64)
    for r in range(n_rows):
        for c in range(n_cols):
            kl, ku = _neighbours(occ, r, c)
            out[:, r * n_cols + c] = occ[:, r, c] * nc_of(config.shape, r, c) + kl * (2 if r > 0 else 1) + ku
    return out


def _mask_slice(t, mask, n, kl, ku):
    """model.py:201-212 on one (u, d, l, r) slice."""
    if mask is None:
        return t
    u, d, l, r = t.shape
    return (t * mask[ku][:u][:, None, None, None] * mask[n][:d][None, :, None, None]
            * mask[kl][:l][None, None, :, None] * mask[n][:r][None, None, None, :])


def flat_tensors(config, tensors):
    """(2, nc, u, d, l, r) -> (2 * nc, u, d, l, r) with the charge masks of each
    slice applied: slice (0, kL * stride + kU) carries (kU, 0, kL, 0), slice
    (1, 0) carries (0, 1, 0, 1)."""
    mask = config.mask()
    out = []
    for r in range(config.shape[0]):
        row = []
        for c in range(config.shape[1]):
            t = np.array(tensors[r][c])
            nc = t.shape[1]
            for n in (0, 1):
                for cfg in range(nc):
                    if n == 0:
                        kl, ku = (cfg // 2, cfg % 2) if r > 0 else (cfg, 0)
                        if c == 0:
                            kl, ku = 0, cfg
                    else:
                        kl = ku = 0
                    t[n, cfg] = _mask_slice(t[n, cfg], mask, n, kl, ku)
            row.append(t.reshape((2 * nc,) + t.shape[2:]))
        out.append(row)
    return out


def literal_amplitude(config, tensors, occ, strategy):
    """compat.py:40-62 with _assemble_site's per-sample masks (model.py:157-212)."""
    mask = config.mask()
    n_rows, n_cols = config.shape
    z = occ.shape[0]
    dtype = np.asarray(tensors[0][0]).dtype
    boundary = P.ones_boundary(z, n_cols, dtype)
    for r in range(n_rows):
        mpo = []
        for c in range(n_cols):
            kl, ku = _neighbours(occ, r, c)
            n = occ[:, r, c]
            cfg = _cfg(n, kl, ku, r)
            sl = np.stack([_mask_slice(np.asarray(tensors[r][c])[n[i], cfg[i]], mask, n[i], kl[i], ku[i])
                           for i in range(z)])
            mpo.append(np.transpose(sl, (0, 3, 4, 1, 2)))
        boundary = strategy.apply(boundary, tuple(mpo))
    return np.where(valid(occ), P.contract_bottom(boundary), 0.0)


def value_and_grad(config, tensors, occ, strategy):
    """blockade_apply and its gradient: amplitude (0 for invalid samples),
    sliced dpsi/dA (env-gradients of the effective tensors, zero for invalid
    samples) and the combined slice indices."""
    ft = flat_tensors(config, tensors)
    amp, grad, _ = GI._amp_and_grad(ft, slices(config, occ), config.shape, strategy, False)
    ok = valid(occ)
    return np.where(ok, amp, 0.0), np.where(ok[:, None], grad, 0.0), active_slice_indices(config, occ)


def _row_mpo(ft, s, row, n_cols):
    return GI._row_mpo(ft, s, row, n_cols)


def _site_mpo(ft, shape, r, c, n, kl, ku):
    """MPO (z, l, r, u, d) of site (r, c) for occupation n and incoming kL, kU."""
    idx = n * nc_of(shape, r, c) + _cfg(n, kl, ku, r)
    return np.transpose(ft[r][c][idx], (0, 3, 4, 1, 2))


def can_flip_to_one(occ, r, c):
    """model.py:220-239."""
    n_rows, n_cols = occ.shape[1:]
    blocked = np.zeros(occ.shape[0], bool)
    if c > 0:
        blocked |= occ[:, r, c - 1] == 1
    if c < n_cols - 1:
        blocked |= occ[:, r, c + 1] == 1
    if r > 0:
        blocked |= occ[:, r - 1, c] == 1
    if r < n_rows - 1:
        blocked |= occ[:, r + 1, c] == 1
    return ~blocked


def _flip_mpos(ft, occ, r, c, n_flip, second_row):
    """MPOs of (r, c) with n_flip, of (r, c+1) with kL = n_flip and, for a row
    pair, of (r+1, c) with kU = n_flip (model.py:870-897)."""
    shape = occ.shape[1:]
    kl, ku = _neighbours(occ, r, c)
    m0c = _site_mpo(ft, shape, r, c, n_flip, kl, ku)
    m0c1 = None
    if c + 1 < shape[1]:
        _, ku1 = _neighbours(occ, r, c + 1)
        m0c1 = _site_mpo(ft, shape, r, c + 1, occ[:, r, c + 1], n_flip, ku1)
    m1c = None
    if second_row:
        kl1, _ = _neighbours(occ, r + 1, c)
        m1c = _site_mpo(ft, shape, r + 1, c, occ[:, r + 1, c], kl1, n_flip)
    return m0c, m1c, m0c1


def _amp_1row_2col(left, top, bot, m0, m1, right, c):
    """model.py:709-717."""
    return np.einsum("zace,zaub,zcduv,zevf,zbpg,zdhpq,zfqi,zghi->z", left, top[c], m0, bot[c], top[c + 1], m1,
                     bot[c + 1], right, optimize=_ROW2_PATH)


def _amp_1row_1col(left, top, bot, m0, c):
    """model.py:753-760 (right environment of ones)."""
    z = left.shape[0]
    one = np.ones((z, 1, 1, 1), dtype=left.dtype)
    return np.einsum("zace,zaub,zcduv,zevf,zbdf->z", left, top[c], m0, bot[c], one, optimize=_ROW1_PATH)


def _metropolis(amp_cur, amp_flip, u):
    ratio = P.metropolis_ratio(np.abs(amp_cur) ** 2, np.abs(amp_flip) ** 2)
    return u < np.minimum(1.0, ratio)


def _sweep_single_row(ft, occ, r, top, bot, seed, sweep, chains):
    """model.py:668-817."""
    n_cols = occ.shape[2]
    dtype = ft[0][0].dtype
    z = occ.shape[0]
    s = slices_of(occ)
    mpo = list(_row_mpo(ft, s, r, n_cols))
    renv = P.compute_right_envs(top, mpo, bot, dtype)
    left = np.ones((z, 1, 1, 1), dtype=dtype)
    for c in range(n_cols):
        n_cur = occ[:, r, c]
        n_flip = 1 - n_cur
        can = np.where(n_flip == 1, can_flip_to_one(occ, r, c), True)
        m0f, _, m1f = _flip_mpos(ft, occ, r, c, n_flip, False)
        if c + 1 < n_cols:
            right = renv[c + 1] if c + 1 < n_cols - 1 else np.ones((z, 1, 1, 1), dtype=dtype)
            amp_cur = _amp_1row_2col(left, top, bot, mpo[c], mpo[c + 1], right, c)
            amp_flip = np.where(can, _amp_1row_2col(left, top, bot, m0f, m1f, right, c), 0.0)
        else:
            amp_cur = _amp_1row_1col(left, top, bot, mpo[c], c)
            amp_flip = np.where(can, _amp_1row_1col(left, top, bot, m0f, c), 0.0)
        u, _ = R.sweep_draws(seed, sweep, BL_BASE + r * n_cols + c, chains)
        acc = _metropolis(amp_cur, amp_flip, u)
        a5 = acc[:, None, None, None, None]
        occ[:, r, c] = np.where(acc, n_flip, n_cur)
        mpo[c] = np.where(a5, m0f, mpo[c])
        if c + 1 < n_cols:
            mpo[c + 1] = np.where(a5, m1f, mpo[c + 1])
        left = P.left_env_step(left, top[c], mpo[c], bot[c])
    return occ


def _sweep_row_pair(ft, occ, r, top, bot, seed, sweep, chains):
    """model.py:820-994."""
    n_cols = occ.shape[2]
    dtype = ft[0][0].dtype
    z = occ.shape[0]
    s = slices_of(occ)
    mpo0 = list(_row_mpo(ft, s, r, n_cols))
    mpo1 = list(_row_mpo(ft, s, r + 1, n_cols))
    renv = P.compute_right_envs_2row(top, mpo0, mpo1, bot, dtype)
    left = np.ones((z, 1, 1, 1, 1), dtype=dtype)
    for c in range(n_cols):
        n_cur = occ[:, r, c]
        n_flip = 1 - n_cur
        can = np.where(n_flip == 1, can_flip_to_one(occ, r, c), True)
        m0f, m1f, m0f1 = _flip_mpos(ft, occ, r, c, n_flip, True)
        if c + 1 < n_cols:
            amp_cur = GI._plaquette_amp(left, top, bot, mpo0[c], mpo1[c], mpo0[c + 1], mpo1[c + 1], renv[c + 1], c)
            amp_flip = np.where(can, GI._plaquette_amp(left, top, bot, m0f, m1f, m0f1, mpo1[c + 1], renv[c + 1], c),
                                0.0)
        else:
            one = np.ones((z, 1, 1, 1, 1), dtype=dtype)
            amp_cur = GI._column_pair_amp(left, top, bot, mpo0[c], mpo1[c], one, c)
            amp_flip = np.where(can, GI._column_pair_amp(left, top, bot, m0f, m1f, one, c), 0.0)
        u, _ = R.sweep_draws(seed, sweep, BL_BASE + r * n_cols + c, chains)
        acc = _metropolis(amp_cur, amp_flip, u)
        a5 = acc[:, None, None, None, None]
        occ[:, r, c] = np.where(acc, n_flip, n_cur)
        mpo0[c] = np.where(a5, m0f, mpo0[c])
        mpo1[c] = np.where(a5, m1f, mpo1[c])
        if c + 1 < n_cols:
            mpo0[c + 1] = np.where(a5, m0f1, mpo0[c + 1])
        left = GI._left2_step(left, top[c], mpo0[c], mpo1[c], bot[c])
    return occ


def slices_of(occ):
    """slices() of the lattice being swept (the shape is the occupation array's)."""
    n_rows, n_cols = occ.shape[1:]
    out = np.zeros((occ.shape[0], n_rows * n_cols), np.int64)
    for r in range(n_rows):
        for c in range(n_cols):
            kl, ku = _neighbours(occ, r, c)
            n = occ[:, r, c]
            out[:, r * n_cols + c] = n * nc_of((n_rows, n_cols), r, c) + _cfg(n, kl, ku, r)
    return out


def init_cache(config, tensors, occ, strategy):
    """kernels.py:34-47: bottom boundaries of every row."""
    return P.init_cache(flat_tensors(config, tensors), slices(config, occ), config.shape, strategy)


def transition(config, tensors, occ, cache, strategy, seed, sweep, chains):
    """model.py:620-665: row pairs (r, r+1) for r < n_rows - 1 over the cached
    bottom boundaries, each row then joining the top boundary; the last row
    with the single-row sweep. Returns (occ, amp, top_envs)."""
    ft = flat_tensors(config, tensors)
    occ = np.array(occ, dtype=np.int64)
    n_rows, n_cols = config.shape
    z = occ.shape[0]
    dtype = ft[0][0].dtype
    top = P.ones_boundary(z, n_cols, dtype)
    tops = [None] * n_rows
    bots = list(cache.bottom_envs)
    for r in range(n_rows - 1):
        tops[r] = top
        occ = _sweep_row_pair(ft, occ, r, top, bots[r + 1], seed, sweep, chains)
        top = strategy.apply(top, tuple(_row_mpo(ft, slices_of(occ), r, n_cols)))
    tops[n_rows - 1] = top
    occ = _sweep_single_row(ft, occ, n_rows - 1, top, P.ones_boundary(z, n_cols, dtype), seed, sweep, chains)
    top = strategy.apply(top, tuple(_row_mpo(ft, slices_of(occ), n_rows - 1, n_cols)))
    return occ, P.contract_bottom(top), tops


def rydberg_terms(shape, omega, delta, v_nnn=0.0):
    """hamiltonian.py:22-79 as oracle terms: (Omega/2) sigma^x and -Delta n on
    every site, V_nnn n_i n_j on the diagonal and anti-diagonal neighbours."""
    n_rows, n_cols = shape
    terms = P.BucketedTerms(shape)
    sx = np.array([[0.0, 1.0], [1.0, 0.0]])
    for r in range(n_rows):
        for c in range(n_cols):
            terms.one_site[r][c].append(0.5 * omega * sx)
            terms.diagonal.append((((r, c),), np.array([0.0, -delta])))
    if v_nnn != 0.0:
        nn = np.array([0.0, 0.0, 0.0, v_nnn])
        for r in range(n_rows):
            for c in range(n_cols):
                if r + 1 < n_rows and c + 1 < n_cols:
                    terms.diagonal.append((((r, c), (r + 1, c + 1)), nn))
                if r + 1 < n_rows and c - 1 >= 0:
                    terms.diagonal.append((((r, c), (r + 1, c - 1)), nn))
    return terms


def estimate(config, tensors, occ, amp, top_envs, strategy, terms):
    """model.py:295-617: env-gradients, diagonal energies from the occupations
    and one-site energies sum_ops op[n_flip, n_cur] amp_flip / amp over two-row
    (or, on the last row, one-row) windows; returns (o, p, eloc, bottom_envs)."""
    ft = flat_tensors(config, tensors)
    n_rows, n_cols = config.shape
    z = occ.shape[0]
    dtype = ft[0][0].dtype
    s = slices(config, occ)
    energy = np.zeros(z, dtype=np.result_type(dtype, np.float64))
    for sites, diag in terms.diagonal:
        idx = np.zeros(z, np.int64)
        for rr, cc in sites:
            idx = idx * 2 + occ[:, rr, cc]
        energy = energy + diag[idx]
    grads = [[None] * n_cols for _ in range(n_rows)]
    bottoms = [None] * n_rows
    bottom = P.ones_boundary(z, n_cols, dtype)
    for row in range(n_rows - 1, -1, -1):
        bottoms[row] = bottom
        top = top_envs[row]
        mpo = _row_mpo(ft, s, row, n_cols)
        renv = P.compute_right_envs(top, mpo, bottom, dtype)
        left = np.ones((z, 1, 1, 1), dtype=dtype)
        has_x = any(terms.one_site[row][c] for c in range(n_cols))
        if has_x and row < n_rows - 1:
            bpair = bottoms[row + 1]
            mpo_next = _row_mpo(ft, s, row + 1, n_cols)
            renv2 = P.compute_right_envs_2row(top, mpo, mpo_next, bpair, dtype)
            left2 = np.ones((z, 1, 1, 1, 1), dtype=dtype)
        for c in range(n_cols):
            grads[row][c] = P.compute_single_gradient(left, renv[c], top[c], bottom[c])
            ops = terms.one_site[row][c]
            if ops:
                n_cur = occ[:, row, c]
                n_flip = 1 - n_cur
                can = np.where(n_flip == 1, can_flip_to_one(occ, row, c), True)
                if row < n_rows - 1:
                    m0f, m1f, m0f1 = _flip_mpos(ft, occ, row, c, n_flip, True)
                    if c + 1 < n_cols:
                        af = GI._plaquette_amp(left2, top, bpair, m0f, m1f, m0f1, mpo_next[c + 1], renv2[c + 1], c)
                    else:
                        one = np.ones((z, 1, 1, 1, 1), dtype=dtype)
                        af = GI._column_pair_amp(left2, top, bpair, m0f, m1f, one, c)
                else:
                    m0f, _, m0f1 = _flip_mpos(ft, occ, row, c, n_flip, False)
                    if c + 1 < n_cols:
                        right = renv[c + 1] if c + 1 < n_cols - 1 else np.ones((z, 1, 1, 1), dtype=dtype)
                        af = _amp_1row_2col(left, top, bottom, m0f, m0f1, right, c)
                    else:
                        af = _amp_1row_1col(left, top, bottom, m0f, c)
                af = np.where(can, af, 0.0)
                for op in ops:
                    energy = energy + op[n_flip, n_cur] * af / amp
            left = P.left_env_step(left, top[c], mpo[c], bottom[c])
            if has_x and row < n_rows - 1:
                left2 = GI._left2_step(left2, top[c], mpo[c], mpo_next[c], bpair[c])
        bottom = P.apply_mpo_from_below(bottom, mpo, strategy)
    o = np.concatenate([grads[r][c].reshape(z, -1) for r in range(n_rows) for c in range(n_cols)], axis=1)
    o = o / amp[:, None]
    return o, active_slice_indices(config, occ), energy, bottoms


def mc_sampler(config, tensors, occ, strategy, terms, seed, n_steps, sweep0=0):
    """make_mc_sampler over build_mc_kernels (kernels.py:17-117): cache from the
    initial occupations, then per step transition + estimate; returns the
    per-step samples, log-derivatives, slice indices and local energies."""
    occ = np.array(occ, dtype=np.int64)
    z = occ.shape[0]
    chains = np.arange(z)
    cache = init_cache(config, tensors, occ, strategy)
    hist = {"samples": [], "o": [], "p": [], "eloc": [], "amp": []}
    for t in range(n_steps):
        occ, amp, tops = transition(config, tensors, occ, cache, strategy, seed, sweep0 + t, chains)
        o, p, e, bottoms = estimate(config, tensors, occ, amp, tops, strategy, terms)
        cache = P.Cache(bottom_envs=tuple(bottoms))
        hist["samples"].append(occ.reshape(z, -1).copy())
        hist["o"].append(o)
        hist["p"].append(p)
        hist["eloc"].append(e)
        hist["amp"].append(amp)
    return occ, cache, hist
