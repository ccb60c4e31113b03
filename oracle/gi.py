"""FP64 restatement of the gauge-invariant PEPS path (src/vmc/peps/gi/*).

Test infrastructure only: imported by tests/, never by the product path.

Follows
* src/vmc/peps/gi/model.py: GIPEPSConfig (:40-64), tensor shapes (:99-118),
  sample layout (:121-141), random configurations (:153-182), cfg index
  (:308-343), site assembly (:276-305), row MPOs (:350-366), plaquette flips
  (:369-384), estimate (:392-612), charge index map (:190-217), the
  plaquette sweep (:623-700), the horizontal / vertical link sweeps with
  matter (:703-878) and transition (:901-1043);
* src/vmc/peps/gi/compat.py:60-99 (gi_apply: Gauss-law check, zero amplitude
  for an invalid sample);
* src/vmc/peps/gi/kernels.py:90-126 (sliced log-derivatives, combined slice
  index sigma * nc + cfg);
* src/vmc/peps/gi/local_terms.py (LinkDiagonalTerm, build_electric_terms).

The effective tensor of a site is one (d, mu_u, mu_d, mu_l, mu_r) slice of the
stored (d, nc, ...) tensor, so amplitudes and gradients are the standard
PEPS ones (oracle.peps) over tensors reshaped to (d * nc, ...) and indexed by
the combined slice index. A non-uniform degeneracy_per_charge masks each
slice once with the charges of its own links (flat_tensors); on Gauss-law-valid
samples, the only ones the sampler and estimate evaluate, these are the
per-sample masks of _assemble_site (literal_amplitude applies those, and the
tests check the two agree).

Draws (the reference splits JAX threefry keys) come from oracle.rng Philox:
* the plaquette (r, c) of sweep t for chain b uses counter
  (PLAQ_BASE + r * n_cols + c, t, b, 0); its uniform is the accept draw and its
  word gives delta = 1 + word % (N - 1);
* the horizontal link (r, c) uses HLINK_BASE + r * (n_cols - 1) + c, the
  vertical link (r, c) VLINK_BASE + r * n_cols + c, in the first counter word:
  block (.., 0) gives the accept uniform and delta as above, block (.., 1) the
  two uniforms that pick the new matter index of the link's first and second
  site among the physical indices of their new charge.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from oracle import peps as P
from oracle import rng as R

PLAQ_BASE = 1 << 20
HLINK_BASE = 2 << 20
VLINK_BASE = 3 << 20


@dataclass(frozen=True)
class GIConfig:
    """gi/model.py:40-64."""

    shape: tuple
    N: int
    phys_dim: int
    Qx: int
    degeneracy_per_charge: tuple
    charge_of_site: tuple

    @property
    def dmax(self):
        return int(max(self.degeneracy_per_charge))

    def mask(self):
        """mask_per_charge (gi/model.py:52-60): (N, dmax) bool, row q keeps the
        first degeneracy_per_charge[q] bond indices; None when all are dmax."""
        dm = self.dmax
        if all(d == dm for d in self.degeneracy_per_charge):
            return None
        return np.arange(dm)[None, :] < np.asarray(self.degeneracy_per_charge)[:, None]


def n_links(shape, r, c):
    n_rows, n_cols = shape
    return int(r > 0) + int(r < n_rows - 1) + int(c > 0) + int(c < n_cols - 1)


def nc_of(config, r, c):
    """gi/model.py:101-103: feasible local link configurations."""
    return int(config.N ** max(n_links(config.shape, r, c) - 1, 0))


def random_gi_tensors(config, seed=0, dtype=np.float64):
    """Site tensors (d, nc, mu_u, mu_d, mu_l, mu_r), i.i.d. N(0,1)/sqrt(dmax^4 d)."""
    rng = np.random.default_rng(seed)
    n_rows, n_cols = config.shape
    scale = 1.0 / np.sqrt(config.dmax ** 4 * config.phys_dim)
    out = []
    for r in range(n_rows):
        row = []
        for c in range(n_cols):
            dm = config.dmax
            shp = (config.phys_dim, nc_of(config, r, c), dm if r > 0 else 1, dm if r < n_rows - 1 else 1,
                   dm if c > 0 else 1, dm if c < n_cols - 1 else 1)
            t = rng.standard_normal(shp) * scale
            if np.issubdtype(dtype, np.complexfloating):
                t = t + 1j * rng.standard_normal(shp) * scale
            row.append(t.astype(dtype))
        out.append(row)
    return out


def n_sample(shape):
    n_rows, n_cols = shape
    return n_rows * n_cols + n_rows * (n_cols - 1) + (n_rows - 1) * n_cols


def unflatten(samples, shape):
    """gi/model.py:131-141 (batched)."""
    n_rows, n_cols = shape
    s = np.asarray(samples).reshape(-1, n_sample(shape))
    ns, nh = n_rows * n_cols, n_rows * (n_cols - 1)
    return (s[:, :ns].reshape(-1, n_rows, n_cols), s[:, ns:ns + nh].reshape(-1, n_rows, n_cols - 1),
            s[:, ns + nh:].reshape(-1, n_rows - 1, n_cols))


def flatten(sites, h, v):
    z = sites.shape[0]
    return np.concatenate([sites.reshape(z, -1), h.reshape(z, -1), v.reshape(z, -1)], axis=1).astype(np.int32)


def random_configurations(config, n, seed=0):
    """gi/model.py:153-182: links from random plaquette deltas on the all-zero
    configuration, then matter sites drawn uniformly within their Gauss charge."""
    rng = np.random.default_rng(seed)
    n_rows, n_cols = config.shape
    N = config.N
    h = np.zeros((n, n_rows, n_cols - 1), np.int64)
    v = np.zeros((n, n_rows - 1, n_cols), np.int64)
    if n_rows > 1 and n_cols > 1:
        dl = rng.integers(0, N, size=(n, n_rows - 1, n_cols - 1))
        h[:, :n_rows - 1, :] += dl
        h[:, 1:, :] -= dl
        v[:, :, :n_cols - 1] += dl
        v[:, :, 1:] -= dl
        h %= N
        v %= N
    div = _divergence(h, v, N)
    charge = (config.Qx - div) % N
    cos = np.asarray(config.charge_of_site) % N
    sites = np.zeros((n, n_rows, n_cols), np.int64)
    for q in range(N):
        idx = np.nonzero(cos == q)[0]
        mask = charge == q
        if mask.any():
            if len(idx) == 0:
                raise ValueError("charge_of_site must include all charges")
            sites[mask] = idx[rng.integers(0, len(idx), size=int(mask.sum()))]
    return flatten(sites, h, v)


def _divergence(h, v, N):
    """gi/compat.py:79-84: (n_left + n_down - n_up - n_right) mod N."""
    nl = np.pad(h, ((0, 0), (0, 0), (1, 0)))
    nr = np.pad(h, ((0, 0), (0, 0), (0, 1)))
    nu = np.pad(v, ((0, 0), (1, 0), (0, 0)))
    nd = np.pad(v, ((0, 0), (0, 1), (0, 0)))
    return (nl + nd - nu - nr) % N


def gauss_valid(config, samples):
    """gi/compat.py:79-88 -> (z,) bool."""
    sites, h, v = unflatten(samples, config.shape)
    div = _divergence(h, v, config.N)
    charge = np.asarray(config.charge_of_site)[sites]
    return np.all((div + charge) % config.N == config.Qx % config.N, axis=(1, 2))


def _link(h, v, r, c, direction):
    """gi/model.py:217-257 (batched)."""
    z = h.shape[0]
    if direction == "left":
        return h[:, r, c - 1] if c > 0 else np.zeros(z, np.int64)
    if direction == "right":
        return h[:, r, c] if c < h.shape[2] else np.zeros(z, np.int64)
    if direction == "up":
        return v[:, r - 1, c] if r > 0 else np.zeros(z, np.int64)
    return v[:, r, c] if r < v.shape[1] else np.zeros(z, np.int64)


def site_cfg_index(config, h, v, r, c):
    """gi/model.py:308-343: the first active link in (right, down, up, left)
    is fixed by Gauss law; the others, in (left, up, down, right) order, are
    the base-N digits of the index."""
    n_rows, n_cols = config.shape
    active = {"left": c > 0, "right": c < n_cols - 1, "up": r > 0, "down": r < n_rows - 1}
    dependent = next((dname for dname in ("right", "down", "up", "left") if active[dname]), None)
    idx = np.zeros(h.shape[0], np.int64)
    for dname in ("left", "up", "down", "right"):
        if not active[dname] or dname == dependent:
            continue
        idx = idx * config.N + _link(h, v, r, c, dname)
    return idx


def slices(config, samples):
    """Combined slice index sigma * nc + cfg of every site, (z, n_sites)
    (gi/kernels.py:116-118)."""
    sites, h, v = unflatten(samples, config.shape)
    n_rows, n_cols = config.shape
    out = np.zeros((sites.shape[0], n_rows * n_cols), np.int64)
    for r in range(n_rows):
        for c in range(n_cols):
            out[:, r * n_cols + c] = sites[:, r, c] * nc_of(config, r, c) + site_cfg_index(config, h, v, r, c)
    return out


def slice_links(config, r, c, sigma, cfg):
    """Link values (k_l, k_u, k_d, k_r) of slice (sigma, cfg) of site (r, c):
    the inverse of site_cfg_index, with the dependent link solved from Gauss
    law (k_l + k_d - k_u - k_r + charge(sigma) = Qx mod N, gi/compat.py:79-88);
    inactive (boundary) links are 0."""
    n_rows, n_cols = config.shape
    N = config.N
    active = {"left": c > 0, "right": c < n_cols - 1, "up": r > 0, "down": r < n_rows - 1}
    dependent = next((dname for dname in ("right", "down", "up", "left") if active[dname]), None)
    k = {"left": 0, "up": 0, "down": 0, "right": 0}
    rest = cfg
    for dname in ("right", "down", "up", "left"):
        if active[dname] and dname != dependent:
            k[dname] = rest % N
            rest //= N
    q = int(config.charge_of_site[sigma]) % N
    qx = config.Qx % N
    if dependent == "right":
        k["right"] = (k["left"] + k["down"] - k["up"] + q - qx) % N
    elif dependent == "down":
        k["down"] = (qx - q - k["left"] + k["up"] + k["right"]) % N
    elif dependent == "up":
        k["up"] = (k["left"] + k["down"] - k["right"] + q - qx) % N
    elif dependent == "left":
        k["left"] = (qx - q - k["down"] + k["up"] + k["right"]) % N
    return k["left"], k["up"], k["down"], k["right"]


def _mask_gi(t, mask, kl, ku, kd, kr):
    """gi/model.py:292-305 on one (u, d, l, r) slice."""
    u, d, l, r = t.shape
    return (t * mask[ku][:u][:, None, None, None] * mask[kd][:d][None, :, None, None]
            * mask[kl][:l][None, None, :, None] * mask[kr][:r][None, None, None, :])


def flat_tensors(tensors, config=None):
    """(d, nc, u, dn, l, r) -> (d * nc, u, dn, l, r): the combined slice index
    selects gi/model.py:_assemble_site's tensor[sigma, cfg]. With a
    non-uniform degeneracy_per_charge every slice carries the charge masks of
    its own links (slice_links): on Gauss-law-valid samples those are the
    links _assemble_site masks with."""
    mask = None if config is None else config.mask()
    out = []
    for r, row in enumerate(tensors):
        orow = []
        for c, t in enumerate(row):
            t = np.array(t)
            if mask is not None:
                for sigma in range(t.shape[0]):
                    for cfg in range(t.shape[1]):
                        kl, ku, kd, kr = slice_links(config, r, c, sigma, cfg)
                        t[sigma, cfg] = _mask_gi(t[sigma, cfg], mask, kl, ku, kd, kr)
            orow.append(t.reshape((t.shape[0] * t.shape[1],) + t.shape[2:]))
        out.append(orow)
    return out


def literal_amplitude(config, tensors, samples, strategy):
    """gi_apply with _assemble_site's per-sample masks (gi/model.py:276-305,
    gi/compat.py:60-99): 0 for Gauss-law violations."""
    sites, h, v = unflatten(samples, config.shape)
    n_rows, n_cols = config.shape
    mask = config.mask()
    z = sites.shape[0]
    dtype = np.asarray(tensors[0][0]).dtype
    boundary = P.ones_boundary(z, n_cols, dtype)
    for r in range(n_rows):
        mpo = []
        for c in range(n_cols):
            cfg = site_cfg_index(config, h, v, r, c)
            kl, kr = _link(h, v, r, c, "left"), _link(h, v, r, c, "right")
            ku, kd = _link(h, v, r, c, "up"), _link(h, v, r, c, "down")
            sl = []
            for i in range(z):
                t = np.asarray(tensors[r][c])[sites[i, r, c], cfg[i]]
                sl.append(t if mask is None else _mask_gi(t, mask, kl[i], ku[i], kd[i], kr[i]))
            mpo.append(np.transpose(np.stack(sl), (0, 3, 4, 1, 2)))
        boundary = strategy.apply(boundary, tuple(mpo))
    return np.where(gauss_valid(config, samples), P.contract_bottom(boundary), 0.0)


def _amp_and_grad(ft, s, shape, strategy, full_gradient, grad=True):
    """compat.py:94-133 over combined slice indices. Unlike the spin evaluator
    there is no spin_to_occupancy step: gi_apply feeds the indices directly
    (gi/compat.py:18-31, 90-95)."""
    spins = np.asarray(s).reshape(-1, *shape)
    amp, top_envs = P.forward_with_cache(ft, spins, strategy)
    if not grad:
        return amp
    z = spins.shape[0]
    n_rows, n_cols = shape
    env_grads = P.compute_all_gradients(ft, spins, strategy, top_envs)
    parts, p_parts = [], []
    for r in range(n_rows):
        for c in range(n_cols):
            g = env_grads[r][c].reshape(z, -1)
            if full_gradient:
                full = np.zeros((z, ft[r][c].shape[0], g.shape[1]), dtype=g.dtype)
                full[np.arange(z), spins[:, r, c]] = g
                parts.append(full.reshape(z, -1))
            else:
                parts.append(g)
                p_parts.append(np.repeat(spins[:, r, c][:, None], g.shape[1], axis=1).astype(np.int8))
    return amp, np.concatenate(parts, axis=1), (None if full_gradient else np.concatenate(p_parts, axis=1))


def gi_value(tensors, samples, config, strategy):
    """gi/compat.py:60-99 (batched): 0 for samples violating Gauss law."""
    amp = _amp_and_grad(flat_tensors(tensors, config), slices(config, samples), config.shape, strategy, False, grad=False)
    return np.where(gauss_valid(config, samples), amp, 0.0)


def gi_value_and_grad(tensors, samples, config, strategy, full_gradient=False):
    """Amplitude, dpsi/dA (sliced or full) and the combined slice indices
    (the custom VJP of gi/compat.py:34-56; sliced layout of gi/kernels.py:104-118)."""
    return _amp_and_grad(flat_tensors(tensors, config), slices(config, samples), config.shape, strategy, full_gradient)


def _row_mpo(ft, s, row, n_cols):
    """gi/model.py:350-366: per-sample MPOs (z, l, r, u, d) of one row."""
    out = []
    for c in range(n_cols):
        sl = ft[row][c][s[:, row * n_cols + c]]
        out.append(np.transpose(sl, (0, 3, 4, 1, 2)))
    return out


def _plaquette_flip(h, v, r, c, delta, N):
    """gi/model.py:369-384 (batched delta)."""
    h = h.copy()
    v = v.copy()
    h[:, r, c] = (h[:, r, c] + delta) % N
    h[:, r + 1, c] = (h[:, r + 1, c] - delta) % N
    v[:, r, c] = (v[:, r, c] + delta) % N
    v[:, r, c + 1] = (v[:, r, c + 1] - delta) % N
    return h, v


_PLAQ_PATH = P._path((1, 5), (3, 6), (1, 2), (1, 2), (0, 2), (2, 4), (1, 3), (0, 2), (0, 1))


def _plaquette_amp(left, top, bot, m0c, m1c, m0n, m1n, renv, c):
    """The two-column amplitude of gi/model.py:669-677 / 548-558."""
    return np.einsum(
        "zalxe,zaub,zlruv,zxyvw,zewf,zbgc,zrsgh,zythi,zfij,zcstj->z",
        left, top[c], m0c, m1c, bot[c], top[c + 1], m0n, m1n, bot[c + 1], renv,
        optimize=_PLAQ_PATH,
    )


def _left2_step(left, top, m0, m1, bot):
    return np.einsum(
        "zalxe,zaub,zlruv,zxyvw,zewf->zbryf", left, top, m0, m1, bot,
        optimize=P._path((0, 1), (0, 3), (0, 2), (0, 1)),
    )


def _flipped_mpos(ft, config, sites, h, v, r, c):
    """MPOs of the four plaquette sites for links (h, v)."""
    n_cols = config.shape[1]
    z = sites.shape[0]
    out = []
    for rr, cc in ((r, c), (r, c + 1), (r + 1, c), (r + 1, c + 1)):
        idx = sites[:, rr, cc] * nc_of(config, rr, cc) + site_cfg_index(config, h, v, rr, cc)
        out.append(np.transpose(ft[rr][cc][idx], (0, 3, 4, 1, 2)))
    del n_cols, z
    return out


def _plaquette_sweep_pair(ft, config, sites, h, v, top, bot, mpo0, mpo1, r, seed, sweep, chains):
    """gi/model.py:623-700 for row pair (r, r + 1)."""
    n_cols = config.shape[1]
    N = config.N
    dtype = mpo0[0].dtype
    z = sites.shape[0]
    renv = P.compute_right_envs_2row(top, mpo0, mpo1, bot, dtype)
    left = np.ones((z, 1, 1, 1, 1), dtype=dtype)
    mpo0 = list(mpo0)
    mpo1 = list(mpo1)
    for c in range(n_cols - 1):
        u, word = R.sweep_draws(seed, sweep, PLAQ_BASE + r * n_cols + c, chains)
        delta = 1 + (word % np.uint32(max(N - 1, 1))).astype(np.int64)
        amp_cur = _plaquette_amp(left, top, bot, mpo0[c], mpo1[c], mpo0[c + 1], mpo1[c + 1], renv[c + 1], c)
        hp, vp = _plaquette_flip(h, v, r, c, delta, N)
        m00, m01, m10, m11 = _flipped_mpos(ft, config, sites, hp, vp, r, c)
        amp_prop = _plaquette_amp(left, top, bot, m00, m10, m01, m11, renv[c + 1], c)
        ratio = P.metropolis_ratio(np.abs(amp_cur) ** 2, np.abs(amp_prop) ** 2)
        acc = u < np.minimum(1.0, ratio)
        a5 = acc[:, None, None, None, None]
        mpo0[c] = np.where(a5, m00, mpo0[c])
        mpo0[c + 1] = np.where(a5, m01, mpo0[c + 1])
        mpo1[c] = np.where(a5, m10, mpo1[c])
        mpo1[c + 1] = np.where(a5, m11, mpo1[c + 1])
        h = np.where(acc[:, None, None], hp, h)
        v = np.where(acc[:, None, None], vp, v)
        left = _left2_step(left, top[c], mpo0[c], mpo1[c], bot[c])
    return mpo0, mpo1, h, v


def charge_index_map(config):
    """gi/model.py:190-206: physical indices of each charge (ascending, -1
    padded) and their count."""
    q = np.asarray(config.charge_of_site, np.int64) % config.N
    deg = np.bincount(q, minlength=config.N)
    table = np.full((config.N, max(int(deg.max()), 1)), -1, np.int64)
    for c in range(config.N):
        idx = np.nonzero(q == c)[0]
        table[c, : len(idx)] = idx
    return table, deg


def _uniform53(x0, x1):
    u53 = (x0.astype(np.uint64) << np.uint64(21)) | (x1.astype(np.uint64) >> np.uint64(11))
    return u53.astype(np.float64) * (2.0 ** -53)


def link_draws(seed, sweep, ctr0, chains):
    """Accept uniform, delta word and the two matter-choice uniforms of one link move."""
    chains = np.asarray(chains, dtype=np.uint32)
    ctr = np.zeros(chains.shape + (4,), dtype=np.uint32)
    ctr[..., 0] = np.uint32(ctr0)
    ctr[..., 1] = np.uint32(sweep)
    ctr[..., 2] = chains
    key = R.seed_key(seed)[None, :]
    x = R.philox4x32(ctr, key)
    ctr[..., 3] = 1
    y = R.philox4x32(ctr, key)
    return _uniform53(x[..., 0], x[..., 1]), x[..., 2], _uniform53(y[..., 0], y[..., 1]), _uniform53(y[..., 2], y[..., 3])


def _pick(table, deg, q, u):
    """gi/model.py:209-217: table[q, floor(u * deg[q])]."""
    k = np.floor(u * deg[q]).astype(np.int64)
    return table[q, k]


def _site_mpo(ft, config, sites, h, v, r, c):
    idx = sites[:, r, c] * nc_of(config, r, c) + site_cfg_index(config, h, v, r, c)
    return np.transpose(ft[r][c][idx], (0, 3, 4, 1, 2))


_H2_PATH = P._path((0, 1), (0, 6), (0, 5), (0, 3), (1, 2), (1, 2), (0, 1))


def _two_site_amp(left, top, bot, m0, m1, renv, c):
    """gi/model.py:734-739: a horizontal pair of sites between boundaries."""
    return np.einsum(
        "zace,zaub,zcduv,zevf,zbgh,zdigw,zfwj,zhij->z",
        left, top[c], m0, bot[c], top[c + 1], m1, bot[c + 1], renv,
        optimize=_H2_PATH,
    )


_V1_PATH = P._path((0, 1), (0, 4), (1, 2), (1, 2), (0, 1))


def _column_pair_amp(left, top, bot, m0, m1, renv, c):
    """gi/model.py:821-825: one column of a row pair."""
    return np.einsum(
        "zalxe,zaub,zlruv,zxyvw,zewf,zbryf->z", left, top[c], m0, m1, bot[c], renv, optimize=_V1_PATH
    )


def _hlink_sweep_row(ft, config, sites, h, v, top, bot, mpo, r, seed, sweep, chains):
    """gi/model.py:703-792: horizontal links of row r with matter moves."""
    n_cols = config.shape[1]
    N = config.N
    charge = np.asarray(config.charge_of_site, np.int64) % N
    table, deg = charge_index_map(config)
    dtype = mpo[0].dtype
    z = sites.shape[0]
    renv = P.compute_right_envs(top, mpo, bot, dtype)
    left = np.ones((z, 1, 1, 1), dtype=dtype)
    mpo = list(mpo)
    for c in range(n_cols - 1):
        u, word, ua, ub = link_draws(seed, sweep, HLINK_BASE + r * (n_cols - 1) + c, chains)
        delta = 1 + (word % np.uint32(max(N - 1, 1))).astype(np.int64)
        amp_cur = _two_site_amp(left, top, bot, mpo[c], mpo[c + 1], renv[c + 1], c)
        hp = h.copy()
        hp[:, r, c] = (h[:, r, c] + delta) % N
        qa, qb = charge[sites[:, r, c]], charge[sites[:, r, c + 1]]
        qa_new, qb_new = (qa + delta) % N, (qb - delta) % N
        sp = sites.copy()
        sp[:, r, c] = _pick(table, deg, qa_new, ua)
        sp[:, r, c + 1] = _pick(table, deg, qb_new, ub)
        m0 = _site_mpo(ft, config, sp, hp, v, r, c)
        m1 = _site_mpo(ft, config, sp, hp, v, r, c + 1)
        amp_prop = _two_site_amp(left, top, bot, m0, m1, renv[c + 1], c)
        ratio = P.metropolis_ratio(np.abs(amp_cur) ** 2, np.abs(amp_prop) ** 2)
        ratio = ratio * (deg[qa_new] * deg[qb_new]) / (deg[qa] * deg[qb])
        acc = u < np.minimum(1.0, ratio)
        a5 = acc[:, None, None, None, None]
        mpo[c] = np.where(a5, m0, mpo[c])
        mpo[c + 1] = np.where(a5, m1, mpo[c + 1])
        h = np.where(acc[:, None, None], hp, h)
        sites = np.where(acc[:, None, None], sp, sites)
        left = P.left_env_step(left, top[c], mpo[c], bot[c])
    return mpo, sites, h


def _vlink_sweep_pair(ft, config, sites, h, v, top, bot, mpo0, mpo1, r, seed, sweep, chains):
    """gi/model.py:795-878: vertical links between rows r and r + 1 with matter moves."""
    n_cols = config.shape[1]
    N = config.N
    charge = np.asarray(config.charge_of_site, np.int64) % N
    table, deg = charge_index_map(config)
    dtype = mpo0[0].dtype
    z = sites.shape[0]
    renv = P.compute_right_envs_2row(top, mpo0, mpo1, bot, dtype)
    left = np.ones((z, 1, 1, 1, 1), dtype=dtype)
    mpo0 = list(mpo0)
    mpo1 = list(mpo1)
    for c in range(n_cols):
        u, word, ua, ub = link_draws(seed, sweep, VLINK_BASE + r * n_cols + c, chains)
        delta = 1 + (word % np.uint32(max(N - 1, 1))).astype(np.int64)
        amp_cur = _column_pair_amp(left, top, bot, mpo0[c], mpo1[c], renv[c], c)
        vp = v.copy()
        vp[:, r, c] = (v[:, r, c] + delta) % N
        qa, qb = charge[sites[:, r, c]], charge[sites[:, r + 1, c]]
        qa_new, qb_new = (qa - delta) % N, (qb + delta) % N
        sp = sites.copy()
        sp[:, r, c] = _pick(table, deg, qa_new, ua)
        sp[:, r + 1, c] = _pick(table, deg, qb_new, ub)
        m0 = _site_mpo(ft, config, sp, h, vp, r, c)
        m1 = _site_mpo(ft, config, sp, h, vp, r + 1, c)
        amp_prop = _column_pair_amp(left, top, bot, m0, m1, renv[c], c)
        ratio = P.metropolis_ratio(np.abs(amp_cur) ** 2, np.abs(amp_prop) ** 2)
        ratio = ratio * (deg[qa_new] * deg[qb_new]) / (deg[qa] * deg[qb])
        acc = u < np.minimum(1.0, ratio)
        a5 = acc[:, None, None, None, None]
        mpo0[c] = np.where(a5, m0, mpo0[c])
        mpo1[c] = np.where(a5, m1, mpo1[c])
        v = np.where(acc[:, None, None], vp, v)
        sites = np.where(acc[:, None, None], sp, sites)
        left = _left2_step(left, top[c], mpo0[c], mpo1[c], bot[c])
    return mpo0, mpo1, sites, v


def _bottom_envs(ft, config, sites, h, v, strategy):
    """gi/model.py:880-897."""
    return list(P.init_cache(ft, slices(config, flatten(sites, h, v)), config.shape, strategy).bottom_envs)


def init_cache(tensors, samples, config, strategy):
    """gi/kernels.py:33-47: bottom boundaries of every row."""
    return P.init_cache(flat_tensors(tensors, config), slices(config, samples), config.shape, strategy)


def transition(tensors, samples, cache, config, strategy, seed, sweep, chain_ids):
    """gi/model.py:901-1043: plaquette sweep over the row pairs; with matter
    (phys_dim > 1) then horizontal link sweeps of every row and vertical link
    sweeps of every row pair, each stage over freshly recomputed bottom
    boundaries; then the final top contraction."""
    n_rows, n_cols = config.shape
    ft = flat_tensors(tensors, config)
    sites, h, v = unflatten(samples, config.shape)
    sites, h, v = sites.copy(), h.copy(), v.copy()
    z = sites.shape[0]
    dtype = ft[0][0].dtype
    pure = config.phys_dim == 1
    top_cache = [None] * n_rows

    def mpo_row(row):
        return _row_mpo(ft, slices(config, flatten(sites, h, v)), row, n_cols)

    top = None
    if n_rows > 1:
        top = P.ones_boundary(z, n_cols, dtype)
        mpo0, mpo1 = mpo_row(0), mpo_row(1)
        for r in range(n_rows - 1):
            if pure:
                top_cache[r] = top
            mpo0, mpo1, h, v = _plaquette_sweep_pair(ft, config, sites, h, v, top, cache.bottom_envs[r + 1], mpo0,
                                                     mpo1, r, seed, sweep, chain_ids)
            top = strategy.apply(top, tuple(mpo0))
            if r + 2 < n_rows:
                mpo0 = mpo1
                mpo1 = mpo_row(r + 2)
        if pure:
            top_cache[n_rows - 1] = top
        top = strategy.apply(top, tuple(mpo1))
    if pure:
        if top is None:
            top = P.ones_boundary(z, n_cols, dtype)
            for row in range(n_rows):
                top_cache[row] = top
                top = strategy.apply(top, tuple(mpo_row(row)))
        return flatten(sites, h, v), P.Context(amp=P.contract_bottom(top), top_envs=tuple(top_cache))

    # horizontal link sweeps (gi/model.py:964-992)
    bottom = _bottom_envs(ft, config, sites, h, v, strategy)
    top = P.ones_boundary(z, n_cols, dtype)
    for row in range(n_rows):
        if n_rows == 1:
            top_cache[row] = top
        mpo, sites, h = _hlink_sweep_row(ft, config, sites, h, v, top, bottom[row], mpo_row(row), row, seed, sweep,
                                         chain_ids)
        top = strategy.apply(top, tuple(mpo))
    if n_rows == 1:
        return flatten(sites, h, v), P.Context(amp=P.contract_bottom(top), top_envs=tuple(top_cache))

    # vertical link sweeps (gi/model.py:994-1036)
    bottom = _bottom_envs(ft, config, sites, h, v, strategy)
    top = P.ones_boundary(z, n_cols, dtype)
    mpo0, mpo1 = mpo_row(0), mpo_row(1)
    for r in range(n_rows - 1):
        top_cache[r] = top
        mpo0, mpo1, sites, v = _vlink_sweep_pair(ft, config, sites, h, v, top, bottom[r + 1], mpo0, mpo1, r, seed,
                                                 sweep, chain_ids)
        top = strategy.apply(top, tuple(mpo0))
        if r + 2 < n_rows:
            mpo0 = mpo1
            mpo1 = mpo_row(r + 2)
    top_cache[n_rows - 1] = top
    top = strategy.apply(top, tuple(mpo1))
    return flatten(sites, h, v), P.Context(amp=P.contract_bottom(top), top_envs=tuple(top_cache))


@dataclass
class GITerms:
    """Plaquettes (gi/model.py:546-608, coefficient per top-left corner),
    link-diagonal terms (local_terms.py:13-36), diagonal terms on matter
    states (gi/model.py:422-431; MatterMassTerm, local_terms.py:39-52) and
    one-site / two-site operators on matter states with the links held fixed
    (gi/model.py:474-505, 517-536; operators applied as op[:, s])."""

    plaquette: dict          # (r, c) -> summed coefficient
    links_h: list            # (r, c, diag[N])
    links_v: list
    site_diag: tuple = ()    # (((r, c), ...), diag[phys_dim ** n_sites])
    one_site: dict = field(default_factory=dict)    # (r, c) -> [d x d operators]
    horizontal: dict = field(default_factory=dict)  # (r, c) -> [d^2 x d^2] on (r, c), (r, c+1)
    vertical: dict = field(default_factory=dict)    # (r, c) -> [d^2 x d^2] on (r, c), (r+1, c)


def electric_terms(shape, coeff, N):
    """local_terms.py:74-90: coeff * (2 - 2 cos(2 pi k / N)) on every link."""
    n_rows, n_cols = shape
    diag = coeff * (2.0 - 2.0 * np.cos(2.0 * np.pi * np.arange(N) / N))
    return ([(r, c, diag) for r in range(n_rows) for c in range(n_cols - 1)],
            [(r, c, diag) for r in range(n_rows - 1) for c in range(n_cols)])


def _eff(ft, config, sites, h, v, r, c):
    """gi/model.py:276-305 for every matter state p of site (r, c) with the
    sample's links: the slices p * nc + cfg, (z, d, up, down, left, right)."""
    nc = nc_of(config, r, c)
    cfg = site_cfg_index(config, h, v, r, c)
    idx = np.arange(config.phys_dim)[None, :] * nc + cfg[:, None]
    return ft[r][c][idx]


def _apply_ops(ops, idx, amps, amp):
    """sum over ops of op[:, s] . amps / amp (energy.py:88, 179, 200 convention)."""
    e = 0.0
    for op in ops:
        e = e + np.einsum("zp,zp->z", np.asarray(op)[:, idx].T, amps) / amp
    return e


def estimate(tensors, samples, context, config, strategy, terms: GITerms):
    """gi/model.py:392-612: env gradients, E_loc (link-diagonal, matter-diagonal,
    matter one-site / two-site and plaquette terms) and the new bottom
    boundaries."""
    n_rows, n_cols = config.shape
    N = config.N
    ft = flat_tensors(tensors, config)
    sites, h, v = unflatten(samples, config.shape)
    s = slices(config, samples)
    z = sites.shape[0]
    dtype = ft[0][0].dtype
    amp = context.amp
    energy = np.zeros(z, dtype=np.result_type(amp.dtype, dtype))
    for r, c, diag in terms.links_h:
        energy = energy + diag[h[:, r, c]]
    for r, c, diag in terms.links_v:
        energy = energy + diag[v[:, r, c]]
    for tsites, diag in terms.site_diag:
        idx = np.zeros(z, np.int64)
        for r, c in tsites:
            idx = idx * config.phys_dim + sites[:, r, c]
        energy = energy + np.asarray(diag)[idx]
    grads = [[None] * n_cols for _ in range(n_rows)]
    bottom_cache = [None] * n_rows
    bottom = P.ones_boundary(z, n_cols, dtype)
    for row in range(n_rows - 1, -1, -1):
        bottom_cache[row] = bottom
        top = context.top_envs[row]
        mpo = _row_mpo(ft, s, row, n_cols)
        one_here = [terms.one_site.get((row, c), []) for c in range(n_cols)]
        h_here = [terms.horizontal.get((row, c), []) for c in range(n_cols - 1)]
        if any(one_here) or any(h_here):
            # gi/model.py:463-512: gradients with one-site and horizontal matter terms
            renv = P.compute_right_envs(top, mpo, bottom, dtype)
            left = np.ones((z, 1, 1, 1), dtype=dtype)
            row_grads = []
            for c in range(n_cols):
                g = P.compute_single_gradient(left, renv[c], top[c], bottom[c])
                row_grads.append(g)
                if one_here[c]:
                    amps = np.einsum("zpudlr,zudlr->zp", _eff(ft, config, sites, h, v, row, c), g)
                    energy = energy + _apply_ops(one_here[c], sites[:, row, c], amps, amp)
                if c < n_cols - 1 and h_here[c]:
                    amps_edge = np.einsum(
                        "zace,zaub,zedf,zpudcr,zqvwrx,zbvg,zfwi,zgxi->zpq",
                        left, top[c], bottom[c], _eff(ft, config, sites, h, v, row, c),
                        _eff(ft, config, sites, h, v, row, c + 1), top[c + 1], bottom[c + 1], renv[c + 1],
                        optimize="greedy",
                    )
                    col = sites[:, row, c] * config.phys_dim + sites[:, row, c + 1]
                    energy = energy + _apply_ops(h_here[c], col, amps_edge.reshape(z, -1), amp)
                left = P.left_env_step(left, top[c], mpo[c], bottom[c])
            grads[row] = row_grads
        else:
            grads[row] = P.compute_all_row_gradients(top, bottom, mpo)
        v_here = [terms.vertical.get((row, c), []) for c in range(n_cols)] if row < n_rows - 1 else []
        if any(v_here):
            # gi/model.py:517-536 (energy.py:47-95 over the GI effective tensors)
            mpo_next = _row_mpo(ft, s, row + 1, n_cols)
            bot_next = bottom_cache[row + 1]
            renv2 = P.compute_right_envs_2row(top, mpo, mpo_next, bot_next, dtype)
            left2 = np.ones((z, 1, 1, 1, 1), dtype=dtype)
            for c in range(n_cols):
                if v_here[c]:
                    amps_edge = np.einsum(
                        "zalmg,zaub,zpuvlr,zqvwmn,zgwf,zbrnf->zpq",
                        left2, top[c], _eff(ft, config, sites, h, v, row, c),
                        _eff(ft, config, sites, h, v, row + 1, c), bot_next[c], renv2[c], optimize="greedy",
                    )
                    col = sites[:, row, c] * config.phys_dim + sites[:, row + 1, c]
                    energy = energy + _apply_ops(v_here[c], col, amps_edge.reshape(z, -1), amp)
                left2 = _left2_step(left2, top[c], mpo[c], mpo_next[c], bot_next[c])
        if row < n_rows - 1 and any((row, c) in terms.plaquette for c in range(n_cols - 1)):
            mpo_next = _row_mpo(ft, s, row + 1, n_cols)
            bot_next = bottom_cache[row + 1]
            renv = P.compute_right_envs_2row(top, mpo, mpo_next, bot_next, dtype)
            left = np.ones((z, 1, 1, 1, 1), dtype=dtype)
            for c in range(n_cols - 1):
                coeff = terms.plaquette.get((row, c))
                if coeff is not None:
                    tot = 0.0
                    for delta in (1, -1):
                        hp, vp = _plaquette_flip(h, v, row, c, np.full(z, delta), N)
                        m00, m01, m10, m11 = _flipped_mpos(ft, config, sites, hp, vp, row, c)
                        tot = tot + _plaquette_amp(left, top, bot_next, m00, m10, m01, m11, renv[c + 1], c)
                    energy = energy + coeff * tot / amp
                left = _left2_step(left, top[c], mpo[c], mpo_next[c], bot_next[c])
        bottom = P.apply_mpo_from_below(bottom, tuple(mpo), strategy)
    o, p = P.assemble_log_derivatives(ft, config.shape, grads, s, amp, False)
    return P.Cache(bottom_envs=tuple(bottom_cache)), P.LocalEstimates(o, energy, p, amp)


def mc_sampler(tensors, samples, cache, config, strategy, terms, seed, n_steps, sweep0=0):
    """sampling.py:22-64 over the GI transition/estimate."""
    z = np.asarray(samples).shape[0]
    chains = np.arange(z, dtype=np.uint32)
    cur = np.asarray(samples)
    hist = {"samples": [], "o": [], "eloc": [], "p": [], "amp": []}
    for t in range(n_steps):
        cur, ctx = transition(tensors, cur, cache, config, strategy, seed, sweep0 + t, chains)
        cache, est = estimate(tensors, cur, ctx, config, strategy, terms)
        hist["samples"].append(cur.copy())
        hist["o"].append(est.local_log_derivatives)
        hist["eloc"].append(est.local_estimate)
        hist["p"].append(est.active_slice_indices)
        hist["amp"].append(est.amp)
    return cur, cache, {k: np.stack(vv) for k, vv in hist.items()}
