"""Pins the oracle's gauge-invariant PEPS restatement (oracle.gi).

* Gauss law holds for random configurations and after every sampler sweep
  (recipe of tests/experimental/test_gi_peps.py:15-66);
* an invalid sample has amplitude 0 (tests/experimental/test_gi_peps_experimental.py:89-109);
* the cfg index maps the feasible local link configurations of every site
  one-to-one onto [0, nc);
* the amplitude is multilinear in the selected slice and the sliced
  gradient matches central finite differences;
* E_loc of plaquette + electric terms equals <x|H|psi>/psi(x) with H built
  densely on the gauge-invariant link basis (plaquette operators as
  permutation matrices), for every basis state;
* the plaquette sampler's histogram matches |psi|^2 over that basis;
* charge masks of a non-uniform degeneracy_per_charge: slice_links inverts the
  cfg index (with matter charges and Qx != 0), per-slice masks give
  _assemble_site's per-sample amplitudes, entries a mask removes never matter
  and unmasked ones do, and the dense-H E_loc holds with masks.
"""
import itertools

import numpy as np
import pytest

from oracle import gi as G
from oracle import peps as P


def _cfg(shape=(2, 3), N=2, D=2, deg=None):
    return G.GIConfig(shape=shape, N=N, phys_dim=1, Qx=0, degeneracy_per_charge=deg or (D,) * N,
                      charge_of_site=(0,))


def _basis(config):
    """All gauge-invariant link configurations: plaquette deltas on zero links."""
    n_rows, n_cols = config.shape
    plaq = [(r, c) for r in range(n_rows - 1) for c in range(n_cols - 1)]
    out = []
    for deltas in itertools.product(range(config.N), repeat=len(plaq)):
        h = np.zeros((1, n_rows, n_cols - 1), np.int64)
        v = np.zeros((1, n_rows - 1, n_cols), np.int64)
        for (r, c), dl in zip(plaq, deltas):
            h, v = G._plaquette_flip(h, v, r, c, np.array([dl]), config.N)
        out.append(G.flatten(np.zeros((1, n_rows, n_cols), np.int64), h, v)[0])
    return np.unique(np.stack(out), axis=0)


def test_random_configurations_satisfy_gauss_law():
    for shape, N in (((3, 3), 2), ((3, 4), 3)):
        cfg = _cfg(shape, N)
        s = G.random_configurations(cfg, 64, seed=1)
        assert G.gauss_valid(cfg, s).all()
        assert len(np.unique(s, axis=0)) > 8


def test_invalid_sample_has_zero_amplitude():
    cfg = _cfg((3, 3))
    T = G.random_gi_tensors(cfg, seed=0)
    s = G.random_configurations(cfg, 4, seed=0)
    bad = s.copy()
    bad[:, 9] = 1 - bad[:, 9]  # flip one horizontal link only
    assert not G.gauss_valid(cfg, bad).any()
    assert np.all(G.gi_value(T, bad, cfg, P.NoTruncation()) == 0)
    assert np.all(G.gi_value(T, s, cfg, P.NoTruncation()) != 0)


@pytest.mark.parametrize("N", [2, 3])
def test_cfg_index_is_a_bijection_on_feasible_links(N):
    cfg = _cfg((3, 4), N)
    n_rows, n_cols = cfg.shape
    for r in range(n_rows):
        for c in range(n_cols):
            dirs = [dname for dname, ok in (("left", c > 0), ("right", c < n_cols - 1), ("up", r > 0),
                                          ("down", r < n_rows - 1)) if ok]
            seen = set()
            for vals in itertools.product(range(N), repeat=len(dirs)):
                k = dict(zip(dirs, vals))
                div = (k.get("left", 0) + k.get("down", 0) - k.get("up", 0) - k.get("right", 0)) % N
                if div != 0:  # pure gauge, Qx = 0: only divergence-free stars are feasible
                    continue
                h = np.zeros((1, n_rows, n_cols - 1), np.int64)
                v = np.zeros((1, n_rows - 1, n_cols), np.int64)
                if "left" in k:
                    h[0, r, c - 1] = k["left"]
                if "right" in k:
                    h[0, r, c] = k["right"]
                if "up" in k:
                    v[0, r - 1, c] = k["up"]
                if "down" in k:
                    v[0, r, c] = k["down"]
                seen.add(int(G.site_cfg_index(cfg, h, v, r, c)[0]))
            assert seen == set(range(G.nc_of(cfg, r, c))), (r, c)


def test_gradient_matches_finite_differences():
    cfg = _cfg((3, 3))
    T = G.random_gi_tensors(cfg, seed=3)
    s = G.random_configurations(cfg, 3, seed=3)
    amp, grad, _ = G.gi_value_and_grad(T, s, cfg, P.NoTruncation(), full_gradient=True)
    flat = np.concatenate([t.ravel() for row in T for t in row])
    rng = np.random.default_rng(0)
    for idx in rng.choice(flat.size, 12, replace=False):
        eps = 1e-6

        def val(x):
            TT, off = [], 0
            for row in T:
                rr = []
                for t in row:
                    rr.append(x[off:off + t.size].reshape(t.shape))
                    off += t.size
                TT.append(rr)
            return G.gi_value(TT, s, cfg, P.NoTruncation())

        xp, xm = flat.copy(), flat.copy()
        xp[idx] += eps
        xm[idx] -= eps
        fd = (val(xp) - val(xm)) / (2 * eps)
        assert np.allclose(grad[:, idx], fd, rtol=1e-6, atol=1e-9 * np.abs(amp).max())


def _dense_eloc(T, config, basis, coeff_p, coeff_e, strategy):
    n_rows, n_cols = config.shape
    psi = G.literal_amplitude(config, T, basis, strategy)  # _assemble_site's per-sample masks
    index = {tuple(b): i for i, b in enumerate(basis)}
    nb = len(basis)
    H = np.zeros((nb, nb))
    links_h, links_v = G.electric_terms(config.shape, coeff_e, config.N)
    _, h, v = G.unflatten(basis, config.shape)
    for r, c, diag in links_h:
        H[np.arange(nb), np.arange(nb)] += diag[h[:, r, c]]
    for r, c, diag in links_v:
        H[np.arange(nb), np.arange(nb)] += diag[v[:, r, c]]
    for r in range(n_rows - 1):
        for c in range(n_cols - 1):
            for delta in (1, -1):
                hp, vp = G._plaquette_flip(h, v, r, c, np.full(nb, delta), config.N)
                y = G.flatten(np.zeros((nb, n_rows, n_cols), np.int64), hp, vp)
                for i in range(nb):
                    H[i, index[tuple(y[i])]] += coeff_p
    return psi, (H @ psi) / psi


@pytest.mark.parametrize("shape,N,deg", [((2, 3), 2, None), ((3, 3), 2, None), ((2, 3), 3, None),
                                         ((3, 3), 2, (2, 1)), ((2, 3), 3, (1, 3, 2))])
def test_local_energy_matches_dense_hamiltonian(shape, N, deg):
    cfg = _cfg(shape, N, deg=deg)
    T = G.random_gi_tensors(cfg, seed=5)
    basis = _basis(cfg)
    strat = P.NoTruncation()
    psi, e_dense = _dense_eloc(T, cfg, basis, -1.3, 0.7, strat)
    links_h, links_v = G.electric_terms(shape, 0.7, N)
    terms = G.GITerms({(r, c): -1.3 for r in range(shape[0] - 1) for c in range(shape[1] - 1)}, links_h, links_v)
    cache = G.init_cache(T, basis, cfg, strat)
    # the estimate reads the transition's top boundaries: run it with u = 2 (never accepts)
    tops = P.forward_with_cache(G.flat_tensors(T, cfg), G.slices(cfg, basis).reshape(-1, *shape), strat)
    ctx = P.Context(amp=psi, top_envs=tops[1])
    _, est = G.estimate(T, basis, ctx, cfg, strat, terms)
    assert np.max(np.abs(est.local_estimate - e_dense)) < 1e-9 * max(1.0, np.max(np.abs(e_dense)))
    del cache


def test_plaquette_sampler_histogram_matches_psi_squared():
    cfg = _cfg((2, 3), 2)
    T = G.random_gi_tensors(cfg, seed=9)
    basis = _basis(cfg)
    strat = P.NoTruncation()
    psi = G.gi_value(T, basis, cfg, strat)
    prob = np.abs(psi) ** 2 / np.sum(np.abs(psi) ** 2)
    n_chains, n_steps = 256, 40
    s0 = G.random_configurations(cfg, n_chains, seed=2)
    cache = G.init_cache(T, s0, cfg, strat)
    links_h, links_v = G.electric_terms(cfg.shape, 1.0, 2)
    terms = G.GITerms({(0, 0): -1.0, (0, 1): -1.0}, links_h, links_v)
    _, _, hist = G.mc_sampler(T, s0, cache, cfg, strat, terms, seed=4, n_steps=n_steps)
    samples = hist["samples"][5:].reshape(-1, hist["samples"].shape[-1])
    assert G.gauss_valid(cfg, samples).all()
    index = {tuple(b): i for i, b in enumerate(basis)}
    counts = np.bincount([index[tuple(x)] for x in samples], minlength=len(basis))
    freq = counts / counts.sum()
    assert np.max(np.abs(freq - prob)) < 0.02, (freq, prob)
    # the sampler's estimate agrees with the evaluator at the sampled states
    amp = G.gi_value(T, hist["samples"][-1], cfg, strat)
    assert np.allclose(hist["amp"][-1], amp, rtol=1e-10)


def _matter_basis(config):
    """All Gauss-valid (matter, link) configurations, by enumeration."""
    n_rows, n_cols = config.shape
    ns, nh, nv = n_rows * n_cols, n_rows * (n_cols - 1), (n_rows - 1) * n_cols
    sites = np.array(list(itertools.product(range(config.phys_dim), repeat=ns)), np.int64)
    links = np.array(list(itertools.product(range(config.N), repeat=nh + nv)), np.int64)
    allc = np.concatenate([np.repeat(sites, len(links), axis=0), np.tile(links, (len(sites), 1))], axis=1)
    return allc[G.gauss_valid(config, allc)].astype(np.int32)


@pytest.mark.parametrize("shape,charges,phys", [((2, 2), (0, 1), 2), ((2, 3), (0, 1), 2), ((2, 2), (0, 1, 1), 3)])
def test_matter_sampler_histogram_matches_psi_squared(shape, charges, phys):
    """Plaquette + horizontal + vertical link sweeps with matter moves
    (gi/model.py:703-1043) sample |psi|^2 over the Gauss sector, including a
    charge with two physical states (proposal ratio deg'/deg)."""
    cfg = G.GIConfig(shape=shape, N=2, phys_dim=phys, Qx=0, degeneracy_per_charge=(2, 2), charge_of_site=charges)
    T = G.random_gi_tensors(cfg, seed=11)
    basis = _matter_basis(cfg)
    strat = P.NoTruncation()
    psi = G.gi_value(T, basis, cfg, strat)
    prob = np.abs(psi) ** 2 / np.sum(np.abs(psi) ** 2)
    n_chains, n_steps = 256, 40
    s0 = G.random_configurations(cfg, n_chains, seed=3)
    cache = G.init_cache(T, s0, cfg, strat)
    links_h, links_v = G.electric_terms(cfg.shape, 1.0, 2)
    terms = G.GITerms({}, links_h, links_v)
    _, _, hist = G.mc_sampler(T, s0, cache, cfg, strat, terms, seed=6, n_steps=n_steps)
    samples = hist["samples"][5:].reshape(-1, hist["samples"].shape[-1])
    assert G.gauss_valid(cfg, samples).all()
    index = {tuple(b): i for i, b in enumerate(basis)}
    counts = np.bincount([index[tuple(x)] for x in samples], minlength=len(basis))
    freq = counts / counts.sum()
    assert np.max(np.abs(freq - prob)) < 0.02, np.max(np.abs(freq - prob))
    # matter indices actually move, and the sampler's amplitude is the evaluator's
    assert len(np.unique(samples[:, : shape[0] * shape[1]], axis=0)) > 1
    amp = G.gi_value(T, hist["samples"][-1], cfg, strat)
    assert np.allclose(hist["amp"][-1], amp, rtol=1e-10)


def test_charge_index_map():
    """gi/model.py:190-206: ascending indices per charge, -1 padded, counts."""
    cfg = G.GIConfig(shape=(2, 2), N=3, phys_dim=5, Qx=0, degeneracy_per_charge=(1, 1, 1),
                     charge_of_site=(2, 0, 2, 1, 2))
    table, deg = G.charge_index_map(cfg)
    assert deg.tolist() == [1, 1, 3]
    assert table.tolist() == [[1, -1, -1], [3, -1, -1], [0, 2, 4]]


def test_matter_local_energy_matches_dense_hamiltonian():
    """Plaquette + electric + matter-diagonal (mass, two-site) E_loc equals
    <x|H|psi>/psi(x) over the Gauss sector with matter, H built densely."""
    cfg = G.GIConfig(shape=(2, 3), N=2, phys_dim=2, Qx=0, degeneracy_per_charge=(2, 2), charge_of_site=(0, 1))
    T = G.random_gi_tensors(cfg, seed=12)
    basis = _matter_basis(cfg)
    strat = P.NoTruncation()
    psi = G.gi_value(T, basis, cfg, strat)
    nb = len(basis)
    index = {tuple(b): i for i, b in enumerate(basis)}
    sites, h, v = G.unflatten(basis, cfg.shape)
    links_h, links_v = G.electric_terms(cfg.shape, 0.4, 2)
    mass = [(((r, c),), np.asarray(cfg.charge_of_site) * (0.3 + 0.1 * (r + c))) for r in range(2) for c in range(3)]
    pair = [(((0, 1), (1, 1)), np.array([0.2, -0.5, 0.7, 1.1]))]
    H = np.zeros((nb, nb))
    for r, c, dg in links_h:
        H[np.arange(nb), np.arange(nb)] += dg[h[:, r, c]]
    for r, c, dg in links_v:
        H[np.arange(nb), np.arange(nb)] += dg[v[:, r, c]]
    for tsites, dg in mass + pair:
        idx = np.zeros(nb, np.int64)
        for r, c in tsites:
            idx = idx * 2 + sites[:, r, c]
        H[np.arange(nb), np.arange(nb)] += dg[idx]
    for c in range(2):
        for delta in (1, -1):  # gi/model.py:560-606 adds both flips (equal for Z2)
            hp, vp = G._plaquette_flip(h, v, 0, c, np.full(nb, delta), 2)
            y = G.flatten(sites, hp, vp)
            for i in range(nb):
                H[i, index[tuple(y[i])]] += -0.9
    e_dense = (H @ psi) / psi
    terms = G.GITerms({(0, 0): -0.9, (0, 1): -0.9}, links_h, links_v, tuple(mass + pair))
    tops = P.forward_with_cache(G.flat_tensors(T, cfg), G.slices(cfg, basis).reshape(-1, 2, 3), strat)
    _, est = G.estimate(T, basis, P.Context(amp=psi, top_envs=tops[1]), cfg, strat, terms)
    assert np.max(np.abs(est.local_estimate - e_dense)) < 1e-9 * max(1.0, np.max(np.abs(e_dense)))


def test_matter_operators_match_dense_hamiltonian():
    """One-site, horizontal and vertical operators on matter states (links held
    fixed, gi/model.py:474-536) plus a plaquette: E_loc equals <x|H|psi>/psi(x)
    with H built densely over the Gauss sector. phys_dim 3 with charges (0, 1, 1):
    the operators preserve every site's charge (block diagonal in the charge,
    mixing the two charge-1 states), so H maps the sector to itself; H applies
    op[:, s] like the reference (energy.py:88, 179, 200)."""
    cfg = G.GIConfig(shape=(2, 3), N=2, phys_dim=3, Qx=0, degeneracy_per_charge=(1, 1), charge_of_site=(0, 1, 1))
    T = G.random_gi_tensors(cfg, seed=13)
    basis = _matter_basis(cfg)
    strat = P.NoTruncation()
    psi = G.gi_value(T, basis, cfg, strat)
    nb = len(basis)
    index = {tuple(b): i for i, b in enumerate(basis)}
    sites, h, v = G.unflatten(basis, cfg.shape)
    rng = np.random.default_rng(3)

    def one_op():
        a = np.zeros((3, 3))
        a[0, 0] = rng.normal()
        a[1:, 1:] = rng.normal(size=(2, 2))  # mixes the two charge-1 states, non-symmetric
        return a

    two_op = lambda: np.kron(one_op(), one_op()) + 0.3 * np.kron(one_op(), one_op())
    one = {(0, 1): [one_op()], (1, 2): [one_op(), one_op()]}
    hor = {(0, 0): [two_op()], (1, 1): [two_op()]}
    ver = {(0, 2): [two_op()]}
    H = np.zeros((nb, nb))

    def add(tsites, op):
        d = 3
        sidx = np.zeros(nb, np.int64)
        for r, c in tsites:
            sidx = sidx * d + sites[:, r, c]
        for pidx in range(op.shape[0]):
            coeff = op[pidx, sidx]
            if not np.any(coeff):
                continue
            sp = sites.copy()
            rem = pidx
            for r, c in reversed(tsites):
                sp[:, r, c] = rem % d
                rem //= d
            y = G.flatten(sp, h, v)
            for i in np.nonzero(coeff)[0]:
                H[i, index[tuple(y[i])]] += coeff[i]

    for (r, c), ops in one.items():
        for op in ops:
            add([(r, c)], op)
    for (r, c), ops in hor.items():
        for op in ops:
            add([(r, c), (r, c + 1)], op)
    for (r, c), ops in ver.items():
        for op in ops:
            add([(r, c), (r + 1, c)], op)
    for delta in (1, -1):
        hp, vp = G._plaquette_flip(h, v, 0, 1, np.full(nb, delta), 2)
        y = G.flatten(sites, hp, vp)
        for i in range(nb):
            H[i, index[tuple(y[i])]] += -0.7
    e_dense = (H @ psi) / psi
    terms = G.GITerms({(0, 1): -0.7}, [], [], (), one, hor, ver)
    tops = P.forward_with_cache(G.flat_tensors(T, cfg), G.slices(cfg, basis).reshape(-1, 2, 3), strat)
    _, est = G.estimate(T, basis, P.Context(amp=psi, top_envs=tops[1]), cfg, strat, terms)
    assert np.max(np.abs(est.local_estimate - e_dense)) < 1e-9 * max(1.0, np.max(np.abs(e_dense)))


def test_slice_links_invert_the_cfg_index():
    cfg = G.GIConfig(shape=(3, 4), N=3, phys_dim=2, Qx=1, degeneracy_per_charge=(2, 1, 2), charge_of_site=(0, 2))
    n_rows, n_cols = cfg.shape
    for r in range(n_rows):
        for c in range(n_cols):
            dirs = [dname for dname, ok in (("left", c > 0), ("right", c < n_cols - 1), ("up", r > 0),
                                          ("down", r < n_rows - 1)) if ok]
            hits = 0
            for sigma in range(2):
                q = cfg.charge_of_site[sigma]
                for vals in itertools.product(range(3), repeat=len(dirs)):
                    k = dict(zip(dirs, vals))
                    div = k.get("left", 0) + k.get("down", 0) - k.get("up", 0) - k.get("right", 0)
                    if (div + q) % 3 != cfg.Qx:
                        continue
                    h = np.zeros((1, n_rows, n_cols - 1), np.int64)
                    v = np.zeros((1, n_rows - 1, n_cols), np.int64)
                    if "left" in k:
                        h[0, r, c - 1] = k["left"]
                    if "right" in k:
                        h[0, r, c] = k["right"]
                    if "up" in k:
                        v[0, r - 1, c] = k["up"]
                    if "down" in k:
                        v[0, r, c] = k["down"]
                    idx = int(G.site_cfg_index(cfg, h, v, r, c)[0])
                    got = G.slice_links(cfg, r, c, sigma, idx)
                    assert got == (k.get("left", 0), k.get("up", 0), k.get("down", 0), k.get("right", 0))
                    hits += 1
            assert hits == 2 * G.nc_of(cfg, r, c)


@pytest.mark.parametrize("N,deg,phys,charges,Qx,dtype", [
    (2, (2, 1), 1, (0,), 0, np.float64),
    (3, (1, 3, 2), 1, (0,), 0, np.complex128),
    (2, (1, 2), 2, (0, 1), 1, np.float64),
])
def test_slice_masks_equal_per_sample_masks(N, deg, phys, charges, Qx, dtype):
    cfg = G.GIConfig(shape=(3, 3), N=N, phys_dim=phys, Qx=Qx, degeneracy_per_charge=deg, charge_of_site=charges)
    T = G.random_gi_tensors(cfg, seed=4, dtype=dtype)
    s = G.random_configurations(cfg, 32, seed=5)
    strat = P.ZipUp(81)
    lit = G.literal_amplitude(cfg, T, s, strat)
    flat = G.gi_value(T, s, cfg, strat)
    scale = np.max(np.abs(lit))
    assert np.max(np.abs(flat - lit)) <= 1e-13 * scale
    # entries the per-slice masks remove never matter
    m = cfg.mask()
    T2 = [[t.copy() for t in row] for row in T]
    for r in range(3):
        for c in range(3):
            t = T2[r][c]
            for sigma in range(t.shape[0]):
                for k in range(t.shape[1]):
                    kl, ku, kd, kr = G.slice_links(cfg, r, c, sigma, k)
                    keep = (m[ku][:t.shape[2]][:, None, None, None] & m[kd][:t.shape[3]][None, :, None, None]
                            & m[kl][:t.shape[4]][None, None, :, None] & m[kr][:t.shape[5]][None, None, None, :])
                    t[sigma, k][~keep] += 7.0
    assert np.max(np.abs(G.literal_amplitude(cfg, T2, s, strat) - lit)) <= 1e-13 * scale
    assert np.max(np.abs(G.gi_value(T2, s, cfg, strat) - lit)) <= 1e-13 * scale
    # an entry inside every sector is kept
    T3 = [[t.copy() for t in row] for row in T]
    T3[1][1][:, :, 0, 0, 0, 0] += 1.0
    assert np.max(np.abs(G.gi_value(T3, s, cfg, strat) - lit)) > 1e-6 * scale
