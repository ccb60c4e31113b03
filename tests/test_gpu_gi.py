"""GPU parity of the gauge-invariant PEPS path (vmc_gi_* C ABI) against oracle.gi.

oracle.gi restates src/vmc/peps/gi/* and is pinned on the CPU
(tests/test_oracle_gi.py: Gauss law, dense-Hamiltonian E_loc, histogram). The
device sampler draws the oracle's Philox stream, so in FP64 the sampled link
configurations must be identical and amp, o and E_loc agree to rounding.
"""
import numpy as np
import pytest
import torch

from oracle import gi as G
from oracle import peps as P

pytestmark = pytest.mark.gpu


def _cfg(shape, N=2, dmax=2, d=1, deg=None):
    return G.GIConfig(shape=shape, N=N, phys_dim=d, Qx=0, degeneracy_per_charge=deg or (dmax,) * N,
                      charge_of_site=tuple(range(d)) if d > 1 else (0,))


def _flat(T):
    return np.concatenate([t.ravel() for row in T for t in row])


def _strategy(name, chi):
    from vmc_amd import _lib

    return {"zipup": (_lib.VMC_ZIPUP, P.ZipUp(chi)), "exact": (_lib.VMC_NO_TRUNCATION, P.NoTruncation()),
            "variational": (_lib.VMC_VARIATIONAL, P.Variational(chi, n_sweeps=2))}[name]


@pytest.mark.parametrize("shape,N,dmax,d,chi,strat,deg", [
    ((3, 3), 2, 2, 1, 0, "exact", None),
    ((4, 4), 2, 2, 1, 4, "zipup", None),
    ((6, 6), 2, 2, 1, 8, "zipup", None),
    ((4, 4), 3, 1, 1, 3, "zipup", None),
    ((3, 4), 2, 1, 2, 4, "zipup", None),
    ((6, 6), 2, 2, 1, 8, "variational", None),
    # charge masks (degeneracy_per_charge below dmax); untruncated, since the
    # masks make theta exactly rank deficient and a truncating cut then
    # depends on the factorisation's null vectors
    ((4, 4), 2, 2, 1, 16, "zipup", (2, 1)),
    ((3, 4), 3, 2, 1, 0, "exact", (1, 2, 2)),
    ((3, 3), 2, 2, 2, 0, "exact", (1, 2)),
])
def test_gi_amp_grad_matches_oracle(shape, N, dmax, d, chi, strat, deg):
    from vmc_amd.engine import GIPlan

    code, ostrat = _strategy(strat, chi)
    print(f"[gi] amp_grad shape={shape} N={N} dmax={dmax} d={d} chi={chi} {strat} deg={deg}", flush=True)
    cfg = _cfg(shape, N, dmax, d, deg)
    T = G.random_gi_tensors(cfg, seed=1)
    s = G.random_configurations(cfg, 16, seed=2)
    bad = s[:2].copy()
    bad[:, shape[0] * shape[1]] = (bad[:, shape[0] * shape[1]] + 1) % N  # break Gauss law on one link
    samples = np.concatenate([s, bad])
    plan = GIPlan(shape, N, d, dmax, max(chi, 1), cfg.charge_of_site, strategy=code, dtype=torch.float64,
                  degeneracy_per_charge=deg)
    params = torch.tensor(_flat(T), dtype=torch.float64, device="cuda")
    la, o, p = plan.amp_grad(params, torch.tensor(samples), want_p=True)
    sl, valid = plan.slices(torch.tensor(samples))
    torch.cuda.synchronize()
    la, o, p = la.cpu().numpy(), o.cpu().numpy(), p.cpu().numpy()
    assert np.array_equal(sl.cpu().numpy(), G.slices(cfg, samples))
    assert np.array_equal(valid.cpu().numpy(), G.gauss_valid(cfg, samples))
    amp, grad, p_ref = G.gi_value_and_grad(T, s, cfg, ostrat)
    n = len(s)
    rel_log = np.max(np.abs(la[:n, 0] - np.log(np.abs(amp))) / np.maximum(1.0, np.abs(np.log(np.abs(amp)))))
    sign_ok = np.array_equal(np.where(la[:n, 1] != 0, -1.0, 1.0), np.sign(amp))
    o_ref = grad / amp[:, None]
    err_o = np.max(np.abs(o[:n] - o_ref)) / np.max(np.abs(o_ref))
    print(f"  rel err log|psi| {rel_log:.2e}  o {err_o:.2e}", flush=True)
    assert sign_ok and rel_log < 1e-11 and err_o < 1e-9
    assert np.array_equal(p[:n], p_ref)
    # Gauss-law violations: amplitude 0 (gi/compat.py:90-99), zero log-derivatives
    assert np.all(np.isneginf(la[n:, 0])) and np.all(o[n:] == 0)


@pytest.mark.parametrize("shape,N,dmax,chi,strat,deg", [
    ((4, 4), 2, 2, 8, "zipup", None),
    ((3, 4), 3, 1, 9, "zipup", None),
    ((4, 4), 2, 2, 8, "variational", None),
    ((4, 4), 2, 2, 16, "zipup", (1, 2)),
])
def test_gi_plaquette_sampler_matches_oracle(shape, N, dmax, chi, strat, deg):
    from vmc_amd.engine import GIPlan

    code, ostrat = _strategy(strat, chi)
    print(f"[gi] sampler shape={shape} N={N} dmax={dmax} chi={chi} {strat} deg={deg}", flush=True)
    cfg = _cfg(shape, N, dmax, deg=deg)
    R, C = shape
    T = G.random_gi_tensors(cfg, seed=3)
    n_chains, n_steps, seed = 16, 3, 11
    s0 = G.random_configurations(cfg, n_chains, seed=4)
    coeff = np.full((R - 1) * (C - 1), -1.0)
    coeff[1] = np.nan  # one plaquette without a term
    lh, lv = G.electric_terms(shape, 0.6, N)
    link = np.concatenate([np.stack([dg for _, _, dg in lh]), np.stack([dg for _, _, dg in lv])])
    terms = G.GITerms({(i // (C - 1), i % (C - 1)): coeff[i] for i in range(len(coeff)) if not np.isnan(coeff[i])},
                      lh, lv)
    plan = GIPlan(shape, N, 1, dmax, chi, (0,), strategy=code, dtype=torch.float64, degeneracy_per_charge=deg)
    plan.set_gi_terms(coeff, link)
    params = torch.tensor(_flat(T), dtype=torch.float64, device="cuda")
    samples = torch.tensor(s0, dtype=torch.int32, device="cuda")
    cache = plan.new_cache(n_chains)
    plan.init_cache(params, samples, cache)
    outs = []
    for t in range(n_steps):
        la, o, p, e = plan.mc_sweep(params, samples, cache, seed, t)
        outs.append((samples.cpu().numpy().copy(), la.cpu().numpy(), o.cpu().numpy(), p.cpu().numpy(),
                     e.cpu().numpy()))
    cref = G.init_cache(T, s0, cfg, ostrat)
    _, _, hist = G.mc_sampler(T, s0, cref, cfg, ostrat, terms, seed, n_steps)
    moved = 0
    for t, (sm, la, o, p, e) in enumerate(outs):
        assert np.array_equal(sm, hist["samples"][t]), f"sweep {t}: samples differ"
        assert G.gauss_valid(cfg, sm).all()
        moved += int(np.sum(np.any(sm != (outs[t - 1][0] if t else s0), axis=1)))
        dl = np.max(np.abs(la[:, 0] - np.log(np.abs(hist["amp"][t]))))
        err_o = np.max(np.abs(o - hist["o"][t])) / np.max(np.abs(hist["o"][t]))
        err_e = np.max(np.abs(e - hist["eloc"][t])) / max(1.0, np.max(np.abs(hist["eloc"][t])))
        print(f"  sweep {t}: dlog {dl:.1e} o {err_o:.1e} eloc {err_e:.1e}", flush=True)
        assert dl < 1e-10 and err_o < 1e-9 and err_e < 1e-9
        assert np.array_equal(p, hist["p"][t])
    assert moved > 0  # the sweep accepted plaquette moves


@pytest.mark.parametrize("shape,N,dmax,d,charges,chi,strat,deg", [
    ((3, 3), 2, 2, 2, (0, 1), 0, "exact", None),
    ((4, 4), 2, 2, 2, (0, 1), 8, "zipup", None),
    ((3, 4), 2, 1, 3, (0, 1, 1), 4, "zipup", None),
    ((4, 4), 2, 2, 2, (0, 1), 8, "variational", None),
    ((3, 4), 2, 2, 2, (0, 1), 0, "exact", (2, 1)),
])
def test_gi_matter_sampler_matches_oracle(shape, N, dmax, d, charges, chi, strat, deg):
    """Plaquette, horizontal-link and vertical-link sweeps with matter moves
    (gi/model.py:703-1043): same Philox stream as the oracle, so in FP64 the
    sampled (matter, link) configurations are identical; E_loc includes
    plaquette, electric, matter-diagonal and one-site / two-site matter terms."""
    from vmc_amd.engine import GIPlan

    code, ostrat = _strategy(strat, chi)
    print(f"[gi] matter sampler shape={shape} N={N} dmax={dmax} d={d} charges={charges} {strat} deg={deg}",
          flush=True)
    cfg = G.GIConfig(shape=shape, N=N, phys_dim=d, Qx=0, degeneracy_per_charge=deg or (dmax,) * N,
                     charge_of_site=charges)
    R, C = shape
    T = G.random_gi_tensors(cfg, seed=7)
    n_chains, n_steps, seed = 16, 3, 13
    s0 = G.random_configurations(cfg, n_chains, seed=8)
    coeff = np.full((R - 1) * (C - 1), -0.8)
    lh, lv = G.electric_terms(shape, 0.5, N)
    link = np.concatenate([np.stack([dg for _, _, dg in lh]), np.stack([dg for _, _, dg in lv])])
    # matter-diagonal terms: a mass on every site (MatterMassTerm) and one two-site diagonal term
    mass = [(((r, c),), np.asarray(charges, np.float64) * (0.3 + 0.05 * (r * C + c))) for r in range(R) for c in range(C)]
    pair = [(((0, 0), (1, 0)), np.arange(d * d, dtype=np.float64) * 0.25 - 0.4)]
    # operators on matter states with the links held fixed (gi/model.py:474-536); off the Gauss
    # sector both sides contract the raw network, so any operator is a parity test
    rng = np.random.default_rng(17)
    one = {(0, 1): [rng.normal(size=(d, d))], (R - 1, C - 1): [rng.normal(size=(d, d))]}
    hor = {(1, 0): [rng.normal(size=(d * d, d * d))]}
    ver = {(0, C - 1): [rng.normal(size=(d * d, d * d))]}
    terms = G.GITerms({(i // (C - 1), i % (C - 1)): coeff[i] for i in range(len(coeff))}, lh, lv,
                      tuple(mass + pair), one, hor, ver)
    plan = GIPlan(shape, N, d, dmax, max(chi, 1), charges, strategy=code, dtype=torch.float64,
                  degeneracy_per_charge=deg)
    plan.set_gi_terms(coeff, link)
    dsites, dvals = [], []
    for tsites, dg in mass + pair:
        fl = [r * C + c for r, c in tsites]
        dsites += [fl[0], fl[1] if len(fl) == 2 else -1]
        vals = np.zeros(d * d)
        vals[: dg.size] = dg
        dvals.append(vals)
    ei, ed = np.zeros(0, np.int32), np.zeros(0)
    def flat_ops(dct):
        sites_ = [r * C + c for (r, c), ops in dct.items() for _ in ops]
        ops_ = [op.ravel() for ops in dct.values() for op in ops]
        return np.asarray(sites_, np.int32), np.concatenate(ops_)

    os_, oo = flat_ops(one)
    hs, ho = flat_ops(hor)
    vs, vo = flat_ops(ver)
    plan.set_terms({"diag_sites": np.asarray(dsites, np.int32), "diag_vals": np.concatenate(dvals),
                    "one_sites": os_, "one_ops": oo, "h_sites": hs, "h_ops": ho, "v_sites": vs, "v_ops": vo})
    params = torch.tensor(_flat(T), dtype=torch.float64, device="cuda")
    samples = torch.tensor(s0, dtype=torch.int32, device="cuda")
    cache = plan.new_cache(n_chains)
    plan.init_cache(params, samples, cache)
    outs = []
    for t in range(n_steps):
        la, o, p, e = plan.mc_sweep(params, samples, cache, seed, t)
        outs.append((samples.cpu().numpy().copy(), la.cpu().numpy(), o.cpu().numpy(), p.cpu().numpy(),
                     e.cpu().numpy()))
    cref = G.init_cache(T, s0, cfg, ostrat)
    _, _, hist = G.mc_sampler(T, s0, cref, cfg, ostrat, terms, seed, n_steps)
    ns = R * C
    matter_moved = 0
    for t, (sm, la, o, p, e) in enumerate(outs):
        assert np.array_equal(sm, hist["samples"][t]), f"sweep {t}: samples differ"
        assert G.gauss_valid(cfg, sm).all()
        prev = outs[t - 1][0] if t else s0
        matter_moved += int(np.sum(sm[:, :ns] != prev[:, :ns]))
        dl = np.max(np.abs(la[:, 0] - np.log(np.abs(hist["amp"][t]))))
        err_o = np.max(np.abs(o - hist["o"][t])) / np.max(np.abs(hist["o"][t]))
        err_e = np.max(np.abs(e - hist["eloc"][t])) / max(1.0, np.max(np.abs(hist["eloc"][t])))
        print(f"  sweep {t}: dlog {dl:.1e} o {err_o:.1e} eloc {err_e:.1e}", flush=True)
        assert dl < 1e-10 and err_o < 1e-9 and err_e < 1e-9
        assert np.array_equal(p, hist["p"][t])
    assert matter_moved > 0  # link moves carried matter with them


def test_gi_matter_sampler_requires_every_charge():
    """gi/model.py:92-93: with matter, every charge needs a physical state."""
    from vmc_amd.engine import GIPlan

    cfg = G.GIConfig(shape=(2, 2), N=3, phys_dim=2, Qx=0, degeneracy_per_charge=(1, 1, 1), charge_of_site=(0, 1))
    T = G.random_gi_tensors(cfg, seed=1)
    plan = GIPlan((2, 2), 3, 2, 1, 0, (0, 1), strategy=_strategy("exact", 0)[0], dtype=torch.float64)
    params = torch.tensor(_flat(T), dtype=torch.float64, device="cuda")
    samples = torch.zeros((2, G.n_sample((2, 2))), dtype=torch.int32, device="cuda")
    cache = plan.new_cache(2)
    with pytest.raises(RuntimeError, match="all charges"):
        plan.mc_sweep(params, samples, cache, 1, 0)


def _model(cfg, T, strat):
    from vmc_amd.peps import ZipUp
    from vmc_amd.peps.gi import GIPEPS, GIPEPSConfig

    mc = GIPEPSConfig(shape=cfg.shape, N=cfg.N, phys_dim=cfg.phys_dim, Qx=cfg.Qx,
                      degeneracy_per_charge=cfg.degeneracy_per_charge, charge_of_site=cfg.charge_of_site,
                      dtype=torch.float64)
    return GIPEPS(config=mc, contraction_strategy=strat or ZipUp(8), tensors=T)


def test_gi_apply_vjp_matches_oracle_full_gradient():
    from vmc_amd.peps.gi import gi_apply

    print("[gi] gi_apply vjp 3x4 Z2", flush=True)
    cfg = _cfg((3, 4), 2, 2)
    T = G.random_gi_tensors(cfg, seed=6)
    model = _model(cfg, T, None)
    s = G.random_configurations(cfg, 8, seed=6)
    flat = model.flat_params().clone().requires_grad_(True)
    amp = gi_apply(model, flat, torch.tensor(s))
    g = torch.arange(1, 9, dtype=torch.float64, device="cuda")
    (amp * g).sum().backward()
    ramp, rgrad, _ = G.gi_value_and_grad(T, s, cfg, P.ZipUp(8), full_gradient=True)
    assert np.allclose(amp.detach().cpu().numpy(), ramp, rtol=1e-11)
    ref = (rgrad * np.arange(1, 9)[:, None]).sum(0)
    assert np.allclose(flat.grad.cpu().numpy(), ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_gi_build_mc_kernels_and_make_mc_sampler():
    from vmc_amd.core import make_mc_sampler
    from vmc_amd.operators import PlaquetteTerm
    from vmc_amd.peps import ChainKeys, build_mc_kernels
    from vmc_amd.peps.gi import GILocalHamiltonian, build_electric_terms

    print("[gi] build_mc_kernels + make_mc_sampler 4x4 Z2", flush=True)
    shape = (4, 4)
    cfg = _cfg(shape, 2, 2)
    T = G.random_gi_tensors(cfg, seed=7)
    model = _model(cfg, T, None)
    terms = tuple(PlaquetteTerm(row=r, col=c, coeff=-1.0) for r in range(3) for c in range(3))
    op = GILocalHamiltonian(shape=shape, terms=terms + build_electric_terms(shape, 0.4, 2))
    init_cache, transition, estimate = build_mc_kernels(model, op)
    s0 = G.random_configurations(cfg, 8, seed=8)
    cache = init_cache(None, torch.tensor(s0))
    sampler = make_mc_sampler(transition, estimate)
    (_, _, _), (hist, est) = sampler(None, torch.tensor(s0), ChainKeys(5, 0), cache, n_steps=2)
    lh, lv = G.electric_terms(shape, 0.4, 2)
    gterms = G.GITerms({(r, c): -1.0 for r in range(3) for c in range(3)}, lh, lv)
    _, _, ref = G.mc_sampler(T, s0, G.init_cache(T, s0, cfg, P.ZipUp(8)), cfg, P.ZipUp(8), gterms, 5, 2)
    assert np.array_equal(hist.cpu().numpy(), ref["samples"])
    assert np.allclose(est.local_estimate.cpu().numpy(), ref["eloc"], rtol=1e-9, atol=1e-9)
    assert np.allclose(est.local_log_derivatives.cpu().numpy(), ref["o"], rtol=1e-9,
                       atol=1e-10 * np.abs(ref["o"]).max())


@pytest.mark.parametrize("space", ["parameter", "sample"])
def test_gi_sr_with_per_site_slices(space):
    """SR over the GI sliced Jacobian: sliced_dims differ per site (d * nc)."""
    from oracle import sr as S
    from vmc_amd.peps.gi import gi_value_and_grad
    from vmc_amd.preconditioners import SRPreconditioner
    from vmc_amd.qgt import ParameterSpace, SampleSpace

    print(f"[gi] SR per-site slices {space}", flush=True)
    cfg = _cfg((3, 3), 2, 1)
    T = G.random_gi_tensors(cfg, seed=9)
    model = _model(cfg, T, None)
    s = G.random_configurations(cfg, 48, seed=9)
    amp, grad, p = gi_value_and_grad(model, torch.tensor(s))
    o = grad / amp[:, None]
    e = torch.tensor(np.random.default_rng(1).standard_normal(48), device="cuda")
    sd = model.sliced_dims()
    assert len(set(sd)) > 1
    pre = SRPreconditioner(ParameterSpace() if space == "parameter" else SampleSpace(), diag_shift=1e-3)
    x, _ = pre.apply(model, None, None, o, p, e, grad_factor=-1.0)
    ref, _ = S.sr_apply(o.cpu().numpy(), p.cpu().numpy(), e.cpu().numpy(), model.params_per_site(), sd,
                        space=space, diag_shift=1e-3, grad_factor=-1.0)
    err = np.max(np.abs(x.cpu().numpy() - ref)) / np.max(np.abs(ref))
    print(f"  rel err {err:.2e}", flush=True)
    assert x.numel() == model.flat_params().numel() and err < 1e-8


def test_gi_cfg5_8x8_dmax4_chi32_variational():
    """The batched mixed-size form's middle member: Z2 pure-gauge GI-PEPS, 8x8,
    bond dmax = 4, chi = 2 dmax^2 = 32, Variational(n_sweeps=2), FP64 against the
    oracle; FP32 contractions against the FP64 path on the same FP32-cast values
    (median within 1e-5; the tail follows the conditioning of the truncation)."""
    from vmc_amd.engine import GIPlan

    shape, N, dmax, chi = (8, 8), 2, 4, 32
    code, ostrat = _strategy("variational", chi)
    print(f"[gi] cfg5 {shape} dmax={dmax} chi={chi} variational", flush=True)
    cfg = _cfg(shape, N, dmax, 1)
    T = G.random_gi_tensors(cfg, seed=61)
    s = G.random_configurations(cfg, 3, seed=62)
    plan = GIPlan(shape, N, 1, dmax, chi, cfg.charge_of_site, strategy=code, dtype=torch.float64)
    la, o, p = plan.amp_grad(torch.tensor(_flat(T), dtype=torch.float64, device="cuda"), torch.tensor(s),
                             want_p=True)
    la, o = la.cpu().numpy(), o.cpu().numpy()
    amp, grad, p_ref = G.gi_value_and_grad(T, s, cfg, ostrat)
    rel_log = np.max(np.abs(la[:, 0] - np.log(np.abs(amp))) / np.maximum(1.0, np.abs(np.log(np.abs(amp)))))
    o_ref = grad / amp[:, None]
    err_o = np.max(np.abs(o - o_ref)) / np.max(np.abs(o_ref))
    print(f"  FP64 rel err log|psi| {rel_log:.2e}  o {err_o:.2e}", flush=True)
    assert rel_log < 1e-10 and err_o < 1e-7
    assert np.array_equal(p.cpu().numpy(), p_ref)
    # FP32 contractions vs FP64 on identical values, 64 samples
    cast = _flat(T).astype(np.float32).astype(np.float64)
    s64 = torch.tensor(G.random_configurations(cfg, 64, seed=63))
    out = {}
    for dt in (torch.float32, torch.float64):
        pl = GIPlan(shape, N, 1, dmax, chi, cfg.charge_of_site, strategy=code, dtype=dt)
        out[dt] = pl.amp_grad(torch.tensor(cast, dtype=dt, device="cuda"), s64, grad=False)[0].cpu().numpy()
    e32 = np.abs(out[torch.float32][:, 0] - out[torch.float64][:, 0]) / np.maximum(1.0, np.abs(out[torch.float64][:, 0]))
    print(f"  FP32 vs FP64 rel err log|psi| median {np.median(e32):.2e} max {e32.max():.2e}", flush=True)
    assert np.median(e32) < 1e-5


def test_gi_cfg5_10x10_dmax6_chi72_variational():
    """The batched mixed-size form's largest member: Z2 pure-gauge GI-PEPS 10x10,
    bond dmax = 6, chi = 2 dmax^2 = 72, Variational(n_sweeps=2): QR panels up to
    432 x 72 in global scratch. FP64 against the oracle."""
    from vmc_amd.engine import GIPlan

    shape, N, dmax, chi = (10, 10), 2, 6, 72
    code, ostrat = _strategy("variational", chi)
    print(f"[gi] cfg5 {shape} dmax={dmax} chi={chi} variational", flush=True)
    cfg = _cfg(shape, N, dmax, 1)
    T = G.random_gi_tensors(cfg, seed=71)
    s = G.random_configurations(cfg, 1, seed=72)
    plan = GIPlan(shape, N, 1, dmax, chi, cfg.charge_of_site, strategy=code, dtype=torch.float64)
    la, o, p = plan.amp_grad(torch.tensor(_flat(T), dtype=torch.float64, device="cuda"), torch.tensor(s),
                             want_p=True)
    la, o = la.cpu().numpy(), o.cpu().numpy()
    amp, grad, p_ref = G.gi_value_and_grad(T, s, cfg, ostrat)
    rel_log = np.max(np.abs(la[:, 0] - np.log(np.abs(amp))) / np.maximum(1.0, np.abs(np.log(np.abs(amp)))))
    o_ref = grad / amp[:, None]
    err_o = np.max(np.abs(o - o_ref)) / np.max(np.abs(o_ref))
    print(f"  FP64 rel err log|psi| {rel_log:.2e}  o {err_o:.2e}", flush=True)
    assert rel_log < 1e-10 and err_o < 1e-7
    assert np.array_equal(p.cpu().numpy(), p_ref)


def test_gi_degeneracy_is_validated():
    from vmc_amd.engine import GIPlan

    plan = GIPlan((3, 3), 3, 1, 2, 4, (0,), dtype=torch.float64)
    with pytest.raises(RuntimeError, match=r"\[1, dmax\]"):
        plan.set_degeneracy((2, 0, 2))
    with pytest.raises(RuntimeError, match="must equal the plan's dmax"):
        plan.set_degeneracy((1, 1, 1))
    plan.set_degeneracy((2, 1, 2))
