"""The reference-named Python surface on the device engine, against the oracle.

_value/_value_and_grad (compat.py:67-133), local_estimate (compat.py:136-194),
the custom VJP of peps_apply (compat.py:31-64) and build_mc_kernels +
make_mc_sampler (kernels.py:91-247, sampling.py:22-82).
"""
import numpy as np
import pytest
import torch

from oracle import hamiltonians as HH
from oracle import peps as P

pytestmark = pytest.mark.gpu


def _model(shape, D, chi, seed, dtype=torch.float64):
    from vmc_amd.peps import PEPS, ZipUp

    T = P.random_peps(shape, D, seed=seed)
    return PEPS(shape=shape, bond_dim=D, contraction_strategy=ZipUp(chi), tensors=T, dtype=dtype), T


def _heis(shape):
    from vmc_amd.operators import HorizontalTwoSiteTerm, LocalHamiltonian, VerticalTwoSiteTerm

    terms = []
    for r in range(shape[0]):
        for c in range(shape[1]):
            if c + 1 < shape[1]:
                terms.append(HorizontalTwoSiteTerm(r, c, HH.HEISENBERG_BOND))
            if r + 1 < shape[0]:
                terms.append(VerticalTwoSiteTerm(r, c, HH.HEISENBERG_BOND))
    return LocalHamiltonian(shape=shape, terms=tuple(terms))


def test_value_and_grad_and_full_gradient():
    from vmc_amd.peps.compat import _value, _value_and_grad

    print("[api] value_and_grad 3x3 D=2", flush=True)
    shape, D, chi = (3, 3), 2, 4
    model, T = _model(shape, D, chi, 1)
    cfg = P.random_configs(9, 32, seed=1)
    amp, grad, p = _value_and_grad(model, torch.tensor(cfg))
    ra, rg, rp = P.value_and_grad(T, cfg, shape, P.ZipUp(chi), D)
    assert np.allclose(amp.cpu().numpy(), ra, rtol=1e-11, atol=0)
    assert np.allclose(grad.cpu().numpy(), rg, rtol=1e-9, atol=1e-12 * np.abs(rg).max())
    assert np.array_equal(p.cpu().numpy(), rp)
    _, gf, pf = _value_and_grad(model, torch.tensor(cfg), full_gradient=True)
    _, rgf, _ = P.value_and_grad(T, cfg, shape, P.ZipUp(chi), D, full_gradient=True)
    assert pf is None and np.allclose(gf.cpu().numpy(), rgf, rtol=1e-9, atol=1e-12 * np.abs(rgf).max())
    assert np.allclose(_value(model, torch.tensor(cfg)).cpu().numpy(), ra, rtol=1e-11)
    # spins in {-1, +1} map to occupancies like spin_to_occupancy
    amp_pm, _, _ = _value_and_grad(model, torch.tensor(2 * cfg - 1))
    assert np.allclose(amp_pm.cpu().numpy(), ra, rtol=1e-11)


def test_peps_apply_custom_vjp():
    from vmc_amd.peps.compat import peps_apply

    print("[api] peps_apply vjp", flush=True)
    shape, D, chi = (3, 3), 2, 4
    model, T = _model(shape, D, chi, 2)
    cfg = P.random_configs(9, 8, seed=2)
    flat = model.flat_params().clone().requires_grad_(True)
    amp = peps_apply(model, flat, torch.tensor(cfg))
    g = torch.arange(1, 9, dtype=torch.float64, device="cuda")
    (amp * g).sum().backward()
    _, rgf, _ = P.value_and_grad(T, cfg, shape, P.ZipUp(chi), D, full_gradient=True)
    ref = (rgf * np.arange(1, 9)[:, None]).sum(0)
    assert np.allclose(flat.grad.cpu().numpy(), ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_local_estimate_matches_oracle():
    from vmc_amd.peps.compat import _value, local_estimate

    print("[api] local_estimate 4x4 Heisenberg", flush=True)
    shape, D, chi = (4, 4), 2, 4
    model, T = _model(shape, D, chi, 3)
    cfg = P.random_configs(16, 64, seed=3)
    amps = _value(model, torch.tensor(cfg))
    e = local_estimate(model, torch.tensor(cfg), _heis(shape), amps).cpu().numpy()
    ra = P.value(T, cfg, shape, P.ZipUp(chi))
    re = P.local_estimate(T, cfg, shape, P.ZipUp(chi), HH.heisenberg(shape), ra)
    assert np.max(np.abs(e - re)) < 1e-9 * max(1.0, np.abs(re).max())


def test_build_mc_kernels_sampler_matches_oracle():
    from vmc_amd.core import _sample_counts, _trim_samples, make_mc_sampler
    from vmc_amd.peps import ChainKeys, build_mc_kernels

    print("[api] build_mc_kernels + make_mc_sampler", flush=True)
    shape, D, chi = (3, 3), 2, 4
    model, T = _model(shape, D, chi, 4)
    op = _heis(shape)
    init_cache, transition, estimate = build_mc_kernels(model, op)
    n_samples, n_chains = 50, 8
    _, _, chain_length, total = _sample_counts(n_samples, n_chains)
    cfg0 = P.random_configs(9, n_chains, seed=21)
    cache = init_cache(None, torch.tensor(cfg0))
    sampler = make_mc_sampler(transition, estimate)
    (cfg_end, key_end, _), (hist, est) = sampler(None, torch.tensor(cfg0), ChainKeys(7), cache, n_steps=chain_length)
    assert key_end.sweep == chain_length
    samples = _trim_samples(hist, total, n_samples).cpu().numpy()
    eloc = _trim_samples(est.local_estimate, total, n_samples).cpu().numpy()
    cache_ref = P.init_cache(T, cfg0, shape, P.ZipUp(chi))
    _, _, h = P.mc_sampler(T, cfg0, cache_ref, shape, P.ZipUp(chi), HH.heisenberg(shape), 7, chain_length)
    ref_samples = P.trim_samples(h["samples"], total, n_samples)
    ref_e = P.trim_samples(h["eloc"], total, n_samples)
    assert np.array_equal(samples, ref_samples)
    assert np.allclose(eloc, ref_e, rtol=1e-9, atol=1e-10)
    o = _trim_samples(est.local_log_derivatives, total, n_samples).cpu().numpy()
    ref_o = P.trim_samples(h["o"], total, n_samples)
    assert np.allclose(o, ref_o, rtol=1e-8, atol=1e-10 * np.abs(ref_o).max())
