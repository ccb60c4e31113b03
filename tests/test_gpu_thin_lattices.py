"""Lattices of one row or one column (chains) and a single site, against the oracle.

The reference runs PEPS chains of shape (1, n) in its physics tests
(tests/test_physics_models.py:376-382) and a (1, 1) PEPS in its driver tests
(tests/test_tdvp_driver.py:39-83). On these shapes every boundary bond is 1
along the short direction, the first column is also the last (one column) and
no cut truncates anything (one row), so the device's edge handling is what is
exercised: amplitudes, env-gradients, the sampler's Philox stream and E_loc,
for the standard and the blockade PEPS.
"""
import numpy as np
import pytest
import torch

from oracle import blockade as B
from oracle import exact as E
from oracle import hamiltonians as HH
from oracle import peps as P

pytestmark = pytest.mark.gpu


def _flat(T):
    return np.concatenate([t.ravel() for row in T for t in row])


@pytest.mark.parametrize("shape,D,chi,strat", [
    ((1, 6), 3, 0, "exact"), ((1, 6), 3, 2, "zipup"), ((6, 1), 3, 0, "exact"), ((6, 1), 2, 2, "zipup"),
    ((1, 1), 2, 0, "exact"), ((2, 1), 2, 0, "exact"), ((1, 2), 2, 0, "exact"),
])
def test_thin_amp_grad_matches_oracle(shape, D, chi, strat):
    from vmc_amd import _lib
    from vmc_amd.engine import AmpGradPlan

    print(f"[thin] amp_grad {shape} D={D} {strat} chi={chi}", flush=True)
    code, ostrat = ((_lib.VMC_NO_TRUNCATION, P.NoTruncation()) if strat == "exact"
                    else (_lib.VMC_ZIPUP, P.ZipUp(chi)))
    T = P.random_peps(shape, D, seed=3)
    cfg = P.random_configs(shape[0] * shape[1], 16, seed=4)
    plan = AmpGradPlan(shape, D, max(chi, 1), strategy=code, dtype=torch.float64)
    params = torch.tensor(_flat(T), dtype=torch.float64, device="cuda")
    la, o, p = plan.amp_grad(params, torch.tensor(cfg), want_p=True)
    torch.cuda.synchronize()
    la, o, p = la.cpu().numpy(), o.cpu().numpy(), p.cpu().numpy()
    amp, grad, p_ref = P.value_and_grad(T, cfg, shape, ostrat, D)
    if strat == "exact":
        full = np.array([E.full_network_amplitude(T, x, shape) for x in cfg])
        assert np.max(np.abs(amp - full)) <= 1e-12 * np.max(np.abs(full))
    err_l = np.max(np.abs(la[:, 0] - np.log(np.abs(amp))))
    assert np.array_equal(np.where(la[:, 1] != 0, -1.0, 1.0), np.sign(amp))
    err_o = np.max(np.abs(o - grad / amp[:, None])) / np.max(np.abs(grad / amp[:, None]))
    print(f"  |dlog| {err_l:.1e}  o {err_o:.1e}", flush=True)
    assert err_l < 1e-12 and err_o < 1e-11
    assert np.array_equal(p, p_ref)


@pytest.mark.parametrize("shape,D,terms", [
    ((1, 6), 2, "heisenberg"), ((5, 1), 2, "heisenberg"), ((1, 1), 2, "tfim"), ((2, 1), 2, "tfim"),
])
def test_thin_sampler_matches_oracle(shape, D, terms):
    from vmc_amd import _lib
    from vmc_amd.engine import AmpGradPlan

    from test_gpu_sampler import _tables  # BucketedTerms -> device tables

    print(f"[thin] sampler {shape} D={D} {terms}", flush=True)
    H = HH.heisenberg(shape) if terms == "heisenberg" else HH.tfim(shape, 1.0, 3.04)
    T = P.random_peps(shape, D, seed=7)
    n_chains, n_steps, seed = 16, 3, 9
    cfg0 = P.random_configs(shape[0] * shape[1], n_chains, seed=8)
    plan = AmpGradPlan(shape, D, 1, strategy=_lib.VMC_NO_TRUNCATION, dtype=torch.float64)
    plan.set_terms(_tables(H, shape))
    params = torch.tensor(_flat(T), dtype=torch.float64, device="cuda")
    cfg = torch.tensor(cfg0, dtype=torch.int32, device="cuda")
    cache = plan.new_cache(n_chains)
    plan.init_cache(params, cfg, cache)
    outs = []
    for t in range(n_steps):
        la, o, p, e = plan.mc_sweep(params, cfg, cache, seed, t)
        outs.append((cfg.cpu().numpy().copy(), la.cpu().numpy(), o.cpu().numpy(), p.cpu().numpy(), e.cpu().numpy()))
    strat = P.NoTruncation()
    _, _, hist = P.mc_sampler(T, cfg0, P.init_cache(T, cfg0, shape, strat), shape, strat, H, seed, n_steps)
    for t, (sm, la, o, p, e) in enumerate(outs):
        assert np.array_equal(sm, hist["samples"][t]), f"sweep {t}: samples differ"
        err_l = np.max(np.abs(la[:, 0] - np.log(np.abs(hist["amp"][t]))))
        err_o = np.max(np.abs(o - hist["o"][t])) / np.max(np.abs(hist["o"][t]))
        err_e = np.max(np.abs(e - hist["eloc"][t])) / max(1.0, np.max(np.abs(hist["eloc"][t])))
        print(f"  sweep {t}: |dlog| {err_l:.1e} o {err_o:.1e} eloc {err_e:.1e}", flush=True)
        assert err_l < 1e-11 and err_o < 1e-10 and err_e < 1e-10
        assert np.array_equal(p, hist["p"][t])


@pytest.mark.parametrize("shape,D0,D1", [((1, 5), 2, 1), ((5, 1), 2, 2), ((1, 1), 2, 2), ((2, 1), 2, 1)])
def test_thin_blockade_matches_oracle(shape, D0, D1):
    from vmc_amd import _lib
    from vmc_amd.engine import BLPlan
    from vmc_amd.operators.local_terms import to_tables
    from vmc_amd.peps.blockade import rydberg_hamiltonian

    print(f"[thin] blockade {shape} D0={D0} D1={D1}", flush=True)
    cfg = B.BLConfig(shape, D0, D1)
    T = B.random_tensors(cfg, seed=5)
    occ = B.random_independent_sets(shape, 16, seed=6)
    plan = BLPlan(shape, D0, D1, 1, strategy=_lib.VMC_NO_TRUNCATION, dtype=torch.float64)
    params = torch.tensor(_flat(T), dtype=torch.float64, device="cuda")
    la, o, p = plan.amp_grad(params, torch.tensor(occ.reshape(16, -1)), want_p=True)
    amp, grad, _ = B.value_and_grad(cfg, T, occ, P.NoTruncation())
    la, o = la.cpu().numpy(), o.cpu().numpy()
    assert np.max(np.abs(la[:, 0] - np.log(np.abs(amp)))) < 1e-12
    assert np.max(np.abs(o - grad / amp[:, None])) <= 1e-11 * np.max(np.abs(grad / amp[:, None]))
    # the sampler: the last row's single-row sweep only (one row) or one column of row pairs
    omega, delta, vnnn = 1.2, 0.4, 0.3
    plan.set_terms(to_tables(rydberg_hamiltonian(shape, omega, delta, vnnn), 2))
    samples = torch.tensor(occ.reshape(16, -1), dtype=torch.int32, device="cuda")
    cache = plan.new_cache(16)
    plan.init_cache(params, samples, cache)
    outs = []
    for t in range(3):
        la, o, p, e = plan.mc_sweep(params, samples, cache, 4, t)
        outs.append((samples.cpu().numpy().copy(), la.cpu().numpy(), e.cpu().numpy()))
    _, _, hist = B.mc_sampler(cfg, T, occ, P.NoTruncation(), B.rydberg_terms(shape, omega, delta, vnnn), 4, 3)
    for t, (sm, la, e) in enumerate(outs):
        assert np.array_equal(sm, hist["samples"][t]), f"sweep {t}: samples differ"
        assert np.max(np.abs(la[:, 0] - np.log(np.abs(hist["amp"][t])))) < 1e-11
        assert np.max(np.abs(e - hist["eloc"][t])) <= 1e-10 * max(1.0, np.max(np.abs(hist["eloc"][t])))
