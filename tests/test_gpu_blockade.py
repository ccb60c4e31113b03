"""GPU parity of the blockade PEPS path (vmc_bl_* C ABI) against oracle.blockade.

oracle.blockade restates src/vmc/peps/blockade/* and is pinned on the CPU
(tests/test_oracle_blockade.py: per-sample masks, dense-Hamiltonian Rydberg
E_loc, finite differences, histogram). The device sampler draws the oracle's
Philox stream, so in FP64 the sampled occupations must be identical and amp,
o, p and E_loc agree to rounding.

With D0 != D1 the charge masks make theta matrices exactly rank deficient, and
a truncating zip-up then depends on which orthonormal vectors the
factorisation puts beside the range of theta (the reference leaves them to
LAPACK's SVD; the next row's local truncation sees them). D0 != D1 is therefore
pinned without truncation (NoTruncation, or chi at least the exact boundary
rank), and truncating strategies are pinned with D0 == D1.
"""
import numpy as np
import pytest
import torch

from oracle import blockade as B
from oracle import peps as P

pytestmark = pytest.mark.gpu

_NP = {torch.float64: np.float64, torch.float32: np.float64, torch.complex128: np.complex128,
       torch.complex64: np.complex128}


def _flat(T):
    return np.concatenate([t.ravel() for row in T for t in row])


def _strategy(name, chi):
    from vmc_amd import _lib

    return {"zipup": (_lib.VMC_ZIPUP, P.ZipUp(chi)), "exact": (_lib.VMC_NO_TRUNCATION, P.NoTruncation()),
            "variational": (_lib.VMC_VARIATIONAL, P.Variational(chi, n_sweeps=2)),
            "density": (_lib.VMC_DENSITY_MATRIX, P.DensityMatrix(chi))}[name]


def _per_element(cfg, p_sites):
    """Per-site combined slice indices repeated over each site's sliced parameters."""
    R, C = cfg.shape
    reps = [int(np.prod(B.site_shape(cfg, r, c)[2:])) for r in range(R) for c in range(C)]
    return np.repeat(p_sites, reps, axis=1).astype(np.int8)


def _tensors(cfg, seed, dtype):
    T = B.random_tensors(cfg, seed=seed, dtype=_NP[dtype])
    if dtype == torch.float32 or dtype == torch.complex64:
        T = [[t.astype(np.complex64 if dtype.is_complex else np.float32).astype(_NP[dtype]) for t in row] for row in T]
    return T


@pytest.mark.parametrize("shape,D0,D1,chi,strat,dtype", [
    ((2, 3), 2, 2, 0, "exact", torch.float64),
    ((3, 3), 2, 1, 0, "exact", torch.float64),
    ((4, 4), 2, 1, 16, "zipup", torch.float64),
    ((4, 5), 1, 2, 16, "zipup", torch.float64),
    ((4, 5), 2, 2, 4, "zipup", torch.float64),
    ((5, 5), 2, 2, 4, "variational", torch.float64),
    ((3, 4), 2, 1, 8, "zipup", torch.complex128),
    ((4, 4), 3, 3, 6, "zipup", torch.complex128),
    ((4, 4), 2, 2, 4, "zipup", torch.float32),
    ((4, 4), 2, 2, 2, "density", torch.float64),  # chi <= D: no kept eigenvector in a null space
    ((4, 4), 2, 1, 16, "density", torch.float64),
])
def test_blockade_amp_grad_matches_oracle(shape, D0, D1, chi, strat, dtype):
    from vmc_amd.engine import BLPlan

    code, ostrat = _strategy(strat, chi)
    print(f"[bl] amp_grad shape={shape} D0={D0} D1={D1} chi={chi} {strat} {dtype}", flush=True)
    cfg = B.BLConfig(shape, D0, D1)
    T = _tensors(cfg, 1, dtype)
    occ = B.random_independent_sets(shape, 16, seed=2)
    bad = occ[:3].copy()
    bad[0, 0, 0] = bad[0, 0, 1] = 1  # horizontal neighbours
    bad[1, 0, 1] = bad[1, 1, 1] = 1  # vertical neighbours
    bad[2, shape[0] - 1, shape[1] - 2] = bad[2, shape[0] - 1, shape[1] - 1] = 1
    samples = np.concatenate([occ, bad])
    plan = BLPlan(shape, D0, D1, max(chi, 1), strategy=code, dtype=dtype)
    params = torch.tensor(_flat(T), dtype=dtype, device="cuda")
    assert params.numel() == plan.num_params
    flat_s = torch.tensor(samples.reshape(len(samples), -1))
    la, o, p = plan.amp_grad(params, flat_s, want_p=True)
    sl, ok = plan.slices(flat_s)
    torch.cuda.synchronize()
    la, o, p = la.cpu().numpy(), o.cpu().numpy(), p.cpu().numpy()
    assert np.array_equal(sl.cpu().numpy()[:16], B.slices(cfg, occ))
    assert np.array_equal(ok.cpu().numpy(), B.valid(samples))
    amp, grad, p_ref = B.value_and_grad(cfg, T, occ, ostrat)
    n = len(occ)
    logabs = np.log(np.abs(amp))
    rel_log = np.max(np.abs(la[:n, 0] - logabs) / np.maximum(1.0, np.abs(logabs)))
    if np.iscomplexobj(amp):
        phase_err = np.max(np.abs(np.exp(1j * la[:n, 1]) - amp / np.abs(amp)))
    else:
        phase_err = float(not np.array_equal(np.where(la[:n, 1] != 0, -1.0, 1.0), np.sign(amp)))
    o_ref = grad / amp[:, None]
    err_o = np.max(np.abs(o[:n] - o_ref)) / np.max(np.abs(o_ref))
    print(f"  rel err log|psi| {rel_log:.2e} phase {phase_err:.1e} o {err_o:.2e}", flush=True)
    tol_l, tol_o = (1e-11, 1e-9) if dtype in (torch.float64, torch.complex128) else (1e-4, 1e-3)
    assert phase_err < max(tol_l, 1e-12) * 100 and rel_log < tol_l and err_o < tol_o
    assert np.array_equal(p[:n], _per_element(cfg, p_ref))
    # blockade violations: amplitude 0 (compat.py:47-62), zero log-derivatives
    assert np.all(np.isneginf(la[n:, 0])) and np.all(o[n:] == 0)


@pytest.mark.parametrize("shape,D0,D1,chi,strat,dtype", [
    ((3, 3), 2, 1, 0, "exact", torch.float64),
    ((4, 4), 2, 2, 4, "zipup", torch.float64),
    ((4, 5), 2, 1, 16, "zipup", torch.float64),
    ((4, 5), 2, 2, 4, "zipup", torch.float64),
    ((4, 4), 2, 2, 4, "variational", torch.float64),
    ((3, 4), 2, 2, 4, "zipup", torch.complex128),
])
def test_blockade_sampler_matches_oracle(shape, D0, D1, chi, strat, dtype):
    """Row-pair sweeps, the last row's single-row sweep, env-gradients and the
    Rydberg E_loc (model.py:295-994): same Philox stream as the oracle."""
    from vmc_amd.engine import BLPlan
    from vmc_amd.operators.local_terms import to_tables
    from vmc_amd.peps.blockade import rydberg_hamiltonian

    code, ostrat = _strategy(strat, chi)
    print(f"[bl] sampler shape={shape} D0={D0} D1={D1} chi={chi} {strat} {dtype}", flush=True)
    cfg = B.BLConfig(shape, D0, D1)
    T = _tensors(cfg, 3, dtype)
    n_chains, n_steps, seed = 16, 3, 7
    occ0 = B.random_independent_sets(shape, n_chains, seed=4)
    omega, delta, vnnn = 1.1, 0.6, 0.35
    plan = BLPlan(shape, D0, D1, max(chi, 1), strategy=code, dtype=dtype)
    plan.set_terms(to_tables(rydberg_hamiltonian(shape, omega, delta, vnnn), 2))
    params = torch.tensor(_flat(T), dtype=dtype, device="cuda")
    samples = torch.tensor(occ0.reshape(n_chains, -1), dtype=torch.int32, device="cuda")
    cache = plan.new_cache(n_chains)
    plan.init_cache(params, samples, cache)
    outs = []
    for t in range(n_steps):
        la, o, p, e = plan.mc_sweep(params, samples, cache, seed, t)
        outs.append((samples.cpu().numpy().copy(), la.cpu().numpy(), o.cpu().numpy(), p.cpu().numpy(),
                     e.cpu().numpy()))
    _, _, hist = B.mc_sampler(cfg, T, occ0, ostrat, B.rydberg_terms(shape, omega, delta, vnnn), seed, n_steps)
    moved = 0
    for t, (sm, la, o, p, e) in enumerate(outs):
        assert np.array_equal(sm, hist["samples"][t]), f"sweep {t}: samples differ"
        assert B.valid(sm.reshape(n_chains, *shape)).all()
        moved += int(np.sum(np.any(sm != (outs[t - 1][0] if t else occ0.reshape(n_chains, -1)), axis=1)))
        amp = hist["amp"][t]
        dl = np.max(np.abs(la[:, 0] - np.log(np.abs(amp))))
        err_o = np.max(np.abs(o - hist["o"][t])) / np.max(np.abs(hist["o"][t]))
        err_e = np.max(np.abs(e - hist["eloc"][t])) / max(1.0, np.max(np.abs(hist["eloc"][t])))
        print(f"  sweep {t}: dlog {dl:.1e} o {err_o:.1e} eloc {err_e:.1e}", flush=True)
        assert dl < 1e-10 and err_o < 1e-9 and err_e < 1e-9
        assert np.array_equal(p, _per_element(cfg, hist["p"][t]))
    assert moved > 0  # the sweeps accepted flips


def test_blockade_sweep_refuses_two_site_terms():
    from vmc_amd.engine import BLPlan
    from vmc_amd._lib import VMCError
    from vmc_amd.operators import HorizontalTwoSiteTerm, LocalHamiltonian
    from vmc_amd.operators.local_terms import to_tables

    plan = BLPlan((3, 3), 2, 2, 4)
    op = LocalHamiltonian(shape=(3, 3), terms=(HorizontalTwoSiteTerm(row=0, col=0, op=np.eye(4)),))
    plan.set_terms(to_tables(op, 2))
    params = 0.3 * torch.randn(plan.num_params, dtype=torch.float64, device="cuda")
    samples = torch.zeros((2, 9), dtype=torch.int32, device="cuda")
    cache = plan.new_cache(2)
    plan.init_cache(params, samples, cache)
    with pytest.raises(VMCError, match="diagonal and one-site"):
        plan.mc_sweep(params, samples, cache, 1, 0)


def _model(cfg, T, strat, dtype=torch.float64):
    from vmc_amd.peps.blockade import BlockadePEPS, BlockadePEPSConfig

    return BlockadePEPS(config=BlockadePEPSConfig(cfg.shape, cfg.D0, cfg.D1, dtype=dtype),
                        contraction_strategy=strat, tensors=T)


def test_blockade_apply_vjp_matches_oracle_full_gradient():
    from vmc_amd.peps.blockade import blockade_apply
    from vmc_amd.peps.strategy import ZipUp

    print("[bl] blockade_apply VJP 4x4 D0=2 D1=1", flush=True)
    cfg = B.BLConfig((4, 4), 2, 1)
    T = B.random_tensors(cfg, seed=5)
    model = _model(cfg, T, ZipUp(16))
    occ = B.random_independent_sets((4, 4), 8, seed=6)
    flat = model.flat_params().requires_grad_(True)
    amp = blockade_apply(flat, torch.tensor(occ.reshape(8, -1)), (4, 4), model.config, model.strategy, model=model)
    amp.sum().backward()
    a_ref, g_ref, p_ref = B.value_and_grad(cfg, T, occ, P.ZipUp(16))
    assert np.max(np.abs(amp.detach().cpu().numpy() - a_ref)) <= 1e-10 * np.max(np.abs(a_ref))
    # scatter the oracle's sliced gradients into the full parameter vector
    full = np.zeros(flat.numel())
    off_full, off_sl = 0, 0
    for r in range(4):
        for c in range(4):
            shp = B.site_shape(cfg, r, c)
            n = int(np.prod(shp[2:]))
            for z in range(8):
                full[off_full + p_ref[z, r * 4 + c] * n + np.arange(n)] += g_ref[z, off_sl:off_sl + n]
            off_sl += n
            off_full += n * shp[0] * shp[1]
    g = flat.grad.cpu().numpy()
    assert np.max(np.abs(g - full)) <= 1e-9 * np.max(np.abs(full))


def test_blockade_build_mc_kernels_and_make_mc_sampler():
    from vmc_amd.core import make_mc_sampler
    from vmc_amd.peps.blockade import build_mc_kernels, rydberg_hamiltonian
    from vmc_amd.peps.kernels import ChainKeys
    from vmc_amd.peps.strategy import ZipUp

    print("[bl] build_mc_kernels + make_mc_sampler 4x4 D0=2 D1=2", flush=True)
    cfg = B.BLConfig((4, 4), 2, 2)
    T = B.random_tensors(cfg, seed=9)
    model = _model(cfg, T, ZipUp(4))
    H = rydberg_hamiltonian((4, 4), 1.0, 0.8, 0.3)
    init_cache, transition, estimate = build_mc_kernels(model, H)
    sampler = make_mc_sampler(transition, estimate)
    occ0 = B.random_independent_sets((4, 4), 8, seed=10)
    s0 = torch.tensor(occ0.reshape(8, -1), dtype=torch.int32, device="cuda")
    cache = init_cache(None, s0)
    _, _, ref = B.mc_sampler(cfg, T, occ0, P.ZipUp(4), B.rydberg_terms((4, 4), 1.0, 0.8, 0.3), 5, 2)
    (_, _, _), (hist, est) = sampler(None, s0, ChainKeys(5, 0), cache, n_steps=2)
    assert np.array_equal(hist.cpu().numpy(), np.stack(ref["samples"]))
    e = est.local_estimate.cpu().numpy()
    assert np.allclose(e, np.stack(ref["eloc"]), rtol=1e-9, atol=1e-9)
    assert np.allclose(est.local_log_derivatives.cpu().numpy(), np.stack(ref["o"]), rtol=1e-9,
                       atol=1e-10 * np.abs(np.stack(ref["o"])).max())
