"""GPU parity of the Variational strategy (QR-initialised one-site ALS,
src/vmc/peps/common/strategy.py:179-427) against the FP64 oracle.

Oracle: oracle.peps with P.Variational(chi, n_sweeps), whose restatement is
pinned on the CPU (chi >= exact rank == exact contraction; two-site ALS
converges to the rank-k truncated SVD; invariance under QR phase gauges).
The device QR differs from LAPACK's by such a gauge, so amplitudes, o and
E_loc must agree to rounding.
"""
import numpy as np
import pytest
import torch

from oracle import hamiltonians as HH
from oracle import peps as P

pytestmark = pytest.mark.gpu


def _flat(tensors):
    return np.concatenate([t.ravel() for row in tensors for t in row])


def _amp_grad(shape, D, chi, dtype, B, seed, n_sweeps=2, complex_=False):
    from vmc_amd import _lib
    from vmc_amd.engine import AmpGradPlan

    print(f"[variational] amp_grad shape={shape} D={D} chi={chi} sweeps={n_sweeps} dtype={dtype} B={B}",
          flush=True)
    T = P.random_peps(shape, D, seed=seed, dtype=np.complex128 if complex_ else np.float64)
    cfg = P.random_configs(shape[0] * shape[1], B, seed=seed)
    plan = AmpGradPlan(shape, D, chi, dtype=dtype, strategy=_lib.VMC_VARIATIONAL, n_sweeps=n_sweeps)
    la, o, p = plan.amp_grad(torch.tensor(_flat(T), dtype=dtype, device="cuda"), torch.tensor(cfg), want_p=True)
    torch.cuda.synchronize()
    amp, grad, p_ref = P.value_and_grad(T, cfg, shape, P.Variational(chi, n_sweeps=n_sweeps), D)
    la = la.cpu().numpy()
    rel_log = np.max(np.abs(la[:, 0] - np.log(np.abs(amp))) / np.maximum(1.0, np.abs(np.log(np.abs(amp)))))
    dph = np.abs(np.angle(np.exp(1j * (la[:, 1] - np.angle(amp)))))
    o_ref = grad / amp[:, None]
    err_o = np.max(np.abs(o.cpu().numpy() - o_ref)) / np.max(np.abs(o_ref))
    print(f"  rel err log|psi| {rel_log:.2e}  phase {dph.max():.2e}  o {err_o:.2e}", flush=True)
    assert np.array_equal(p.cpu().numpy(), p_ref)
    return rel_log, dph.max(), err_o


@pytest.mark.parametrize("n_sweeps", [0, 1, 2])
def test_variational_4x4_f64_sweeps(n_sweeps):
    rl, ph, eo = _amp_grad((4, 4), 2, 4, torch.float64, 64, 0, n_sweeps=n_sweeps)
    assert rl < 1e-11 and ph < 1e-9 and eo < 1e-9


def test_variational_6x6_D3_f64():
    rl, ph, eo = _amp_grad((6, 6), 3, 9, torch.float64, 16, 1)
    assert rl < 1e-11 and ph < 1e-9 and eo < 1e-8


def test_variational_10x10_D4_chi32_f64():
    rl, ph, eo = _amp_grad((10, 10), 4, 32, torch.float64, 2, 1)
    assert rl < 1e-10 and ph < 1e-9 and eo < 1e-7


def test_variational_4x4_f32():
    # FP32 contractions (FP64 QR): the ALS sweeps add ~5x the zip-up's FP32
    # contractions, so the bound is the 1e-5 relative log-psi bar
    rl, ph, eo = _amp_grad((4, 4), 2, 4, torch.float32, 64, 0)
    assert rl < 1e-5 and ph < 1e-9 and eo < 1e-3


def test_variational_8x8_D4_chi16_complex128():
    rl, ph, eo = _amp_grad((8, 8), 4, 16, torch.complex128, 4, 2, complex_=True)
    assert rl < 1e-11 and ph < 1e-8 and eo < 1e-8


def test_variational_sampler_tfim_4x4():
    from tests.test_gpu_sampler import _tables
    from vmc_amd import _lib
    from vmc_amd.engine import AmpGradPlan

    shape, D, chi, n_chains, n_steps, seed = (4, 4), 2, 4, 16, 2, 3
    print(f"[variational] sampler 4x4 TFIM chains={n_chains}", flush=True)
    terms = HH.tfim(shape, 1.0, 3.04)
    T = P.random_peps(shape, D, seed=seed)
    cfg0 = P.random_configs(16, n_chains, seed=seed + 17)
    plan = AmpGradPlan(shape, D, chi, dtype=torch.float64, strategy=_lib.VMC_VARIATIONAL, n_sweeps=2)
    plan.set_terms(_tables(terms, shape))
    params = torch.tensor(_flat(T), dtype=torch.float64, device="cuda")
    cfg = torch.tensor(cfg0, dtype=torch.int32, device="cuda")
    cache = plan.new_cache(n_chains)
    plan.init_cache(params, cfg, cache)
    outs = []
    for t in range(n_steps):
        la, o, p, e = plan.mc_sweep(params, cfg, cache, seed, t)
        outs.append((cfg.cpu().numpy().copy(), la.cpu().numpy(), o.cpu().numpy(), e.cpu().numpy()))
    strat = P.Variational(chi, n_sweeps=2)
    cref = P.init_cache(T, cfg0, shape, strat)
    _, _, hist = P.mc_sampler(T, cfg0, cref, shape, strat, terms, seed, n_steps)
    for t, (c, la, o, e) in enumerate(outs):
        assert np.array_equal(c, hist["samples"][t])
        dl = np.max(np.abs(la[:, 0] - np.log(np.abs(hist["amp"][t]))))
        err_e = np.max(np.abs(e - hist["eloc"][t])) / max(1.0, np.max(np.abs(hist["eloc"][t])))
        err_o = np.max(np.abs(o - hist["o"][t])) / np.max(np.abs(hist["o"][t]))
        print(f"  sweep {t}: dlog {dl:.1e} o {err_o:.1e} eloc {err_e:.1e}", flush=True)
        assert dl < 1e-10 and err_o < 1e-9 and err_e < 1e-9


# Panels beyond the 128 x 32 register tile run in global scratch (vqr_factor_g):
# chi = 40 gives 160 x 40 QR panels and 160-row QR initialisations.
def test_variational_generic_panels_6x6_chi40_f64():
    rl, ph, eo = _amp_grad((6, 6), 4, 40, torch.float64, 3, 31)
    assert rl < 1e-11 and ph < 1e-9 and eo < 1e-8


def test_variational_generic_panels_6x6_chi40_c128():
    rl, ph, eo = _amp_grad((6, 6), 4, 40, torch.complex128, 2, 32, complex_=True)
    assert rl < 1e-11 and ph < 1e-9 and eo < 1e-8


def test_variational_generic_panels_fp32_vs_fp64():
    from vmc_amd import _lib
    from vmc_amd.engine import AmpGradPlan

    shape, D, chi, B = (6, 6), 4, 40, 32
    T = P.random_peps(shape, D, seed=33)
    cast = _flat(T).astype(np.float32).astype(np.float64)
    cfg = torch.tensor(P.random_configs(36, B, seed=34), device="cuda")
    out = {}
    for dt in (torch.float32, torch.float64):
        plan = AmpGradPlan(shape, D, chi, dtype=dt, strategy=_lib.VMC_VARIATIONAL)
        out[dt] = plan.amp_grad(torch.tensor(cast, dtype=dt, device="cuda"), cfg, grad=False)[0].cpu().numpy()
    e = np.abs(out[torch.float32][:, 0] - out[torch.float64][:, 0]) / np.maximum(1.0, np.abs(out[torch.float64][:, 0]))
    print(f"  generic variational FP32 vs FP64: median {np.median(e):.2e} max {e.max():.2e}", flush=True)
    assert np.median(e) < 1e-5
