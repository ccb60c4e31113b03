"""Complex (complex64 / complex128) plans against the FP64 complex oracle.

Covers the precision-form workload shape (8x8 OBC, D=4, chi=16, complex
tensors; zip-up truncations of 64x64 complex theta matrices) for amp + sliced
gradient, and the sampler / local energies with complex amplitudes.
"""
import numpy as np
import pytest
import torch

from oracle import hamiltonians as HH
from oracle import peps as P

pytestmark = pytest.mark.gpu


def _flat(T):
    return np.concatenate([t.ravel() for row in T for t in row])


def _logdiff(la, amp):
    dl = np.abs(la[:, 0] - np.log(np.abs(amp)))
    dph = np.abs(np.angle(np.exp(1j * (la[:, 1] - np.angle(amp)))))
    return dl, dph


@pytest.mark.parametrize("dtype,tol_log,tol_o", [(torch.complex128, 1e-10, 1e-8), (torch.complex64, 1e-5, 3e-2)])
def test_amp_grad_complex_8x8(dtype, tol_log, tol_o):
    from vmc_amd.engine import AmpGradPlan

    shape, D, chi, B = (8, 8), 4, 16, 8
    print(f"[complex] amp_grad 8x8 D4 chi16 {dtype}", flush=True)
    T = P.random_peps(shape, D, seed=2, dtype=np.complex128)
    cfg = P.random_configs(64, B, seed=2)
    plan = AmpGradPlan(shape, D, chi, dtype=dtype)
    la, o, p = plan.amp_grad(torch.tensor(_flat(T), dtype=dtype, device="cuda"), torch.tensor(cfg), want_p=True)
    torch.cuda.synchronize()
    amp, grad, p_ref = P.value_and_grad(T, cfg, shape, P.ZipUp(chi), D)
    la = la.cpu().numpy()
    dl, dph = _logdiff(la, amp)
    o_ref = grad / amp[:, None]
    err_o = np.max(np.abs(o.cpu().numpy() - o_ref)) / np.max(np.abs(o_ref))
    rel_log = np.max(dl / np.abs(np.log(np.abs(amp))))
    print(f"  rel err log|psi| {rel_log:.2e}  phase {dph.max():.2e}  o {err_o:.2e}", flush=True)
    assert rel_log < tol_log and dph.max() < max(tol_log * 100, 1e-8) * 10
    assert err_o < tol_o
    assert np.array_equal(p.cpu().numpy(), p_ref)


@pytest.mark.parametrize("terms_name", ["tfim", "heisenberg"])
def test_sampler_complex128(terms_name):
    from tests.test_gpu_sampler import _tables
    from vmc_amd.engine import AmpGradPlan

    shape, D, chi, n_chains, n_steps, seed = (4, 4), 2, 4, 8, 2, 7
    print(f"[complex] sampler 4x4 {terms_name}", flush=True)
    terms = HH.tfim(shape, 1.0, 3.04) if terms_name == "tfim" else HH.heisenberg(shape)
    T = P.random_peps(shape, D, seed=seed, dtype=np.complex128)
    cfg0 = P.random_configs(16, n_chains, seed=seed + 17)
    plan = AmpGradPlan(shape, D, chi, dtype=torch.complex128)
    plan.set_terms(_tables(terms, shape))
    params = torch.tensor(_flat(T), dtype=torch.complex128, device="cuda")
    cfg = torch.tensor(cfg0, dtype=torch.int32, device="cuda")
    cache = plan.new_cache(n_chains)
    plan.init_cache(params, cfg, cache)
    outs = []
    for t in range(n_steps):
        la, o, p, e = plan.mc_sweep(params, cfg, cache, seed, t)
        outs.append((cfg.cpu().numpy().copy(), la.cpu().numpy(), o.cpu().numpy(), e.cpu().numpy()))
    cref = P.init_cache(T, cfg0, shape, P.ZipUp(chi))
    _, _, hist = P.mc_sampler(T, cfg0, cref, shape, P.ZipUp(chi), terms, seed, n_steps)
    for t, (c, la, o, e) in enumerate(outs):
        assert np.array_equal(c, hist["samples"][t])
        dl, dph = _logdiff(la, hist["amp"][t])
        err_e = np.max(np.abs(e - hist["eloc"][t])) / max(1.0, np.max(np.abs(hist["eloc"][t])))
        err_o = np.max(np.abs(o - hist["o"][t])) / np.max(np.abs(hist["o"][t]))
        print(f"  sweep {t}: dlog {dl.max():.1e} dphase {dph.max():.1e} o {err_o:.1e} eloc {err_e:.1e}", flush=True)
        assert dl.max() < 1e-10 and dph.max() < 1e-10 and err_o < 1e-9 and err_e < 1e-9
