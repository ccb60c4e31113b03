"""GPU parity of vmc_amp_grad (C ABI) against the FP64 oracle.

Oracle: oracle.peps.value_and_grad, the restatement of _value_and_grad
(src/vmc/peps/standard/compat.py:94-133) with LAPACK SVD zip-up.
"""
import numpy as np
import pytest
import torch

from oracle import peps as P

pytestmark = pytest.mark.gpu


def _flat(tensors):
    return np.concatenate([t.ravel() for row in tensors for t in row])


def _run(shape, D, chi, dtype, B, seed, strategy=None):
    from vmc_amd import _lib
    from vmc_amd.engine import AmpGradPlan

    tensors = P.random_peps(shape, D, seed=seed)
    cfg = P.random_configs(shape[0] * shape[1], B, seed=seed)
    print(f"[amp_grad] shape={shape} D={D} chi={chi} dtype={dtype} B={B} seed={seed}", flush=True)
    plan = AmpGradPlan(shape, D, chi, dtype=dtype,
                       strategy=_lib.VMC_ZIPUP if strategy is None else strategy)
    params = torch.tensor(_flat(tensors), dtype=dtype, device="cuda")
    log_amp, o, p = plan.amp_grad(params, torch.tensor(cfg), want_p=True)
    torch.cuda.synchronize()
    strat = P.ZipUp(chi) if strategy is None else P.NoTruncation()
    amp, grad, p_ref = P.value_and_grad(tensors, cfg, shape, strat, D)
    return log_amp.cpu().numpy(), o.cpu().numpy(), p.cpu().numpy(), amp, grad, p_ref


def _check(shape, D, chi, dtype, B, seed, tol_log, tol_o, strategy=None):
    la, o, p, amp, grad, p_ref = _run(shape, D, chi, dtype, B, seed, strategy)
    ref_log = np.log(np.abs(amp))
    ref_sign = np.sign(amp)
    err_log = np.max(np.abs(la[:, 0] - ref_log) / np.maximum(1.0, np.abs(ref_log)))
    sign = np.where(la[:, 1] != 0, -1.0, 1.0)
    assert np.array_equal(sign, ref_sign), "amplitude signs differ"
    o_ref = grad / amp[:, None]
    err_o = np.max(np.abs(o - o_ref)) / np.max(np.abs(o_ref))
    print(f"  rel err log psi {err_log:.3e}  rel err o {err_o:.3e}", flush=True)
    assert err_log < tol_log
    assert err_o < tol_o
    assert np.array_equal(p, p_ref)


def test_cfg1_4x4_f64():
    _check((4, 4), 2, 4, torch.float64, 64, 0, 1e-12, 1e-10)


def test_cfg1_4x4_f32():
    _check((4, 4), 2, 4, torch.float32, 64, 0, 1e-6, 1e-4)


def test_exact_3x3_D3_f64():
    from vmc_amd import _lib

    _check((3, 3), 3, 0, torch.float64, 32, 5, 1e-12, 1e-10, strategy=_lib.VMC_NO_TRUNCATION)


def test_cfg3_10x10_f64():
    _check((10, 10), 4, 32, torch.float64, 4, 1, 1e-10, 1e-7)


def test_cfg3_10x10_f32_full_batch_vs_conditioning():
    """FP32 contractions vs the FP64 path on identical (FP32-representable) inputs
    over the metric's whole batch (10x10, D=4, chi=32, 4096 samples).

    The 1e-5 bar on log psi holds at the median but not in the tail. The tail is
    the conditioning of the truncated amplitude: the FP64 path alone, fed the
    unrounded FP64 draw instead of its FP32 cast, moves log psi by the same order.
    Asserted: median <= 1e-5, and the p99 and max of the FP32 error are within 10x
    of the p99 and max of that input-rounding change."""
    from vmc_amd import synthetic
    from vmc_amd.engine import AmpGradPlan

    shape, D, chi, B = (10, 10), 4, 32, 4096
    print(f"[amp_grad] fp32 vs fp64 full batch {shape} D={D} chi={chi} B={B}", flush=True)
    draw = synthetic.flatten(synthetic.random_site_tensors(shape, D, seed=0))
    cast = draw.astype(np.float32).astype(np.float64)
    cfg = torch.tensor(synthetic.random_configs(100, B, seed=0), device="cuda")
    la = {}
    for name, dtype, vals in (("f32", torch.float32, cast), ("f64", torch.float64, cast),
                              ("f64_draw", torch.float64, draw)):
        plan = AmpGradPlan(shape, D, chi, dtype=dtype)
        la[name] = plan.amp_grad(torch.tensor(vals, dtype=dtype, device="cuda"), cfg, grad=False)[0].cpu().numpy()
    ref = np.maximum(1.0, np.abs(la["f64"][:, 0]))
    e32 = np.abs(la["f32"][:, 0] - la["f64"][:, 0]) / ref
    ein = np.abs(la["f64_draw"][:, 0] - la["f64"][:, 0]) / ref
    q = lambda x: (np.median(x), np.quantile(x, 0.99), np.max(x))
    print("  fp32 error median/p99/max %.2e %.2e %.2e" % q(e32), flush=True)
    print("  input-rounding change median/p99/max %.2e %.2e %.2e" % q(ein), flush=True)
    assert np.median(e32) <= 1e-5
    assert np.quantile(e32, 0.99) <= 10 * np.quantile(ein, 0.99)
    assert np.max(e32) <= 10 * np.max(ein)
