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


def test_cfg3_10x10_f32():
    _check((10, 10), 4, 32, torch.float32, 4, 1, 1e-5, 1e-3)
