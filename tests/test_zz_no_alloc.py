"""vmc_amp_grad_noalloc (static device scratch, no allocation) against vmc_amp_grad.

Both entry points run the same kernels on the same workspace layout, so the
outputs must be bitwise identical. Two calls in a row per case, the second with
a smaller batch, exercise reuse of the shared scratch; a batch whose workspace
exceeds the scratch is refused with the library's message.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import peps as P

pytestmark = pytest.mark.gpu


def _flat(tensors):
    return np.concatenate([t.ravel() for row in tensors for t in row])


def _noalloc(plan, params, cfg, want_o=True):
    from vmc_amd._lib import LIB, check

    B = cfg.shape[0]
    la = torch.full((B, 2), float("nan"), dtype=torch.float64, device="cuda")
    o = torch.full((B, plan.num_sliced), float("nan"), dtype=plan.dtype, device="cuda") if want_o else None
    p = torch.full((B, plan.num_sliced), -1, dtype=torch.int8, device="cuda") if want_o else None
    stream = torch.cuda.current_stream().cuda_stream
    ptr = lambda t: ctypes.c_void_p(t.data_ptr() if t is not None else 0)
    check(LIB.vmc_amp_grad_noalloc(plan._h, ptr(params), ptr(cfg), int(B), ptr(la), ptr(o), ptr(p),
                                   ctypes.c_void_p(stream)))
    return la, o, p


@pytest.mark.parametrize("shape,D,chi,dtype,batches", [
    ((10, 10), 4, 32, torch.float64, (16, 5)),
    ((4, 4), 2, 4, torch.float32, (256, 100)),
    ((6, 6), 3, 9, torch.complex128, (32, 7)),
])
def test_noalloc_matches_amp_grad(shape, D, chi, dtype, batches):
    from vmc_amd.engine import AmpGradPlan

    cplx = dtype.is_complex
    T = P.random_peps(shape, D, seed=3, dtype=np.complex128 if cplx else np.float64)
    plan = AmpGradPlan(shape, D, chi, dtype=dtype)
    params = torch.tensor(_flat(T), dtype=dtype, device="cuda")
    for i, B in enumerate(batches):
        print(f"[noalloc] {shape} D={D} chi={chi} {dtype} call {i} batch {B}", flush=True)
        cfg = torch.tensor(P.random_configs(shape[0] * shape[1], B, seed=10 + i), dtype=torch.int32, device="cuda")
        la, o, p = _noalloc(plan, params, cfg)
        ref_la, ref_o, ref_p = plan.amp_grad(params, cfg, want_p=True)
        torch.cuda.synchronize()
        assert torch.equal(la, ref_la)
        assert torch.equal(o, ref_o)
        assert torch.equal(p, ref_p)
        la_only, _, _ = _noalloc(plan, params, cfg, want_o=False)
        torch.cuda.synchronize()
        assert torch.equal(la_only, ref_la)


def test_noalloc_refuses_batches_beyond_the_scratch():
    from vmc_amd import _lib
    from vmc_amd.engine import AmpGradPlan

    plan = AmpGradPlan((10, 10), 4, 32, dtype=torch.float64)
    B = 4096
    need = int(_lib.LIB.vmc_amp_grad_workspace_size(plan._h, B))
    assert need > _lib.VMC_NOALLOC_SCRATCH_BYTES
    cfg = torch.zeros((B, 100), dtype=torch.int32, device="cuda")
    la = torch.empty((B, 2), dtype=torch.float64, device="cuda")
    params = torch.zeros(plan.num_params, dtype=torch.float64, device="cuda")
    rc = _lib.LIB.vmc_amp_grad_noalloc(plan._h, ctypes.c_void_p(params.data_ptr()), ctypes.c_void_p(cfg.data_ptr()),
                                       B, ctypes.c_void_p(la.data_ptr()), None, None,
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc != 0
    assert b"static scratch" in _lib.LIB.vmc_last_error()
