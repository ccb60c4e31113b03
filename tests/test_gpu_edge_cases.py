"""Edge cases of the amplitude path through the C ABI.

* Invalid calls fail with the library's message instead of running: an empty
  batch, a workspace below vmc_amp_grad_workspace_size, truncations beyond the
  kernels' limits (512 rows, 128 kept vectors; 128 x 32 variational panels).
* Rank-deficient boundaries: PEPS whose site tensors vanish unless every bond
  index is 0 make every theta matrix exactly rank deficient (zero columns,
  exactly zero Gram eigenvalues, zero Householder columns, Sturm pivots that
  hit 0). The amplitude is then the product of the selected entries and
  dlog psi / dA is their reciprocal at bond index 0 and 0 elsewhere, a closed
  form independent of any truncation gauge.
* Product states (rank-1 tensors with dense entries) against the oracle's
  full-sum contraction.
* 16384 samples at 10x10 D=4 chi=32: sampled rows across the batch match the
  oracle.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import peps as P

pytestmark = pytest.mark.gpu


def _flat(tensors):
    return np.concatenate([t.ravel() for row in tensors for t in row])


def test_invalid_calls_raise():
    from vmc_amd import _lib
    from vmc_amd.engine import AmpGradPlan, _ptr

    print("[edge] invalid calls", flush=True)
    plan = AmpGradPlan((4, 4), 2, 4, dtype=torch.float64)
    params = torch.zeros(plan.num_params, dtype=torch.float64, device="cuda")
    cfg = torch.zeros((8, 16), dtype=torch.int32, device="cuda")
    log_amp = torch.empty((8, 2), dtype=torch.float64, device="cuda")
    ws = plan.workspace(8)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    with pytest.raises(RuntimeError, match="batch must be positive"):
        _lib.check(_lib.LIB.vmc_amp_grad(plan._h, _ptr(params), _ptr(cfg), 0, _ptr(log_amp), None, None,
                                         _ptr(ws), ctypes.c_size_t(ws.numel()), stream))
    with pytest.raises(RuntimeError, match="workspace too small"):
        _lib.check(_lib.LIB.vmc_amp_grad(plan._h, _ptr(params), _ptr(cfg), 8, _ptr(log_amp), None, None,
                                         _ptr(ws), ctypes.c_size_t(64), stream))
    # truncations beyond the generic kernel (more than 512 rows: D=6 chi=100 gives
    # 600-row thetas; more than 128 kept vectors) are refused with the limits,
    # and so are variational panels beyond 512 x 128 (chi=150: 900-row panels)
    with pytest.raises(RuntimeError, match="exceeds the truncation kernels"):
        AmpGradPlan((6, 6), 6, 100, dtype=torch.float64)
    with pytest.raises(RuntimeError, match="exceeds the truncation kernels"):
        AmpGradPlan((8, 8), 4, 200, dtype=torch.float64)
    with pytest.raises(RuntimeError, match="variational"):
        AmpGradPlan((6, 6), 6, 150, dtype=torch.float64, strategy=_lib.VMC_VARIATIONAL)
    # shapes the register kernels do not take run on the generic ones
    AmpGradPlan((10, 10), 4, 40, dtype=torch.float64)
    AmpGradPlan((10, 10), 5, 32, dtype=torch.float64)
    AmpGradPlan((10, 10), 4, 40, dtype=torch.float64, strategy=_lib.VMC_VARIATIONAL)


def _corner_peps(shape, D, seed, complex_=False):
    """Site tensors that vanish unless all bond indices are 0."""
    rng = np.random.default_rng(seed)
    tensors = P.random_peps(shape, D, seed=seed)
    out = []
    for row in tensors:
        rr = []
        for t in row:
            z = np.zeros(t.shape, np.complex128 if complex_ else np.float64)
            mag = rng.uniform(0.5, 1.5, size=t.shape[0])
            z[:, 0, 0, 0, 0] = mag * (np.exp(1j * rng.uniform(0, 2 * np.pi, size=t.shape[0])) if complex_
                                      else rng.choice([-1.0, 1.0], size=t.shape[0]))
            rr.append(z)
        out.append(rr)
    return out


@pytest.mark.parametrize("dtype,tol,strategy,D,chi", [
    (torch.float64, 1e-13, "zipup", 3, 9), (torch.float32, 1e-6, "zipup", 3, 9),
    (torch.complex128, 1e-13, "zipup", 3, 9), (torch.float64, 1e-13, "density", 3, 9),
    (torch.float64, 1e-13, "variational", 3, 9),
    # chi=6 < 9: the 18x9 cuts discard through theta^T theta (k < n < m)
    (torch.float64, 1e-13, "zipup", 3, 6), (torch.float32, 1e-6, "zipup", 3, 6),
    # the generic truncation kernel: 40 kept vectors (real), 72-row complex thetas
    (torch.float64, 1e-13, "zipup", 4, 40), (torch.complex128, 1e-13, "zipup", 3, 24),
])
def test_exactly_rank_deficient_boundaries(dtype, tol, strategy, D, chi):
    from vmc_amd import _lib
    from vmc_amd.engine import AmpGradPlan

    shape, B = (6, 6), 32
    cplx = dtype.is_complex
    print(f"[edge] rank-deficient boundaries {shape} D={D} chi={chi} {dtype} {strategy}", flush=True)
    tensors = _corner_peps(shape, D, seed=4, complex_=cplx)
    cfg = P.random_configs(shape[0] * shape[1], B, seed=5)
    code = {"zipup": _lib.VMC_ZIPUP, "density": _lib.VMC_DENSITY_MATRIX, "variational": _lib.VMC_VARIATIONAL}
    plan = AmpGradPlan(shape, D, chi, dtype=dtype, strategy=code[strategy])
    params = torch.tensor(_flat(tensors), dtype=dtype, device="cuda")
    la, o, p = plan.amp_grad(params, torch.tensor(cfg), want_p=True)
    la, o, p = la.cpu().numpy(), o.cpu().numpy().astype(np.complex128 if cplx else np.float64), p.cpu().numpy()
    # closed form: psi = prod_s A_s[sigma_s, 0, 0, 0, 0]
    sel = np.stack([tensors[r][c][cfg[:, r * shape[1] + c], 0, 0, 0, 0]
                    for r in range(shape[0]) for c in range(shape[1])], axis=1)
    ref_log = np.sum(np.log(np.abs(sel)), axis=1)
    assert np.max(np.abs(la[:, 0] - ref_log)) < tol * np.max(np.abs(ref_log))
    if cplx:
        dphase = np.angle(np.exp(1j * (la[:, 1] - np.sum(np.angle(sel), axis=1))))
        assert np.max(np.abs(dphase)) < 1e-12
    else:
        assert np.array_equal(la[:, 1] != 0, np.prod(np.sign(sel), axis=1) < 0)
    # o: 1 / A_s[sigma_s, 0, 0, 0, 0] at the first entry of each site's slice, 0 elsewhere
    off = 0
    for r in range(shape[0]):
        for c in range(shape[1]):
            n = tensors[r][c][0].size
            blk = o[:, off:off + n]
            assert np.allclose(blk[:, 0], 1.0 / sel[:, r * shape[1] + c], rtol=tol * 10, atol=0)
            assert np.max(np.abs(blk[:, 1:])) <= tol * np.max(np.abs(blk[:, 0]))
            assert np.all(p[:, off:off + n] == cfg[:, r * shape[1] + c][:, None])
            off += n


def test_product_state_matches_full_sum():
    from vmc_amd import _lib
    from vmc_amd.engine import AmpGradPlan

    shape, D, B = (4, 4), 3, 32
    print(f"[edge] product state {shape} D={D}", flush=True)
    rng = np.random.default_rng(7)
    tensors = []
    for r in range(shape[0]):
        row = []
        for c in range(shape[1]):
            dims = P.site_dims(r, c, shape[0], shape[1], D)
            t = rng.standard_normal(2)[:, None, None, None, None]
            for k, dk in enumerate(dims):
                v = rng.standard_normal(dk).reshape([1] * (k + 1) + [dk] + [1] * (3 - k))
                t = t * v
            row.append(t)
        tensors.append(row)
    cfg = P.random_configs(16, B, seed=8)
    ref = P.value(tensors, cfg, shape, P.NoTruncation())
    for strategy, chi in ((_lib.VMC_ZIPUP, 4), (_lib.VMC_NO_TRUNCATION, 0)):
        plan = AmpGradPlan(shape, D, chi, dtype=torch.float64, strategy=strategy)
        params = torch.tensor(_flat(tensors), dtype=torch.float64, device="cuda")
        la, _, _ = plan.amp_grad(params, torch.tensor(cfg), grad=False)
        la = la.cpu().numpy()
        err = np.max(np.abs(la[:, 0] - np.log(np.abs(ref))))
        print(f"  strategy {strategy}: max |dlog psi| {err:.2e}", flush=True)
        assert err < 1e-11
        assert np.array_equal(la[:, 1] != 0, ref < 0)


def test_large_batch_offsets():
    """16384 samples at 10x10 D=4 chi=32 (o is 2.4 GB in FP64): the first and
    last samples match the oracle, so no per-sample offset wraps."""
    from vmc_amd.engine import AmpGradPlan

    shape, D, chi, B = (10, 10), 4, 32, 16384
    print(f"[edge] large batch {shape} D={D} chi={chi} B={B}", flush=True)
    tensors = P.random_peps(shape, D, seed=21)
    cfg = P.random_configs(100, B, seed=22)
    plan = AmpGradPlan(shape, D, chi, dtype=torch.float64)
    params = torch.tensor(_flat(tensors), dtype=torch.float64, device="cuda")
    la, o, _ = plan.amp_grad(params, torch.tensor(cfg))
    pick = np.array([0, 1, B // 2, B - 2, B - 1])
    la = la[pick].cpu().numpy()
    o = o[pick].cpu().numpy()
    amp, grad, _ = P.value_and_grad(tensors, cfg[pick], shape, P.ZipUp(chi), D)
    err_log = np.max(np.abs(la[:, 0] - np.log(np.abs(amp))) / np.abs(np.log(np.abs(amp))))
    err_o = np.max(np.abs(o - grad / amp[:, None])) / np.max(np.abs(grad / amp[:, None]))
    print(f"  rel err log psi {err_log:.2e} o {err_o:.2e}", flush=True)
    assert err_log < 1e-10 and err_o < 1e-7
